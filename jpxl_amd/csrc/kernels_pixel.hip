// gfx950 kernels for the VarDCT pixel path: sRGB -> XYB, forward DCT over the
// block map, quantization (LF, HF nearest rule, trailing truncation, CfL),
// reconstruction (dequant, CfL, LLF, inverse DCT), Gaborish and the EPF.
//
// Layouts in HBM:
//  * XYB / linear planes: f32, W x H, row-major, stride W.
//  * Coefficient planes (f32 and i32): the "footprint" layout.  A varblock of
//    side s at block (bx, by) stores cell k (row k/s, column k%s of its
//    coefficient matrix) at plane[(8*by + k/s) * 8*bw + 8*bx + k%s].  Every
//    coefficient of the frame therefore has exactly one slot and the planes
//    are 8*bw x 8*bh, so each stage reads and writes dense, coalesced rows.
//  * LF planes: i32, bw x bh per channel.
//  * Varblock work lists: one list per transform size, entries (by << 16) | bx.
//
// One workgroup of 192 threads (three wave64s) processes 64/N varblocks of
// size N, so every 1-D transform pass keeps all 192 lanes busy (3 channels x
// N lines x 64/N varblocks).
#include "dct_matrix.h"
#include "internal.h"

namespace jpxlg {

namespace {

constexpr int kThreads = 192;

// ----------------------------------------------------------------------------
// sRGB -> linear -> XYB  (jpxl-encode-policy/src/source.rs:226-331, :355;
// jpxl-core/src/color.rs:84-97)
// ----------------------------------------------------------------------------

template <bool kEight>
__global__ __launch_bounds__(256) void k_xyb_fwd(int n, const void* __restrict__ rgb,
                                                 const float* __restrict__ lut,
                                                 float* __restrict__ X, float* __restrict__ Y,
                                                 float* __restrict__ B, float* __restrict__ R,
                                                 float* __restrict__ G, float* __restrict__ Bl,
                                                 int* __restrict__ gray) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool g = true;
  if (i < n) {
    uint32_t cr, cg, cb;
    if constexpr (kEight) {
      const uint8_t* p = (const uint8_t*)rgb + 3 * (size_t)i;
      cr = p[0]; cg = p[1]; cb = p[2];
    } else {
      const uint16_t* p = (const uint16_t*)rgb + 3 * (size_t)i;
      cr = p[0]; cg = p[1]; cb = p[2];
    }
    g = cr == cg && cg == cb;
    float r = lut[cr], gg = lut[cg], b = lut[cb];
    R[i] = r; G[i] = gg; Bl[i] = b;
    const float bias = 0.0037930734f, bias_cbrt = 0.1559542f;
    float lm = 0.3f * r + 0.622f * gg + 0.078f * b + bias;
    float mm = 0.23f * r + 0.692f * gg + 0.078f * b + bias;
    float sm = 0.2434227f * r + 0.2047674f * gg + 0.5518099f * b + bias;
    float lg = cbrt_core(lm) - bias_cbrt;
    float mg = cbrt_core(mm) - bias_cbrt;
    float sg = cbrt_core(sm) - bias_cbrt;
    X[i] = 0.5f * (lg - mg);
    Y[i] = 0.5f * (lg + mg);
    B[i] = sg;
  }
  // One atomic per wave at most, and none once another wave has cleared it.
  if (__any(!g) && (threadIdx.x & 63) == 0 && __atomic_load_n(gray, __ATOMIC_RELAXED) != 0)
    atomicAnd(gray, 0);
}

// ----------------------------------------------------------------------------
// Square 2-D transforms (jpxl-core/src/dct.rs:956-997) on an LDS tile of
// N rows x (N+1) padded columns.  Each call is one thread's 1-D line.
// ----------------------------------------------------------------------------

// DCT_2D pass 1: column `col` of the samples, transformed in place.
template <int N>
__device__ __forceinline__ void fwd_pass1(float* tile, int col, const DctTables& t) {
  constexpr int P = N + 1;
  float v[N];
#pragma unroll
  for (int r = 0; r < N; r++) v[r] = tile[r * P + col];
  dct_ii<N>(v, t);
  const float f = fwd_factor<N>();
#pragma unroll
  for (int r = 0; r < N; r++) tile[r * P + col] = v[r] * f;
}

// DCT_2D pass 2: row `row` of the pass-1 result; out[k] is coefficient cell
// k*N + row (the transposed orientation dct_2d_in_place keeps for squares).
template <int N>
__device__ __forceinline__ void fwd_pass2(const float* tile, int row, float* out,
                                          const DctTables& t) {
  constexpr int P = N + 1;
#pragma unroll
  for (int k = 0; k < N; k++) out[k] = tile[row * P + k];
  dct_ii<N>(out, t);
  const float f = fwd_factor<N>();
#pragma unroll
  for (int k = 0; k < N; k++) out[k] = out[k] * f;
}

// IDCT_2D pass 1: coefficient column r (cells k*N + r), inverse along k,
// written back in place (tile[c][r] then holds the row-r, column-c partial).
template <int N>
__device__ __forceinline__ void inv_pass1(float* tile, int r, const DctTables& t) {
  constexpr int P = N + 1;
  float v[N];
#pragma unroll
  for (int k = 0; k < N; k++) v[k] = tile[k * P + r];
  dct_iii<N>(v, t);
  const float f = inv_factor<N>();
#pragma unroll
  for (int k = 0; k < N; k++) tile[k * P + r] = v[k] * f;
}

// IDCT_2D pass 2: sample column c; out[r] is sample (row r, column c).
template <int N>
__device__ __forceinline__ void inv_pass2(const float* tile, int c, float* out,
                                          const DctTables& t) {
  constexpr int P = N + 1;
#pragma unroll
  for (int r = 0; r < N; r++) out[r] = tile[c * P + r];
  dct_iii<N>(out, t);
  const float f = inv_factor<N>();
#pragma unroll
  for (int r = 0; r < N; r++) out[r] = out[r] * f;
}

// Small n x n DCT_2D / IDCT_2D held in registers (the LF <-> LLF maps of
// jpxl-core/src/varblock.rs:817-855 and forward.rs:779-830), same operation
// order as dct_2d_in_place / idct_2d_in_place.
template <int n>
__device__ __forceinline__ void small_dct2_fwd(float* w, const DctTables& t) {
  if constexpr (n == 1) return;
  const float f = fwd_factor<n>();
  float v[n], s[n * n];
#pragma unroll
  for (int c = 0; c < n; c++) {
#pragma unroll
    for (int r = 0; r < n; r++) v[r] = w[r * n + c];
    dct_ii<n>(v, t);
#pragma unroll
    for (int r = 0; r < n; r++) w[r * n + c] = v[r] * f;
  }
#pragma unroll
  for (int r = 0; r < n; r++)
#pragma unroll
    for (int c = 0; c < n; c++) s[c * n + r] = w[r * n + c];
#pragma unroll
  for (int r = 0; r < n; r++) {
#pragma unroll
    for (int k = 0; k < n; k++) v[k] = s[k * n + r];
    dct_ii<n>(v, t);
#pragma unroll
    for (int k = 0; k < n; k++) s[k * n + r] = v[k] * f;
  }
#pragma unroll
  for (int i = 0; i < n * n; i++) w[i] = s[i];
}

template <int n>
__device__ __forceinline__ void small_dct2_inv(float* w, const DctTables& t) {
  if constexpr (n == 1) return;
  const float f = inv_factor<n>();
  float v[n], s[n * n];
#pragma unroll
  for (int i = 0; i < n * n; i++) s[i] = w[i];
#pragma unroll
  for (int r = 0; r < n; r++) {
#pragma unroll
    for (int k = 0; k < n; k++) v[k] = s[k * n + r];
    dct_iii<n>(v, t);
#pragma unroll
    for (int k = 0; k < n; k++) s[k * n + r] = v[k] * f;
  }
#pragma unroll
  for (int r = 0; r < n; r++)
#pragma unroll
    for (int c = 0; c < n; c++) w[r * n + c] = s[c * n + r];
#pragma unroll
  for (int c = 0; c < n; c++) {
#pragma unroll
    for (int r = 0; r < n; r++) v[r] = w[r * n + c];
    dct_iii<n>(v, t);
#pragma unroll
    for (int r = 0; r < n; r++) w[r * n + c] = v[r] * f;
  }
}

__device__ __forceinline__ void unpack(uint32_t o, int* bx, int* by) {
  *bx = (int)(o & 0xffffu);
  *by = (int)(o >> 16);
}

// ----------------------------------------------------------------------------
// Forward DCT over the block map (jpxl-encode-policy/src/lib.rs:3009-3098).
// ----------------------------------------------------------------------------

template <int N, int FORM>
__global__ __launch_bounds__(kThreads) void k_dct_fwd(Geom g, const uint32_t* __restrict__ list,
                                                      int count_host, const StructInfo* __restrict__ si,
                                                      const ConstTables* __restrict__ tab,
                                                      const DctMats* __restrict__ mats,
                                                      const float* __restrict__ X,
                                                      const float* __restrict__ Y,
                                                      const float* __restrict__ B,
                                                      float* __restrict__ CX,
                                                      float* __restrict__ CY,
                                                      float* __restrict__ CB) {
  constexpr int VB = 64 / N, P = N + 1, NN = N * N;
  __shared__ float lds[VB * 3 * N * P];
  __shared__ int vbx[VB], vby[VB];
  // the count from the device structure when given (the grid is then an
  // upper bound and workgroups past the count exit)
  const int count = si ? si->count[size_index(N)] : count_host;
  const int base = blockIdx.x * VB;
  if (base >= count) return;
  if (threadIdx.x < VB) {
    int bx = 0, by = 0;
    if (base + (int)threadIdx.x < count) unpack(list[base + threadIdx.x], &bx, &by);
    vbx[threadIdx.x] = bx;
    vby[threadIdx.x] = by;
  }
  __syncthreads();
  const float* planes[3] = {X, Y, B};
  for (int e = threadIdx.x; e < VB * 3 * NN; e += kThreads) {
    int j = e / (3 * NN), rem = e % (3 * NN), c = rem / NN, k = rem % NN;
    int dy = k / N, dx = k % N;
    float v = 0.0f;
    if (base + j < count) {
      int x = min(8 * vbx[j] + dx, g.w - 1), y = min(8 * vby[j] + dy, g.h - 1);
      v = planes[c][(size_t)y * g.w + x];
    }
    lds[((j * 3 + c) * N + dy) * P + dx] = v;
  }
  __syncthreads();
  if constexpr (FORM != kDctExact) {
    const float* L = N == 8 ? mats->fwd8 : (N == 16 ? mats->fwd16 : mats->fwd32);
    float* dsts[3] = {CX, CY, CB};
    matrix_transform_tiles<N, FORM>(lds, VB * 3, L, [&](int t, int k, int r, float v) {
      const int j = t / 3;
      if (base + j < count) stg32(dsts[t % 3], (uint32_t)((8 * vby[j] + k) * g.cstride + 8 * vbx[j] + r), v);
    });
    return;
  }
  const DctTables& t = tab->dct;
  const int j = threadIdx.x / (3 * N), c = (threadIdx.x / N) % 3, line = threadIdx.x % N;
  float* tile = lds + (j * 3 + c) * N * P;
  fwd_pass1<N>(tile, line, t);
  __syncthreads();
  float out[N];
  fwd_pass2<N>(tile, line, out, t);
  if (base + j < count) {
    float* dst = c == 0 ? CX : (c == 1 ? CY : CB);
#pragma unroll
    for (int k = 0; k < N; k++) dst[(size_t)(8 * vby[j] + k) * g.cstride + 8 * vbx[j] + line] = out[k];
  }
}

// The per-block first-guess features of the quality search, one pass over
// the XYB planes: each aligned 8x8 block's sample variance per channel (the
// analysis atlas's atoms, over the samples the frame has, in raster order:
// quality_features.rs:173-213, analysis.rs:97-157) and its DCT8 summary terms
// (quality_transform_summary, jpxl-encode-policy/src/lib.rs:2769-2874: the
// exact DCT8 of the block, edge blocks replicating the last column and row
// as the cover's candidates do; per block, in cell order, the Y AC energy,
// low/high/row/column band energies, near-zero counts, DC and X+B AC
// energy).  A workgroup stages kSummaryBlocks consecutive raster blocks'
// samples in LDS with row-contiguous loads, reduces each block on a lane of
// its own, transforms eight blocks at a time in place, and sums its blocks in
// block order; the histogram of ln(Y AC energy) bins is an integer count,
// kept in LDS per workgroup.
constexpr int kSummaryBlocks = 64;

__global__ __launch_bounds__(kThreads) void k_dct8_summary(Geom g, const ConstTables* __restrict__ tab,
                                                           const float* __restrict__ X,
                                                           const float* __restrict__ Y,
                                                           const float* __restrict__ B,
                                                           float* __restrict__ luma_var,
                                                           float* __restrict__ chroma_var,
                                                           SummaryPartial* __restrict__ parts,
                                                           uint32_t* __restrict__ hist) {
  constexpr int N = 8, VB = 8, P = N + 1, NN = 64, CP = kSummaryBlocks + 1;
  constexpr int kStrip = kSummaryBlocks * N;  // samples per staged row
  __shared__ float tile[VB * 3 * N * P];
  __shared__ float smp[3][NN][CP];  // channel, cell, block: samples, then coefficients
  __shared__ double bsum[9][kSummaryBlocks];  // y, low, high, row, col, X, dc, dc^2, B
  __shared__ uint32_t bcnt[2][kSummaryBlocks];
  __shared__ uint32_t lhist[kSummaryBins];
  __shared__ float vxb[2][kSummaryBlocks];
  for (int i = threadIdx.x; i < kSummaryBins; i += kThreads) lhist[i] = 0u;
  const float* planes[3] = {X, Y, B};
  const int nblocks = g.bw * g.bh;
  const int base = (int)blockIdx.x * kSummaryBlocks;
  const int nb = min(kSummaryBlocks, nblocks - base);
  // stage: consecutive lanes read consecutive samples of a row strip; each
  // lane requests a batch of samples before storing any
  static_assert((3 * N * kStrip) % (kThreads * 16) == 0, "whole staging batches");
  for (int e0 = 0; e0 < 3 * N * kStrip; e0 += kThreads * 16) {
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int e = e0 + u * kThreads + (int)threadIdx.x;
      const int c = e / (N * kStrip), rem = e % (N * kStrip), dy = rem / kStrip, xx = rem % kStrip;
      const int jb = xx / N, dx = xx % N;
      const int blk = min(base + jb, nblocks - 1);
      const int bx = blk % g.bw, by = blk / g.bw;
      const int x = min(8 * bx + dx, g.w - 1), y = min(8 * by + dy, g.h - 1);
      v[u] = planes[c][(size_t)y * g.w + x];
    }
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int e = e0 + u * kThreads + (int)threadIdx.x;
      const int c = e / (N * kStrip), rem = e % (N * kStrip), dy = rem / kStrip, xx = rem % kStrip;
      smp[c][dy * N + xx % N][xx / N] = v[u];
    }
  }
  __syncthreads();
  // atom variances: a lane per (channel, block)
  {
    const int c = threadIdx.x >> 6, jb = threadIdx.x & 63;
    if (jb < nb) {
      const int bx = (base + jb) % g.bw, by = (base + jb) / g.bw;
      const int vw = min(N, g.w - 8 * bx), vh = min(N, g.h - 8 * by);
      double sum = 0.0, sum_sq = 0.0;
      for (int dy = 0; dy < vh; dy++)
        for (int dx = 0; dx < vw; dx++) {
          const double v = (double)smp[c][dy * N + dx][jb];
          sum += v;
          sum_sq += v * v;
        }
      const double count = (double)vw * (double)vh;
      const double mean = sum / count;
      double variance = sum_sq / count - mean * mean;
      if (!(variance > 0.0)) variance = 0.0;
      if (c == 1) luma_var[base + jb] = (float)variance;
      else vxb[c >> 1][jb] = (float)variance;
    }
  }
  __syncthreads();
  if (threadIdx.x < nb) chroma_var[base + threadIdx.x] = vxb[0][threadIdx.x] + vxb[1][threadIdx.x];
  const DctTables& t = tab->dct;
  for (int sub = 0; sub * VB < nb; sub++) {
    for (int e = threadIdx.x; e < VB * 3 * NN; e += kThreads) {
      const int j = e / (3 * NN), rem = e % (3 * NN), c = rem / NN, k = rem % NN;
      tile[((j * 3 + c) * N + k / N) * P + k % N] = smp[c][k][sub * VB + j];
    }
    __syncthreads();
    const int j = threadIdx.x / (3 * N), c = (threadIdx.x / N) % 3, line = threadIdx.x % N;
    float* tl = tile + (j * 3 + c) * N * P;
    fwd_pass1<N>(tl, line, t);
    __syncthreads();
    float out[N];
    fwd_pass2<N>(tl, line, out, t);
#pragma unroll
    for (int k = 0; k < N; k++) smp[c][k * N + line][sub * VB + j] = out[k];
    __syncthreads();
  }
  // Per block, three waves in parallel (a lane per block): wave 0 the Y AC
  // energy in cell order, the near-zero counts and the histogram bin; wave 1
  // the low/high bands and X; wave 2 the row/column sides and B.
  {
    const int wv = threadIdx.x >> 6, jb = threadIdx.x & 63;
    if (jb < nb) {
      if (wv == 0) {
        double ey = 0.0;
        uint32_t a3 = 0, a2 = 0;
        for (int cell = 1; cell < NN; cell++) {
          const double v = (double)smp[1][cell][jb];
          ey += v * v;
          const double m = fabs(v);
          a3 += m < 1e-3 ? 1u : 0u;
          a2 += m < 1e-2 ? 1u : 0u;
        }
        bsum[0][jb] = ey;
        const double dc = (double)smp[1][0][jb];
        bsum[6][jb] = dc;
        bsum[7][jb] = dc * dc;
        bcnt[0][jb] = a3;
        bcnt[1][jb] = a2;
        const double bin = fmin(fmax((log(ey + 1e-30) - (-46.0)) / 0.125, 0.0), (double)(kSummaryBins - 1));
        atomicAdd(&lhist[(int)bin], 1u);
      } else if (wv == 1) {
        double lo = 0.0, hi = 0.0, xx = 0.0;
        for (int cell = 1; cell < NN; cell++) {
          const int row = cell >> 3, col = cell & 7;
          const double v = (double)smp[1][cell][jb];
          const double en = v * v;
          if (row + col <= 2) lo += en;
          if (max(row, col) >= 4) hi += en;
          const double x = (double)smp[0][cell][jb];
          xx += x * x;
        }
        bsum[1][jb] = lo;
        bsum[2][jb] = hi;
        bsum[5][jb] = xx;
      } else {
        double rw = 0.0, cl = 0.0, bb = 0.0;
        for (int cell = 1; cell < NN; cell++) {
          const int row = cell >> 3, col = cell & 7;
          const double v = (double)smp[1][cell][jb];
          const double en = v * v;
          if (row > col) rw += en;
          else if (col > row) cl += en;
          const double b = (double)smp[2][cell][jb];
          bb += b * b;
        }
        bsum[3][jb] = rw;
        bsum[4][jb] = cl;
        bsum[8][jb] = bb;
      }
    }
  }
  __syncthreads();
  // the workgroup's sums in block order, one quantity per thread
  const int q = threadIdx.x;
  if (q < 8) {
    double acc = 0.0;
    for (int jb = 0; jb < nb; jb++) acc += q == 5 ? bsum[5][jb] + bsum[8][jb] : bsum[q][jb];
    double* o = &parts[blockIdx.x].total_y;
    o[q] = acc;
  } else if (q == 8) {
    uint32_t n3 = 0, n2 = 0;
    for (int jb = 0; jb < nb; jb++) {
      n3 += bcnt[0][jb];
      n2 += bcnt[1][jb];
    }
    SummaryPartial& o = parts[blockIdx.x];
    o.near_1e3 = n3;
    o.near_1e2 = n2;
    o.blocks = (uint32_t)nb;
    o.ac_cells = (uint32_t)nb * 63u;
  }
  for (int i = threadIdx.x; i < kSummaryBins; i += kThreads)
    if (lhist[i]) atomicAdd(&hist[i], lhist[i]);
}

__global__ void k_fill_u32(uint32_t* p, int n, uint32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Forward or inverse DCT of `count` contiguous N x N blocks, through the same
// device passes the pipeline uses (the C-ABI's transform entry point).
template <int N>
__global__ __launch_bounds__(kThreads) void k_dct_blocks(int inverse, int count,
                                                         const ConstTables* __restrict__ tab,
                                                         const float* __restrict__ in,
                                                         float* __restrict__ out) {
  constexpr int VB = 64 / N, P = N + 1, NN = N * N;
  __shared__ float lds[VB * 3 * N * P];
  const int base = blockIdx.x * VB * 3;  // 3 consecutive blocks per "varblock"
  for (int e = threadIdx.x; e < VB * 3 * NN; e += kThreads) {
    int blk = e / NN, k = e % NN;
    lds[(blk * N + k / N) * P + k % N] = base + blk < count ? in[(size_t)(base + blk) * NN + k] : 0.0f;
  }
  __syncthreads();
  const DctTables& t = tab->dct;
  const int blk = threadIdx.x / N, line = threadIdx.x % N;
  float* tile = lds + blk * N * P;
  float v[N];
  if (!inverse) {
    fwd_pass1<N>(tile, line, t);
    __syncthreads();
    fwd_pass2<N>(tile, line, v, t);
    if (base + blk < count)
#pragma unroll
      for (int k = 0; k < N; k++) out[(size_t)(base + blk) * NN + k * N + line] = v[k];
  } else {
    inv_pass1<N>(tile, line, t);
    __syncthreads();
    inv_pass2<N>(tile, line, v, t);
    if (base + blk < count)
#pragma unroll
      for (int r = 0; r < N; r++) out[(size_t)(base + blk) * NN + r * N + line] = v[r];
  }
}

// ----------------------------------------------------------------------------
// Quantization (jpxl-encode-policy/src/lib.rs:3237-3384 quantize_square_varblock;
// quantize.rs:150-213 LfQuantizer, :1400-1499 choose, :468-545 truncate_trailing)
// ----------------------------------------------------------------------------

__device__ __constant__ float kQuantBias[3] = {1.0f - 0.054650073f, 1.0f - 0.0700545f,
                                               1.0f - 0.049935103f};

// I.5.3 nearest-reconstruction choice, ties to the smaller magnitude.
__device__ __forceinline__ int32_t hf_choose(float target, float step, float bias, int* bad) {
  float thr = 0.5f * bias * step;
  if (fabsf(target) <= thr) return 0;
  int32_t e = clamp_round(target / step, bad);
  int32_t cand[4] = {0, e - 1, e, e + 1};
  int32_t best = 0;
  float best_err = INFINITY;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int32_t v = cand[i];
    if (abs(v) > kMaxQuant) continue;
    float err = fabsf(bias_adjust(v, bias) * step - target);
    if (err < best_err || (err == best_err && abs(v) < abs(best))) {
      best = v;
      best_err = err;
    }
  }
  return best;
}

// Trailing truncation of one varblock channel by one wave64
// (jpxl-encode-policy/src/quantize.rs:468-545).  The reference walks the
// natural order backwards from the last nonzero and zeroes it while the bits
// it saves (its own magnitude class plus the zero tokens it exposes back to
// the previous nonzero) exceed its weighted distortion gain; it stops at the
// first coefficient that fails.  The walk only ever removes a suffix, so each
// nonzero's "previous nonzero" is that of the unmodified block and every
// coefficient's test is independent: the lanes evaluate all of them at once
// and zero every nonzero after the last one that fails.
template <int N>
__device__ void truncate_wave(int32_t* q, const float* targets, float* recon_y,
                              const uint16_t* order, const float* mat, const float* donor,
                              float sc, float bias, double lambda, float ztb) {
  constexpr int NN = N * N, nb = (N / 8) * (N / 8), R = NN / 64;
  const int lane = threadIdx.x & 63;
  const int pos0 = lane * R;
  int cell[R];
  int32_t v[R];
  int lastnz = -1;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int p = pos0 + r;
    cell[r] = order[p];
    v[r] = p >= nb ? q[cell[r]] : 0;
    if (v[r] != 0) lastnz = p;
  }
  // Inclusive max-scan of the lanes' last nonzero positions.
  int incl = lastnz;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(incl, d, 64);
    if (lane >= d) incl = max(incl, o);
  }
  int prev = __shfl_up(incl, 1, 64);
  if (lane == 0) prev = -1;
  int fail = -1;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int p = pos0 + r;
    if (v[r] == 0) continue;
    const int c = cell[r];
    const float weight = (float)(lambda * (double)NN * (double)donor[c]);
    const int exposed = prev >= 0 ? p - prev - 1 : p - nb;
    const float own = (float)(33 - __clz((unsigned)abs(v[r])));
    const float saved = own + (float)exposed * ztb;
    const float target = targets[c];
    const float step = sc * mat[c];
    const float recon = bias_adjust(v[r], bias) * step;
    const float kept = recon - target;
    const float added = fmaf(target, target, -(kept * kept));
    if (saved <= weight * added) fail = p;
    prev = p;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) fail = max(fail, __shfl_xor(fail, d, 64));
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int p = pos0 + r;
    if (v[r] != 0 && p > fail) {
      const int c = cell[r];
      q[c] = 0;
      if (recon_y) recon_y[c] = bias_adjust(0, bias) * (sc * mat[c]);
    }
  }
}

// Varblocks per quantize workgroup and its width: DCT32 takes one varblock
// per 256-thread workgroup so its LDS footprint (~29 KB) leaves room for
// five workgroups per CU; the smaller sizes keep 64/N varblocks on 192.
constexpr int quant_vb(int n) { return n == 32 ? 1 : 64 / n; }
constexpr int quant_threads(int) { return 256; }

// The same truncation spread over a group of `nw` waves (DCT32, where one
// varblock per workgroup would otherwise leave waves idle): lane L of the
// group holds positions [L*R, L*R + R); the scans and the reduction cross the
// waves through `xs` (LDS, 2 * nw ints per group).  Every wave of the
// workgroup must call it (it synchronises the workgroup).
template <int N>
__device__ void truncate_group(int32_t* q, const float* targets, float* recon_y, const uint16_t* order,
                               const float* mat, const float* donor, float sc, float bias, double lambda,
                               float ztb, int nw, int wi, int* xs) {
  constexpr int NN = N * N, nb = (N / 8) * (N / 8);
  constexpr int RMAX = NN / 64;
  const int R = NN / (64 * nw);
  const int lane = threadIdx.x & 63;
  const int pos0 = (wi * 64 + lane) * R;
  int cell[RMAX];
  int32_t v[RMAX];
  int lastnz = -1;
#pragma unroll
  for (int r = 0; r < RMAX; r++) {
    if (r >= R) break;
    const int p = pos0 + r;
    cell[r] = order[p];
    v[r] = p >= nb ? q[cell[r]] : 0;
    if (v[r] != 0) lastnz = p;
  }
  int incl = lastnz;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(incl, d, 64);
    if (lane >= d) incl = max(incl, o);
  }
  if (lane == 63) xs[wi] = incl;
  __syncthreads();
  int before = -1;  // last nonzero in the group's earlier waves
  for (int k = 0; k < wi; k++) before = max(before, xs[k]);
  int prev = __shfl_up(incl, 1, 64);
  if (lane == 0) prev = -1;
  prev = max(prev, before);
  int fail = -1;
#pragma unroll
  for (int r = 0; r < RMAX; r++) {
    if (r >= R) break;
    const int p = pos0 + r;
    if (v[r] == 0) continue;
    const int c = cell[r];
    const float weight = (float)(lambda * (double)NN * (double)donor[c]);
    const int exposed = prev >= 0 ? p - prev - 1 : p - nb;
    const float own = (float)(33 - __clz((unsigned)abs(v[r])));
    const float saved = own + (float)exposed * ztb;
    const float target = targets[c];
    const float step = sc * mat[c];
    const float recon = bias_adjust(v[r], bias) * step;
    const float kept = recon - target;
    const float added = fmaf(target, target, -(kept * kept));
    if (saved <= weight * added) fail = p;
    prev = p;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) fail = max(fail, __shfl_xor(fail, d, 64));
  if (lane == 0) xs[nw + wi] = fail;
  __syncthreads();
  for (int k = 0; k < nw; k++) fail = max(fail, xs[nw + k]);
#pragma unroll
  for (int r = 0; r < RMAX; r++) {
    if (r >= R) break;
    const int p = pos0 + r;
    if (v[r] != 0 && p > fail) {
      const int c = cell[r];
      q[c] = 0;
      if (recon_y) recon_y[c] = bias_adjust(0, bias) * (sc * mat[c]);
    }
  }
}

// truncate_group with the lane's tables loaded beforehand: cell[r], don[r]
// and mpos[r] are the natural order, donor weight and matrix entry of the
// lane's positions [L*R, L*R + R) (L = wi*64 + lane), so the walk waits on
// no global load.  Same decisions as truncate_group.
template <int N, int R>
__device__ void truncate_pre(int32_t* q, const float* targets, float* recon_y, const int* cell, const float* don,
                             const float* mpos, float sc, float bias, double lambda, float ztb, int nw, int wi,
                             int* xs) {
  constexpr int NN = N * N, nb = (N / 8) * (N / 8);
  const int lane = threadIdx.x & 63;
  const int pos0 = (wi * 64 + lane) * R;
  int32_t v[R];
  int lastnz = -1;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int p = pos0 + r;
    v[r] = p >= nb ? q[cell[r]] : 0;
    if (v[r] != 0) lastnz = p;
  }
  int incl = lastnz;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(incl, d, 64);
    if (lane >= d) incl = max(incl, o);
  }
  if (lane == 63) xs[wi] = incl;
  __syncthreads();
  int before = -1;  // last nonzero in the group's earlier waves
  for (int k = 0; k < wi; k++) before = max(before, xs[k]);
  int prev = __shfl_up(incl, 1, 64);
  if (lane == 0) prev = -1;
  prev = max(prev, before);
  int fail = -1;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int p = pos0 + r;
    if (v[r] == 0) continue;
    const float weight = (float)(lambda * (double)NN * (double)don[r]);
    const int exposed = prev >= 0 ? p - prev - 1 : p - nb;
    const float own = (float)(33 - __clz((unsigned)abs(v[r])));
    const float saved = own + (float)exposed * ztb;
    const float target = targets[cell[r]];
    const float step = sc * mpos[r];
    const float recon = bias_adjust(v[r], bias) * step;
    const float kept = recon - target;
    const float added = fmaf(target, target, -(kept * kept));
    if (saved <= weight * added) fail = p;
    prev = p;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) fail = max(fail, __shfl_xor(fail, d, 64));
  if (lane == 0) xs[nw + wi] = fail;
  __syncthreads();
  for (int k = 0; k < nw; k++) fail = max(fail, xs[nw + k]);
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int p = pos0 + r;
    if (v[r] != 0 && p > fail) {
      q[cell[r]] = 0;
      if (recon_y) recon_y[cell[r]] = bias_adjust(0, bias) * (sc * mpos[r]);
    }
  }
}

// LDS of one quantize workgroup of side N.
template <int N>
struct QuantSmem {
  static constexpr int VB = quant_vb(N), NN = N * N, n2 = (N / 8) * (N / 8);
  union {
    float co[3][VB][NN];
    float tile[VB * 3 * N * (N + 1)];  // the fused reconstruction's IDCT tiles, after the walks
  };
  int32_t qs[3][VB][NN];
  float dylf[VB][n2];
  float lfs[3][VB][n2], lfo[3][VB][n2];
  int32_t lfq[3][VB][n2];  // the LF levels, for the fused reconstruction
  float khf[2][VB];
  float vsc[VB][4];  // each varblock's step scales (its own HfMul under AQ) and EPF sigma
  int vbx[VB], vby[VB];
  int xs[2][8];
};
union QuantSmemAll {
  QuantSmem<8> s8;
  QuantSmem<16> s16;
  QuantSmem<32> s32;
};

// One quantize workgroup (logical index `block` of the size's prefix `off`).
// The reconstruction of a quantize workgroup's varblocks from the levels it
// has just chosen (still in LDS): the stages of render_body -- I.5.3
// dequantization with the HF CfL, the LLF cells from the LF levels, the
// I.9 inverse DCT and the EPF sigma -- with the same operations, so the
// planes equal the render kernel's bit for bit.  T threads, VB varblocks.
template <int N, int T>
__device__ __forceinline__ void render_from_levels(const Geom& g, const ConstTables* __restrict__ tab,
                                                   const ProbeParams* __restrict__ pp,
                                                   const CflParams* __restrict__ cflp, QuantSmem<N>& sm, int nvb,
                                                   float* __restrict__ X, float* __restrict__ Y,
                                                   float* __restrict__ B, float* __restrict__ sigma) {
  constexpr int VB = quant_vb(N), P = N + 1, NN = N * N, n = N / 8, si = size_index(N), cb = kCellBase[si];
  float* tile = sm.tile;
  for (int e = threadIdx.x; e < VB * NN; e += T) {
    const int j = e / NN, k = e % NN;
    float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f;
    if (j < nvb) {
      d0 = bias_adjust(sm.qs[0][j][k], kQuantBias[0]) * sm.vsc[j][0] * tab->matrix[0][cb + k];
      d1 = bias_adjust(sm.qs[1][j][k], kQuantBias[1]) * sm.vsc[j][1] * tab->matrix[1][cb + k];
      d2 = bias_adjust(sm.qs[2][j][k], kQuantBias[2]) * sm.vsc[j][2] * tab->matrix[2][cb + k];
    }
    const int r = k / N, col = k % N;
    tile[((j * 3 + 0) * N + r) * P + col] = d0 + sm.khf[0][j] * d1;
    tile[((j * 3 + 1) * N + r) * P + col] = d1;
    tile[((j * 3 + 2) * N + r) * P + col] = d2 + sm.khf[1][j] * d1;
  }
  __syncthreads();
  const DctTables& t = tab->dct;
  {
    const float kxl = 0.0f + (float)cflp->x_factor_lf / 84.0f;
    const float kbl = 1.0f + (float)cflp->b_factor_lf / 84.0f;
    const float* sf = tab->scale_f[si];
    if constexpr (n == 1) {
      if (threadIdx.x < 3 * VB) {
        const int j = threadIdx.x / 3, c = threadIdx.x % 3;
        if (j < nvb) {
          const float vy = pp->mdc[1] * (float)sm.lfq[1][j][0] / 1.0f;
          float v;
          if (c == 1) v = vy;
          else if (c == 0) v = pp->mdc[0] * (float)sm.lfq[0][j][0] / 1.0f + kxl * vy;
          else v = pp->mdc[2] * (float)sm.lfq[2][j][0] / 1.0f + kbl * vy;
          tile[(j * 3 + c) * N * P] = v;
        }
      }
    } else {
      const float f = fwd_factor<n>();
      const int j = threadIdx.x / (3 * n), c = (threadIdx.x / n) % 3, line = threadIdx.x % n;
      const bool act = threadIdx.x < 3 * VB * n && j < nvb;
      float (*llf)[n * n] = sm.lfs[0];  // the LF passes' scratch is free again
      if (act) {
        float v[n];
#pragma unroll
        for (int r = 0; r < n; r++) {
          const int i = r * n + line;
          const float vy = pp->mdc[1] * (float)sm.lfq[1][j][i] / 1.0f;
          if (c == 1) v[r] = vy;
          else if (c == 0) v[r] = pp->mdc[0] * (float)sm.lfq[0][j][i] / 1.0f + kxl * vy;
          else v[r] = pp->mdc[2] * (float)sm.lfq[2][j][i] / 1.0f + kbl * vy;
        }
        dct_ii<n>(v, t);
#pragma unroll
        for (int r = 0; r < n; r++) llf[j * 3 + c][r * n + line] = v[r] * f;
      }
      __syncthreads();
      if (act) {
        float v[n];
#pragma unroll
        for (int k = 0; k < n; k++) v[k] = llf[j * 3 + c][line * n + k];
        dct_ii<n>(v, t);
        float* tl = tile + (j * 3 + c) * N * P;
#pragma unroll
        for (int k = 0; k < n; k++) tl[k * P + line] = v[k] * f * sf[k] * sf[line];
      }
    }
  }
  __syncthreads();
  {
    const int j = threadIdx.x / (3 * N), c = (threadIdx.x / N) % 3, line = threadIdx.x % N;
    const bool act = threadIdx.x < VB * 3 * N;
    float* tl = tile + (act ? (j * 3 + c) * N * P : 0);
    if (act) inv_pass1<N>(tl, line, t);
    __syncthreads();
    if (act && j < nvb) {
      float out[N];
      inv_pass2<N>(tl, line, out, t);
      float* dst = c == 0 ? X : (c == 1 ? Y : B);
      const int fx = 8 * sm.vbx[j] + line;
      if (fx < g.w) {
#pragma unroll
        for (int r = 0; r < N; r++) {
          const int fy = 8 * sm.vby[j] + r;
          if (fy < g.h) dst[(size_t)fy * g.w + fx] = out[r];
        }
      }
    }
  }
  if (threadIdx.x < VB * n * n) {
    const int jj = threadIdx.x / (n * n), b = threadIdx.x % (n * n);
    if (jj < nvb) {
      const int sx = sm.vbx[jj] + b % n, sy = sm.vby[jj] + b / n;
      if (sx < g.bw && sy < g.bh) sigma[(size_t)sy * g.bw + sx] = sm.vsc[jj][3];
    }
  }
}

template <int N, bool RENDER>
__device__ __forceinline__ void quantize_body(const ProbeDesc* __restrict__ D, const int* __restrict__ off, int nd,
                                              int block, const ConstTables* __restrict__ tab, QuantSmem<N>& sm) {
  const int d = find_desc(off, nd, block);
  DescSpace& P = desc_at(D, d);
  const Geom g = geom_of(P.g);
  const uint32_t* __restrict__ list = P.list[size_index(N)];
  const int count = P.count[size_index(N)];
  const ProbeParams* __restrict__ pp = P.pp;
  const CflParams* __restrict__ cflp = P.cflp;
  const int8_t* __restrict__ cfl_x = P.cfl_x;
  const int8_t* __restrict__ cfl_b = P.cfl_b;
  const float* __restrict__ CX = P.coeff[0];
  const float* __restrict__ CY = P.coeff[1];
  const float* __restrict__ CB = P.coeff[2];
  int32_t* __restrict__ QX = P.quant[0];
  int32_t* __restrict__ QY = P.quant[1];
  int32_t* __restrict__ QB = P.quant[2];
  int32_t* __restrict__ LX = P.lf[0];
  int32_t* __restrict__ LY = P.lf[1];
  int32_t* __restrict__ LB = P.lf[2];
  uint32_t* __restrict__ err = P.err;
  constexpr int VB = quant_vb(N), T = quant_threads(N), NN = N * N, n = N / 8, si = size_index(N),
                cb = kCellBase[si];
  float (*co)[VB][NN] = sm.co;
  int32_t (*qs)[VB][NN] = sm.qs;
  constexpr int n2 = n * n;
  float (*dylf)[n2] = sm.dylf;
  float (*lfs)[VB][n2] = sm.lfs;
  float (*lfo)[VB][n2] = sm.lfo;
  float (*khf)[VB] = sm.khf;
  float (*vsc)[4] = sm.vsc;
  int* vbx = sm.vbx;
  int* vby = sm.vby;
  const int base = (block - off[d]) * VB;
  const int nvb = min(VB, count - base);
  int bad = 0;
  // DCT32 (one varblock per workgroup, known to every thread): its
  // coefficients and the tables this thread needs later are requested before
  // anything waits -- the matrix at its cells k = tid + 256 i and, for the
  // luma truncation walk, the order, donor and matrix entries of its 4
  // positions (the X/B walk's 8 per lane, X on waves 0-1 and B on waves 2-3,
  // are requested before the X/B choices, which keeps the register count
  // within six waves per SIMD).
  constexpr int kR = N == 32 ? 1 : 0;
  float mr[3][4 * kR + 1], cr[12 * kR + 1];
  int cellY[4 * kR + 1], cellXB[8 * kR + 1];
  float donY[4 * kR + 1], matY[4 * kR + 1], donXB[8 * kR + 1], matXB[8 * kR + 1];
  const float* cin[3] = {CX, CY, CB};
  if constexpr (N == 32) {
    const int tid = threadIdx.x;
    int bx = 0, by = 0;
    if (nvb > 0) unpack(list[base], &bx, &by);
#pragma unroll
    for (int i = 0; i < 12; i++) {
      const int k = tid + 256 * (i % 4);
      cr[i] = nvb > 0 ? ldg32(cin[i / 4], (uint32_t)((8 * by + k / N) * g.cstride + 8 * bx + k % N)) : 0.0f;
    }
    const int py = tid * 4;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int c = 0; c < 3; c++) mr[c][i] = tab->matrix[c][cb + tid + 256 * i];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      cellY[r] = tab->order[cb + py + r];
      donY[r] = tab->donor_pos[cb + py + r];
      matY[r] = tab->matrix_pos[1][cb + py + r];
    }
  }
  if (threadIdx.x < VB) {
    int bx = 0, by = 0;
    if (threadIdx.x < nvb) unpack(list[base + threadIdx.x], &bx, &by);
    vbx[threadIdx.x] = bx;
    vby[threadIdx.x] = by;
    int tile = (by / 8) * g.tw + bx / 8;
    khf[0][threadIdx.x] = 0.0f + (float)cfl_x[tile] / 84.0f;
    khf[1][threadIdx.x] = 1.0f + (float)cfl_b[tile] / 84.0f;
    varblock_scales(pp, g.bw, g.bh, bx, by, n, vsc[threadIdx.x], RENDER ? &vsc[threadIdx.x][3] : nullptr);
  }
  if constexpr (N == 32) {
#pragma unroll
    for (int i = 0; i < 12; i++) co[i / 4][0][threadIdx.x + 256 * (i % 4)] = cr[i];
  } else {
    __syncthreads();
    for (int e = threadIdx.x; e < 3 * VB * NN; e += T) {
      int c = e / (VB * NN), rem = e % (VB * NN), j = rem / NN, k = rem % NN;
      float v = 0.0f;
      if (j < nvb) v = cin[c][(size_t)(8 * vby[j] + k / N) * g.cstride + 8 * vbx[j] + k % N];
      co[c][j][k] = v;
    }
  }
  __syncthreads();
  const DctTables& t = tab->dct;
  const float* sf = tab->scale_f[si];
  const float mdc[3] = {pp->mdc[0], pp->mdc[1], pp->mdc[2]};
  // --- LF of all three channels: the n x n inverse DCT of the LLF cells
  // (forward.rs:779-830), one lane per line in each of its two passes, with
  // the operation order of small_dct2_inv ---
  if constexpr (n == 1) {
    if (threadIdx.x < 3 * VB) {
      const int c = threadIdx.x / VB, j = threadIdx.x % VB;
      lfo[c][j][0] = co[c][j][0];
    }
  } else {
    const float f = inv_factor<n>();
    if (threadIdx.x < 3 * VB * n) {
      const int c = threadIdx.x / (VB * n), j = (threadIdx.x / n) % VB, r = threadIdx.x % n;
      float v[n];
#pragma unroll
      for (int k = 0; k < n; k++) v[k] = co[c][j][k * N + r] / (sf[k] * sf[r]);
      dct_iii<n>(v, t);
#pragma unroll
      for (int k = 0; k < n; k++) lfs[c][j][k * n + r] = v[k] * f;
    }
    __syncthreads();
    if (threadIdx.x < 3 * VB * n) {
      const int c = threadIdx.x / (VB * n), j = (threadIdx.x / n) % VB, col = threadIdx.x % n;
      float v[n];
#pragma unroll
      for (int r = 0; r < n; r++) v[r] = lfs[c][j][col * n + r];
      dct_iii<n>(v, t);
#pragma unroll
      for (int r = 0; r < n; r++) lfo[c][j][r * n + col] = v[r] * f;
    }
  }
  __syncthreads();
  // --- Y LF quantize, and Y HF ---
  for (int e = threadIdx.x; e < VB * n * n; e += T) {
    const int j = e / (n * n), i = e % (n * n);
    if (j < nvb) {
      int32_t v = clamp_round(lfo[1][j][i] * 1.0f / mdc[1], &bad);
      LY[(size_t)(vby[j] + i / n) * g.bw + vbx[j] + i % n] = v;
      if constexpr (RENDER) sm.lfq[1][j][i] = v;
      dylf[j][i] = mdc[1] * (float)v / 1.0f;
    }
  }
  if constexpr (N == 32) {
    const float bias = kQuantBias[1];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = threadIdx.x + 256 * i;
      const float step = vsc[0][1] * mr[1][i];
      int32_t v = 0;
      if ((k % N) >= n || (k / N) >= n) v = hf_choose(co[1][0][k], step, bias, &bad);
      qs[1][0][k] = v;
    }
  } else {
    const float bias = kQuantBias[1];
    for (int e = threadIdx.x; e < VB * NN; e += T) {
      int j = e / NN, k = e % NN;
      float step = vsc[j][1] * tab->matrix[1][cb + k];
      int32_t v = 0;
      if ((k % N) >= n || (k / N) >= n) v = hf_choose(co[1][j][k], step, bias, &bad);
      qs[1][j][k] = v;
    }
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  int (*xs)[8] = sm.xs;
  if (pp->truncate) {
    if constexpr (N == 32) {
      // one varblock: all T/64 waves walk its luma block together
      truncate_pre<N, 4>(qs[1][0], co[1][0], nullptr, cellY, donY, matY, vsc[0][1], kQuantBias[1], pp->lambda[1],
                         pp->zero_token_bits, T / 64, wave, xs[0]);
    } else {
      for (int j = wave; j < nvb; j += T / 64)
        truncate_wave<N>(qs[1][j], co[1][j], nullptr, tab->order + cb, tab->matrix[1] + cb,
                         tab->donor + cb, vsc[j][1], kQuantBias[1], pp->lambda[1], pp->zero_token_bits);
    }
  }
  if constexpr (N == 32) {
    // the X/B truncation tables, requested before the X/B choices
    const int w4 = threadIdx.x >> 6, cxb = w4 < 2 ? 0 : 2, pxb = ((w4 & 1) * 64 + (threadIdx.x & 63)) * 8;
#pragma unroll
    for (int r = 0; r < 8; r++) {
      cellXB[r] = tab->order[cb + pxb + r];
      donXB[r] = tab->donor_pos[cb + pxb + r];
      matXB[r] = tab->matrix_pos[cxb][cb + pxb + r];
    }
  }
  // --- X and B LF against the reconstructed Y LF ---
  {
    const float klf[2] = {0.0f + (float)cflp->x_factor_lf / 84.0f, 1.0f + (float)cflp->b_factor_lf / 84.0f};
    for (int e = threadIdx.x; e < 2 * VB * n * n; e += T) {
      const int ci = e / (VB * n * n), j = (e / (n * n)) % VB, i = e % (n * n), c = ci == 0 ? 0 : 2;
      if (j < nvb) {
        const float target = lfo[c][j][i] - klf[ci] * dylf[j][i];
        int32_t* L = ci == 0 ? LX : LB;
        const int32_t v = clamp_round(target * 1.0f / mdc[c], &bad);
        L[(size_t)(vby[j] + i / n) * g.bw + vbx[j] + i % n] = v;
        if constexpr (RENDER) sm.lfq[c][j][i] = v;
      }
    }
  }
  __syncthreads();
  // --- X and B HF: targets against the reconstructed dY, then choose ---
  if constexpr (N == 32) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int ci = i / 4, c = ci == 0 ? 0 : 2, k = threadIdx.x + 256 * (i % 4);
      int32_t v = 0;
      if ((k % N) >= n || (k / N) >= n) {
        // the luma reconstruction after its truncation, from qs (a truncated
        // coefficient reconstructs to bias_adjust(0) * step = 0, as stored
        // before)
        const float dyk = bias_adjust(qs[1][0][k], kQuantBias[1]) * (vsc[0][1] * mr[1][i % 4]);
        const float target = co[c][0][k] - khf[ci][0] * dyk;
        co[c][0][k] = target;
        const float step = vsc[0][c] * mr[c][i % 4];
        v = hf_choose(target, step, kQuantBias[c], &bad);
      }
      qs[c][0][k] = v;
    }
  } else
  for (int e = threadIdx.x; e < 2 * VB * NN; e += T) {
    int ci = e / (VB * NN), rem = e % (VB * NN), j = rem / NN, k = rem % NN, c = ci == 0 ? 0 : 2;
    int32_t v = 0;
    if ((k % N) >= n || (k / N) >= n) {
      const float dyk = bias_adjust(qs[1][j][k], kQuantBias[1]) * (vsc[j][1] * tab->matrix[1][cb + k]);
      float target = co[c][j][k] - khf[ci][j] * dyk;
      co[c][j][k] = target;
      float step = vsc[j][c] * tab->matrix[c][cb + k];
      v = hf_choose(target, step, kQuantBias[c], &bad);
    }
    qs[c][j][k] = v;
  }
  __syncthreads();
  if (pp->truncate) {
    if constexpr (N == 32) {
      // X on the first half of the waves, B on the second half
      const int half = (T / 64) / 2, ci = wave / half, c = ci == 0 ? 0 : 2;
      truncate_pre<N, 8>(qs[c][0], co[c][0], nullptr, cellXB, donXB, matXB, vsc[0][c], kQuantBias[c],
                         pp->lambda[c], pp->zero_token_bits, half, wave % half, xs[ci]);
    } else {
      for (int w = wave; w < 2 * VB; w += T / 64) {
        const int ci = w / VB, j = w % VB, c = ci == 0 ? 0 : 2;
        if (j < nvb)
          truncate_wave<N>(qs[c][j], co[c][j], nullptr, tab->order + cb, tab->matrix[c] + cb,
                           tab->donor + cb, vsc[j][c], kQuantBias[c], pp->lambda[c], pp->zero_token_bits);
      }
    }
  }
  __syncthreads();
  int32_t* qout[3] = {QX, QY, QB};
  for (int e = threadIdx.x; e < 3 * VB * NN; e += T) {
    int c = e / (VB * NN), rem = e % (VB * NN), j = rem / NN, k = rem % NN;
    if (j < nvb) qout[c][(size_t)(8 * vby[j] + k / N) * g.cstride + 8 * vbx[j] + k % N] = qs[c][j][k];
  }
  if (bad) atomicOr(err, kErrQuantRange);
  if constexpr (RENDER) render_from_levels<N, T>(g, tab, pp, cflp, sm, nvb, P.rx[0], P.rx[1], P.rx[2], P.sigma);
}

// ----------------------------------------------------------------------------

// RENDER: the reconstruction of the exact DCT form follows in the same
// workgroup, from the levels in LDS (the render kernel's work for these
// varblocks, without reading the levels back from memory).
template <bool RENDER>
__global__ __launch_bounds__(256) void k_quantize_all(const ProbeDesc* __restrict__ D, const int* __restrict__ off8,
                                                      const int* __restrict__ off16, const int* __restrict__ off32,
                                                      int nd, int n32, int n16, const ConstTables* __restrict__ tab) {
  __shared__ QuantSmemAll sm;
  int b = blockIdx.x;  // DCT32 workgroups first (the longest), then DCT16, then DCT8
  if (b < n32) {
    quantize_body<32, RENDER>(D, off32, nd, b, tab, sm.s32);
    return;
  }
  b -= n32;
  if (b < n16) {
    quantize_body<16, RENDER>(D, off16, nd, b, tab, sm.s16);
    return;
  }
  quantize_body<8, RENDER>(D, off8, nd, b - n16, tab, sm.s8);
}

// ----------------------------------------------------------------------------
// Reconstruction (jpxl-plan-render/src/lib.rs:853-1032, :1294-1359)
// ----------------------------------------------------------------------------

// LDS of one render workgroup of side N.
template <int N>
struct RenderSmem {
  static constexpr int VB = 64 / N, P = N + 1;
  float lds[VB * 3 * N * P];
  float llf[VB][3][16];
  int vbx[VB], vby[VB];
  float khf[2][VB];
  float vsc[VB][4];  // step scales and sigma (the varblock's own HfMul under AQ)
};
union RenderSmemAll {
  RenderSmem<8> s8;
  RenderSmem<16> s16;
  RenderSmem<32> s32;
};

// One render workgroup (logical index `block` of the size's prefix `off`).
template <int N, int FORM>
__device__ __forceinline__ void render_body(const ProbeDesc* __restrict__ D, const int* __restrict__ off, int nd,
                                            int block, const ConstTables* __restrict__ tab,
                                            const DctMats* __restrict__ mats, RenderSmem<N>& sm) {
  const int d = find_desc(off, nd, block);
  DescSpace& PD = desc_at(D, d);
  const Geom g = geom_of(PD.g);
  const uint32_t* __restrict__ list = PD.list[size_index(N)];
  const int count = PD.count[size_index(N)];
  const ProbeParams* __restrict__ pp = PD.pp;
  const CflParams* __restrict__ cflp = PD.cflp;
  const int8_t* __restrict__ cfl_x = PD.cfl_x;
  const int8_t* __restrict__ cfl_b = PD.cfl_b;
  const int32_t* __restrict__ QX = PD.quant[0];
  const int32_t* __restrict__ QY = PD.quant[1];
  const int32_t* __restrict__ QB = PD.quant[2];
  const int32_t* __restrict__ LX = PD.lf[0];
  const int32_t* __restrict__ LY = PD.lf[1];
  const int32_t* __restrict__ LB = PD.lf[2];
  float* __restrict__ X = PD.rx[0];
  float* __restrict__ Y = PD.rx[1];
  float* __restrict__ B = PD.rx[2];
  float* __restrict__ sigma = PD.sigma;
  constexpr int VB = 64 / N, P = N + 1, NN = N * N, n = N / 8, si = size_index(N), cb = kCellBase[si];
  float* lds = sm.lds;
  float (*llf)[3][16] = sm.llf;
  int* vbx = sm.vbx;
  int* vby = sm.vby;
  float (*khf)[VB] = sm.khf;
  float (*vsc)[4] = sm.vsc;
  const int base = (block - off[d]) * VB;
  const int nvb = min(VB, count - base);
  // Every thread reads the workgroup's varblock positions (uniform loads) and
  // requests all its quantized coefficients and matrix entries at once,
  // before the per-varblock scales are computed.
  constexpr int IT = (VB * NN + kThreads - 1) / kThreads;
  int32_t qv[3][IT];
  float mv[3][IT];
  {
    int px[VB], py[VB];
#pragma unroll
    for (int j = 0; j < VB; j++) {
      px[j] = py[j] = 0;
      if (j < nvb) unpack(list[base + j], &px[j], &py[j]);
    }
    const int32_t* qin[3] = {QX, QY, QB};
#pragma unroll
    for (int i = 0; i < IT; i++) {
      const int e = threadIdx.x + kThreads * i, j = e / NN, k = e % NN;
      int bx = 0, by = 0;
#pragma unroll
      for (int jj = 0; jj < VB; jj++)
        if (jj == j) bx = px[jj], by = py[jj];
      const bool live = e < VB * NN && j < nvb;
      const uint32_t o = (uint32_t)((8 * by + k / N) * g.cstride + 8 * bx + k % N);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        qv[c][i] = live ? ldg32(qin[c], o) : 0;
        mv[c][i] = live ? tab->matrix[c][cb + k] : 0.0f;
      }
    }
  }
  if (threadIdx.x < VB) {
    int bx = 0, by = 0;
    if (threadIdx.x < nvb) unpack(list[base + threadIdx.x], &bx, &by);
    vbx[threadIdx.x] = bx;
    vby[threadIdx.x] = by;
    int tile = (by / 8) * g.tw + bx / 8;
    khf[0][threadIdx.x] = 0.0f + (float)cfl_x[tile] / 84.0f;
    khf[1][threadIdx.x] = 1.0f + (float)cfl_b[tile] / 84.0f;
    varblock_scales(pp, g.bw, g.bh, bx, by, n, vsc[threadIdx.x], &vsc[threadIdx.x][3]);
  }
  __syncthreads();
  // I.5.3 dequantization of all three channels, then I.6 HF CfL.
#pragma unroll
  for (int i = 0; i < IT; i++) {
    const int e = threadIdx.x + kThreads * i;
    if (e >= VB * NN) break;
    int j = e / NN, k = e % NN;
    float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f;
    if (j < nvb) {
      d0 = bias_adjust(qv[0][i], kQuantBias[0]) * vsc[j][0] * mv[0][i];
      d1 = bias_adjust(qv[1][i], kQuantBias[1]) * vsc[j][1] * mv[1][i];
      d2 = bias_adjust(qv[2][i], kQuantBias[2]) * vsc[j][2] * mv[2][i];
    }
    int r = k / N, col = k % N;
    lds[((j * 3 + 0) * N + r) * P + col] = d0 + khf[0][j] * d1;
    lds[((j * 3 + 1) * N + r) * P + col] = d1;
    lds[((j * 3 + 2) * N + r) * P + col] = d2 + khf[1][j] * d1;
  }
  __syncthreads();
  const DctTables& t = tab->dct;
  // I.8: the LLF cells from the dequantized, LF-CfL-corrected LF planes
  // (varblock.rs:817-855): the n x n DCT_2D of small_dct2_fwd, one lane per
  // line in each pass, then ScaleF.
  {
    const float kxl = 0.0f + (float)cflp->x_factor_lf / 84.0f;
    const float kbl = 1.0f + (float)cflp->b_factor_lf / 84.0f;
    const float* sf = tab->scale_f[si];
    if constexpr (n == 1) {
      if (threadIdx.x < 3 * VB) {
        const int j = threadIdx.x / 3, c = threadIdx.x % 3;
        if (j < nvb) {
          const size_t o = (size_t)vby[j] * g.bw + vbx[j];
          const float vy = pp->mdc[1] * (float)LY[o] / 1.0f;
          float v;
          if (c == 1) v = vy;
          else if (c == 0) v = pp->mdc[0] * (float)LX[o] / 1.0f + kxl * vy;
          else v = pp->mdc[2] * (float)LB[o] / 1.0f + kbl * vy;
          lds[(j * 3 + c) * N * P] = v;
        }
      }
    } else {
      const float f = fwd_factor<n>();
      const int j = threadIdx.x / (3 * n), c = (threadIdx.x / n) % 3, line = threadIdx.x % n;
      const bool act = threadIdx.x < 3 * VB * n && j < nvb;
      if (act) {
        // pass 1: column `line` of the LF samples
        float v[n];
#pragma unroll
        for (int r = 0; r < n; r++) {
          const size_t o = (size_t)(vby[j] + r) * g.bw + vbx[j] + line;
          const float vy = pp->mdc[1] * (float)LY[o] / 1.0f;
          if (c == 1) v[r] = vy;
          else if (c == 0) v[r] = pp->mdc[0] * (float)LX[o] / 1.0f + kxl * vy;
          else v[r] = pp->mdc[2] * (float)LB[o] / 1.0f + kbl * vy;
        }
        dct_ii<n>(v, t);
#pragma unroll
        for (int r = 0; r < n; r++) llf[j][c][r * n + line] = v[r] * f;
      }
      __syncthreads();
      if (act) {
        // pass 2: row `line` of pass 1 -> cells (k, line)
        float v[n];
#pragma unroll
        for (int k = 0; k < n; k++) v[k] = llf[j][c][line * n + k];
        dct_ii<n>(v, t);
        float* tile = lds + (j * 3 + c) * N * P;
#pragma unroll
        for (int k = 0; k < n; k++) tile[k * P + line] = v[k] * f * sf[k] * sf[line];
      }
    }
  }
  __syncthreads();
  // I.9: IDCT_2D.
  if constexpr (FORM != kDctExact) {
    const float* L = N == 8 ? mats->inv8 : (N == 16 ? mats->inv16 : mats->inv32);
    float* dsts[3] = {X, Y, B};
    matrix_transform_tiles<N, FORM>(lds, VB * 3, L, [&](int tt, int y, int x, float v) {
      const int jj = tt / 3;
      const int fx = 8 * vbx[jj] + x, fy = 8 * vby[jj] + y;
      if (jj < nvb && fx < g.w && fy < g.h) stg32(dsts[tt % 3], (uint32_t)(fy * g.w + fx), v);
    });
  } else {
  const int j = threadIdx.x / (3 * N), c = (threadIdx.x / N) % 3, line = threadIdx.x % N;
  float* tile = lds + (j * 3 + c) * N * P;
  inv_pass1<N>(tile, line, t);
  __syncthreads();
  float out[N];
  inv_pass2<N>(tile, line, out, t);
  if (j < nvb) {
    float* dst = c == 0 ? X : (c == 1 ? Y : B);
    const int fx = 8 * vbx[j] + line;
    if (fx < g.w) {
#pragma unroll
      for (int r = 0; r < N; r++) {
        int fy = 8 * vby[j] + r;
        if (fy < g.h) dst[(size_t)fy * g.w + fx] = out[r];
      }
    }
  }
  }
  // J.4.3 sigma of every 8x8 block the varblock covers.
  if (threadIdx.x < VB * n * n) {
    const int jj = threadIdx.x / (n * n), b = threadIdx.x % (n * n);
    if (jj < nvb) {
      int sx = vbx[jj] + b % n, sy = vby[jj] + b / n;
      if (sx < g.bw && sy < g.bh) sigma[(size_t)sy * g.bw + sx] = vsc[jj][3];
    }
  }
}

template <int FORM>
__global__ __launch_bounds__(kThreads) void k_render_all(const ProbeDesc* __restrict__ D,
                                                         const int* __restrict__ off8, const int* __restrict__ off16,
                                                         const int* __restrict__ off32, int nd, int n32, int n16,
                                                         const ConstTables* __restrict__ tab,
                                                         const DctMats* __restrict__ mats) {
  __shared__ RenderSmemAll sm;
  int b = blockIdx.x;  // DCT32 workgroups first (the longest), then DCT16, then DCT8
  if (b < n32) {
    render_body<32, FORM>(D, off32, nd, b, tab, mats, sm.s32);
    return;
  }
  b -= n32;
  if (b < n16) {
    render_body<16, FORM>(D, off16, nd, b, tab, mats, sm.s16);
    return;
  }
  render_body<8, FORM>(D, off8, nd, b - n16, tab, mats, sm.s8);
}

__global__ void k_fill(float* p, float v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ----------------------------------------------------------------------------
// J.3 Gaborish smoothing (jpxl-core/src/restoration.rs:148-193)
// ----------------------------------------------------------------------------

__device__ __forceinline__ void gaborish_px(int w, int h, const float* __restrict__ in, float* __restrict__ out,
                                            int x, int y) {
  const float w1 = 0.115169525f, w2 = 0.061248592f;
  const float sum = fmaf(4.0f, w2, fmaf(4.0f, w1, 1.0f));
  const float scale = 1.0f / sum;
  const float ce = scale, ed = w1 * scale, co = w2 * scale;
  if (x >= w) return;
  float acc = 0.0f;
  for (int dy = -1; dy <= 1; dy++) {
    const float* row = in + (size_t)mirror1d(y + dy, h) * w;
    for (int dx = -1; dx <= 1; dx++) {
      float wgt = (dx == 0 && dy == 0) ? ce : ((dx == 0 || dy == 0) ? ed : co);
      acc = fmaf(wgt, row[mirror1d(x + dx, w)], acc);
    }
  }
  out[(size_t)y * w + x] = acc;
}

__global__ __launch_bounds__(256) void k_gaborish(int w, int h, const float* __restrict__ in0,
                                                  const float* __restrict__ in1,
                                                  const float* __restrict__ in2,
                                                  float* __restrict__ o0, float* __restrict__ o1,
                                                  float* __restrict__ o2) {
  const int c = blockIdx.z;
  gaborish_px(w, h, c == 0 ? in0 : (c == 1 ? in1 : in2), c == 0 ? o0 : (c == 1 ? o1 : o2),
              blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// Encoder-side Gaborish sharpen (jpxl-encode/src/vardct/gaborish.rs:66-115,
// :218-315 as the default "simd" build evaluates it): one Jacobi step
// x' = x + (f - Kx)/centre per launch.  Kx: border pixels by the scalar 3x3
// loop over mirrored samples (products added in dy, dx order), interior
// columns in groups of four from x = 1 as mid*c + (n+s+w+e)*e +
// (nw+ne+sw+se)*k, the rest of a row's interior by the scalar loop; the update
// multiplies by 1/centre except on the last n % 4 samples of the plane,
// which divide.  No fused multiply-add (the reference does not contract).
__global__ __launch_bounds__(256) void k_gaborish_jacobi(int w, int h, const float* __restrict__ f0,
                                                         const float* __restrict__ f1,
                                                         const float* __restrict__ f2,
                                                         const float* __restrict__ x0,
                                                         const float* __restrict__ x1,
                                                         const float* __restrict__ x2, float* __restrict__ o0,
                                                         float* __restrict__ o1, float* __restrict__ o2) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, c = blockIdx.z;
  if (x >= w) return;
  const float* F = c == 0 ? f0 : (c == 1 ? f1 : f2);
  const float* X = c == 0 ? x0 : (c == 1 ? x1 : x2);
  float* O = c == 0 ? o0 : (c == 1 ? o1 : o2);
  const float w1 = 0.115169525f, w2 = 0.061248592f;
  const float a4 = 4.0f * w1, b4 = 4.0f * w2;
  const float unnorm = (1.0f + a4) + b4;
  const float ce = 1.0f / unnorm, ed = w1 * ce, co = w2 * ce;
  const bool edge = x == 0 || y == 0 || x + 1 == w || y + 1 == h;
  const bool grouped = !edge && (x - 1) < 4 * ((w - 2) / 4);
  float acc;
  if (grouped) {
    const float* r0 = X + (size_t)(y - 1) * w;
    const float* r1 = X + (size_t)y * w;
    const float* r2 = X + (size_t)(y + 1) * w;
    const float mid = r1[x] * ce;
    const float nsew = ((r0[x] + r2[x]) + r1[x - 1]) + r1[x + 1];
    const float diag = ((r0[x - 1] + r0[x + 1]) + r2[x - 1]) + r2[x + 1];
    acc = (mid + nsew * ed) + diag * co;
  } else {
    acc = 0.0f;
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const float wgt = (dx == 0 && dy == 0) ? ce : ((dx == 0 || dy == 0) ? ed : co);
        const float prod = wgt * X[(size_t)mirror1d(y + dy, h) * w + mirror1d(x + dx, w)];
        acc = acc + prod;
      }
  }
  const size_t i = (size_t)y * w + x, n = (size_t)w * h;
  const float d = F[i] - acc;
  O[i] = i < n - n % 4 ? X[i] + d * (1.0f / ce) : X[i] + d / ce;
}

// Batched: workgroup (row y, 256-column block) of channel blockIdx.z per desc;
// reads the render planes (src 0) or their ping-pong (src 1), writes the other.
__global__ __launch_bounds__(256) void k_gaborish_b(const ProbeDesc* __restrict__ D, const int* __restrict__ off,
                                                    int nd, int src) {
  const int d = find_desc(off, nd, blockIdx.x);
  DescSpace& P = desc_at(D, d);
  const int lb = blockIdx.x - off[d], xb = (P.g.w + 255) / 256;
  const int c = blockIdx.z;
  gaborish_px(P.g.w, P.g.h, src ? P.aux[c] : P.rx[c], src ? P.rx[c] : P.aux[c], (lb % xb) * 256 + threadIdx.x,
              lb / xb);
}

// ----------------------------------------------------------------------------
// J.4 edge-preserving filter step (jpxl-core/src/restoration.rs:583-731)
// ----------------------------------------------------------------------------

// 64x16 output tile per 256-thread workgroup (each lane four rows), halo 3.
constexpr int kEpfTW = 64, kEpfTH = 16, kEpfR = 3, kEpfRows = kEpfTH / 4;

constexpr int kEpfTapX[13] = {0, -1, 1, 0, 0, 1, 1, -1, -1, -2, 2, 0, 0};
constexpr int kEpfTapY[13] = {0, 0, 0, -1, 1, -1, 1, 1, -1, 0, 0, 2, -2};

// The channel-weighted distance of tap (kx, ky) at tile position (cx, cy):
// the sum over the three channels and (steps 0 and 1) the five-sample cross
// of |centre - tap| * scale, accumulated in that order.
template <int STEP, int LH, int LW>
__device__ __forceinline__ float epf_sad(const float (&tile)[3][LH][LW], int cx, int cy, int kx, int ky) {
  const float cscale[3] = {40.0f, 5.0f, 3.5f};
  float d = 0.0f;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if constexpr (STEP == 2) {
      d = fmaf(fabsf(tile[c][cy][cx] - tile[c][cy + ky][cx + kx]), cscale[c], d);
    } else {
#pragma unroll
      for (int jx = 0; jx < 5; jx++) {
        const int ix = kEpfTapX[jx], iy = kEpfTapY[jx];
        d = fmaf(fabsf(tile[c][cy + iy][cx + ix] - tile[c][cy + ky + iy][cx + kx + ix]), cscale[c], d);
      }
    }
  }
  return d;
}

// One output pixel from its taps' distances dist[tp].
template <int STEP, int LH, int LW>
__device__ __forceinline__ void epf_pixel_d(int w, int h, int x, int y, int tx, int ty,
                                            const float (&tile)[3][LH][LW], const float* __restrict__ sigma,
                                            float* __restrict__ o0, float* __restrict__ o1,
                                            float* __restrict__ o2, const float* dist) {
  if (x >= w || y >= h) return;
  const int bw = (w + 7) / 8;
  const size_t idx = (size_t)y * w + x;
  const int cx = tx + kEpfR, cy = ty + kEpfR;
  float* out[3] = {o0, o1, o2};
  const float s = sigma[(y / 8) * bw + x / 8];
  if (s < 0.3f) {
#pragma unroll
    for (int c = 0; c < 3; c++) out[c][idx] = tile[c][cy][cx];
    return;
  }
  constexpr float kScale = STEP == 0 ? 0.9f : (STEP == 1 ? 1.0f : 6.5f);
  const float step_mul = 1.65f * kScale;
  const float inv_sigma = step_mul * 4.0f * (1.0f - __fsqrt_rn(0.5f)) / s;
  const int rx = x % 8, ry = y % 8;
  const bool at_border = ry == 0 || ry == 7 || rx == 0 || rx == 7;
  const float pm = at_border ? 2.0f / 3.0f : 1.0f;
  constexpr int kTaps = STEP == 0 ? 13 : 5;
  float sum_w = 0.0f, acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f;
#pragma unroll
  for (int tp = 0; tp < kTaps; tp++) {
    const int kx = kEpfTapX[tp], ky = kEpfTapY[tp];
    const float sd = pm * dist[tp];
    float wgt;
    if (sd == 0.0f) {
      wgt = 1.0f;
    } else {
      const float v = fmaf(sd, -inv_sigma, 1.0f);
      wgt = v > 0.0f ? v : 0.0f;
    }
    sum_w += wgt;
    acc0 = fmaf(tile[0][cy + ky][cx + kx], wgt, acc0);
    acc1 = fmaf(tile[1][cy + ky][cx + kx], wgt, acc1);
    acc2 = fmaf(tile[2][cy + ky][cx + kx], wgt, acc2);
  }
  out[0][idx] = acc0 / sum_w;
  out[1][idx] = acc1 / sum_w;
  out[2][idx] = acc2 / sum_w;
}

template <int STEP, int LH, int LW>
__device__ __forceinline__ void epf_pixel(int w, int h, int x, int y, int tx, int ty,
                                          const float (&tile)[3][LH][LW], const float* __restrict__ sigma,
                                          float* __restrict__ o0, float* __restrict__ o1,
                                          float* __restrict__ o2) {
  constexpr int kTaps = STEP == 0 ? 13 : 5;
  float dist[kTaps];
#pragma unroll
  for (int tp = 0; tp < kTaps; tp++)
    dist[tp] = epf_sad<STEP>(tile, tx + kEpfR, ty + kEpfR, kEpfTapX[tp], kEpfTapY[tp]);
  epf_pixel_d<STEP>(w, h, x, y, tx, ty, tile, sigma, o0, o1, o2, dist);
}

// Step 1 over a lane's column of kEpfRows pixels, sharing distances: the
// centre tap's is 0 (the samples are finite), and the upward tap of a row is
// the downward tap of the row above (the same pairs in the same order, and
// |a - b| = |b - a|), so each row computes three of its five.
template <int LH, int LW>
__device__ __forceinline__ void epf_column_step1(int w, int h, int x, int y, int tx, int ty,
                                                 const float (&tile)[3][LH][LW], const float* __restrict__ sigma,
                                                 float* __restrict__ o0, float* __restrict__ o1,
                                                 float* __restrict__ o2) {
  // taps: 0 centre, 1 (-1,0), 2 (1,0), 3 (0,-1), 4 (0,1)
  float up = epf_sad<1>(tile, tx + kEpfR, ty + kEpfR, 0, -1);
#pragma unroll
  for (int r = 0; r < kEpfRows; r++) {
    const int cx = tx + kEpfR, cy = ty + r + kEpfR;
    float dist[5];
    dist[0] = 0.0f;
    dist[1] = epf_sad<1>(tile, cx, cy, -1, 0);
    dist[2] = epf_sad<1>(tile, cx, cy, 1, 0);
    dist[3] = up;
    dist[4] = epf_sad<1>(tile, cx, cy, 0, 1);
    up = dist[4];
    epf_pixel_d<1>(w, h, x, y + r, tx, ty + r, tile, sigma, o0, o1, o2, dist);
  }
}

template <int STEP>
__device__ __forceinline__ void epf_tile(int w, int h, const float* __restrict__ in0, const float* __restrict__ in1,
                                         const float* __restrict__ in2, const float* __restrict__ sigma,
                                         float* __restrict__ o0, float* __restrict__ o1, float* __restrict__ o2,
                                         int x0, int y0) {
  constexpr int LW = kEpfTW + 2 * kEpfR, LH = kEpfTH + 2 * kEpfR;
  __shared__ float tile[3][LH][LW];
  const float* in[3] = {in0, in1, in2};
  // all of this thread's halo-tile samples are requested before any is stored
  constexpr int kLoads = (3 * LH * LW + 255) / 256;
  float v[kLoads];
#pragma unroll
  for (int i = 0; i < kLoads; i++) {
    const int e = threadIdx.x + 256 * i;
    const int c = e / (LH * LW), rem = e % (LH * LW), ly = rem / LW, lx = rem % LW;
    const int sy = mirror1d(y0 + ly - kEpfR, h), sx = mirror1d(x0 + lx - kEpfR, w);
    v[i] = e < 3 * LH * LW ? ldg32(in[c < 3 ? c : 2], (uint32_t)(sy * w + sx)) : 0.0f;
  }
#pragma unroll
  for (int i = 0; i < kLoads; i++) {
    const int e = threadIdx.x + 256 * i;
    if (e < 3 * LH * LW) (&tile[0][0][0])[e] = v[i];
  }
  __syncthreads();
  const int tx = threadIdx.x % kEpfTW;
  if constexpr (STEP == 1) {
    const int ty = (threadIdx.x / kEpfTW) * kEpfRows;
    epf_column_step1(w, h, x0 + tx, y0 + ty, tx, ty, tile, sigma, o0, o1, o2);
  } else {
    for (int ty = (threadIdx.x / kEpfTW) * kEpfRows; ty < (int)(threadIdx.x / kEpfTW + 1) * kEpfRows; ty++)
      epf_pixel<STEP>(w, h, x0 + tx, y0 + ty, tx, ty, tile, sigma, o0, o1, o2);
  }
}

template <int STEP>
__global__ __launch_bounds__(256) void k_epf(int w, int h, const float* __restrict__ in0,
                                             const float* __restrict__ in1,
                                             const float* __restrict__ in2,
                                             const float* __restrict__ sigma,
                                             float* __restrict__ o0, float* __restrict__ o1,
                                             float* __restrict__ o2) {
  epf_tile<STEP>(w, h, in0, in1, in2, sigma, o0, o1, o2, blockIdx.x * kEpfTW, blockIdx.y * kEpfTH);
}

// Batched: one 64x16 tile of one desc per workgroup; src as for Gaborish.
template <int STEP>
__global__ __launch_bounds__(256) void k_epf_b(const ProbeDesc* __restrict__ D, const int* __restrict__ off,
                                               int nd, int src) {
  const int d = find_desc(off, nd, blockIdx.x);
  DescSpace& P = desc_at(D, d);
  const int lb = blockIdx.x - off[d], tx = (P.g.w + kEpfTW - 1) / kEpfTW;
  float* in[3] = {src ? P.aux[0] : P.rx[0], src ? P.aux[1] : P.rx[1], src ? P.aux[2] : P.rx[2]};
  float* out[3] = {src ? P.rx[0] : P.aux[0], src ? P.rx[1] : P.aux[1], src ? P.rx[2] : P.aux[2]};
  epf_tile<STEP>(P.g.w, P.g.h, in[0], in[1], in[2], P.sigma, out[0], out[1], out[2], (lb % tx) * kEpfTW,
                 (lb / tx) * kEpfTH);
}

inline unsigned div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

// ----------------------------------------------------------------------------
// Launchers
// ----------------------------------------------------------------------------

void launch_xyb_fwd(hipStream_t s, const Geom& g, int bits, const void* rgb, const float* lut,
                    float* xyb[3], float* lin[3], int* gray) {
  int n = g.w * g.h;
  if (bits == 8)
    k_xyb_fwd<true><<<div_up(n, 256), 256, 0, s>>>(n, rgb, lut, xyb[0], xyb[1], xyb[2], lin[0],
                                                   lin[1], lin[2], gray);
  else
    k_xyb_fwd<false><<<div_up(n, 256), 256, 0, s>>>(n, rgb, lut, xyb[0], xyb[1], xyb[2], lin[0],
                                                    lin[1], lin[2], gray);
}

int summary_groups(const Geom& g) { return (int)div_up((size_t)g.bw * g.bh, kSummaryBlocks); }

void launch_block_features(hipStream_t s, const Geom& g, const ConstTables* tab, float* const xyb[3],
                           float* luma_var, float* chroma_var, SummaryPartial* parts, uint32_t* hist) {
  k_fill_u32<<<div_up(kSummaryBins, 256), 256, 0, s>>>(hist, kSummaryBins, 0u);
  k_dct8_summary<<<summary_groups(g), kThreads, 0, s>>>(g, tab, xyb[0], xyb[1], xyb[2], luma_var, chroma_var,
                                                         parts, hist);
}

template <int N, int FORM>
static void dct_fwd_form(hipStream_t s, const Geom& g, const uint32_t* list, int count, const StructInfo* si,
                         const ConstTables* tab, const DctMats* mats, float* const xyb[3], float* coeff[3]) {
  const int n = N / 8, bound = si ? (g.bw / n) * (g.bh / n) : count;
  if (bound <= 0) return;
  k_dct_fwd<N, FORM><<<div_up(bound, 64 / N), kThreads, 0, s>>>(g, list, count, si, tab, mats, xyb[0], xyb[1],
                                                                 xyb[2], coeff[0], coeff[1], coeff[2]);
}

template <int N>
static void dct_fwd_size(hipStream_t s, int form, const Geom& g, const uint32_t* list, int count,
                         const StructInfo* si, const ConstTables* tab, const DctMats* mats, float* const xyb[3],
                         float* coeff[3]) {
  if (form == kDctFp32) dct_fwd_form<N, kDctFp32>(s, g, list, count, si, tab, mats, xyb, coeff);
  else if (form == kDctBf16x3) dct_fwd_form<N, kDctBf16x3>(s, g, list, count, si, tab, mats, xyb, coeff);
  else dct_fwd_form<N, kDctExact>(s, g, list, count, si, tab, mats, xyb, coeff);
}

void launch_dct_fwd(hipStream_t s, int side, int form, const Geom& g, const uint32_t* list, int count,
                    const StructInfo* si, const ConstTables* tab, const DctMats* mats, float* const xyb[3],
                    float* coeff[3]) {
  if (!si && count <= 0) return;
  switch (side) {
    case 8: dct_fwd_size<8>(s, form, g, list, count, si, tab, mats, xyb, coeff); break;
    case 16: dct_fwd_size<16>(s, form, g, list, count, si, tab, mats, xyb, coeff); break;
    default: dct_fwd_size<32>(s, form, g, list, count, si, tab, mats, xyb, coeff); break;
  }
}

void launch_dct_blocks(hipStream_t s, int side, int inverse, int count, const ConstTables* tab,
                       const float* in, float* out) {
  if (count <= 0) return;
  switch (side) {
    case 8: k_dct_blocks<8><<<div_up(count, 24), kThreads, 0, s>>>(inverse, count, tab, in, out); break;
    case 16: k_dct_blocks<16><<<div_up(count, 12), kThreads, 0, s>>>(inverse, count, tab, in, out); break;
    default: k_dct_blocks<32><<<div_up(count, 6), kThreads, 0, s>>>(inverse, count, tab, in, out); break;
  }
}

int quant_groups(int side, int count) {
  return side == 8 ? (int)div_up(count, quant_vb(8)) : (side == 16 ? (int)div_up(count, quant_vb(16))
                                                                    : (int)div_up(count, quant_vb(32)));
}
int render_groups(int side, int count) { return (int)div_up(count, 64 / side); }

void launch_quantize_all(hipStream_t s, const ProbeDesc* D, const int* const off[3], int nd, const int nwg[3],
                         const ConstTables* tab, bool render) {
  const int total = nwg[0] + nwg[1] + nwg[2];
  if (total <= 0) return;
  if (render)
    k_quantize_all<true><<<total, 256, 0, s>>>(D, off[0], off[1], off[2], nd, nwg[2], nwg[1], tab);
  else
    k_quantize_all<false><<<total, 256, 0, s>>>(D, off[0], off[1], off[2], nd, nwg[2], nwg[1], tab);
}

void launch_render_all(hipStream_t s, int form, const ProbeDesc* D, const int* const off[3], int nd,
                       const int nwg[3], const ConstTables* tab, const DctMats* mats) {
  const int total = nwg[0] + nwg[1] + nwg[2];
  if (total <= 0) return;
  if (form == kDctFp32)
    k_render_all<kDctFp32><<<total, kThreads, 0, s>>>(D, off[0], off[1], off[2], nd, nwg[2], nwg[1], tab, mats);
  else if (form == kDctBf16x3)
    k_render_all<kDctBf16x3><<<total, kThreads, 0, s>>>(D, off[0], off[1], off[2], nd, nwg[2], nwg[1], tab, mats);
  else
    k_render_all<kDctExact><<<total, kThreads, 0, s>>>(D, off[0], off[1], off[2], nd, nwg[2], nwg[1], tab, mats);
}

void launch_fill(hipStream_t s, float* p, float v, size_t n) {
  if (n == 0) return;
  k_fill<<<div_up(n, 256), 256, 0, s>>>(p, v, n);
}

void launch_gaborish_jacobi(hipStream_t s, int w, int h, float* const f[3], float* const x[3], float* out[3]) {
  dim3 grid(div_up(w, 256), h, 3);
  k_gaborish_jacobi<<<grid, 256, 0, s>>>(w, h, f[0], f[1], f[2], x[0], x[1], x[2], out[0], out[1], out[2]);
}

void launch_gaborish(hipStream_t s, int w, int h, float* const in[3], float* out[3]) {
  dim3 grid(div_up(w, 256), h, 3);
  k_gaborish<<<grid, 256, 0, s>>>(w, h, in[0], in[1], in[2], out[0], out[1], out[2]);
}

int epf_tiles(int w, int h) { return (int)(div_up(w, kEpfTW) * div_up(h, kEpfTH)); }
int gaborish_blocks(int w, int h) { return (int)(div_up(w, 256) * h); }

void launch_gaborish_b(hipStream_t s, const ProbeDesc* D, const int* off, int nd, int nwg, int src) {
  if (nwg <= 0) return;
  k_gaborish_b<<<dim3(nwg, 1, 3), 256, 0, s>>>(D, off, nd, src);
}

void launch_epf_b(hipStream_t s, int step, const ProbeDesc* D, const int* off, int nd, int nwg, int src) {
  if (nwg <= 0) return;
  if (step == 0) k_epf_b<0><<<nwg, 256, 0, s>>>(D, off, nd, src);
  else if (step == 1) k_epf_b<1><<<nwg, 256, 0, s>>>(D, off, nd, src);
  else k_epf_b<2><<<nwg, 256, 0, s>>>(D, off, nd, src);
}

void launch_epf(hipStream_t s, int step, int w, int h, float* const in[3], const float* sigma,
                float* out[3]) {
  dim3 grid(div_up(w, kEpfTW), div_up(h, kEpfTH));
  if (step == 0)
    k_epf<0><<<grid, 256, 0, s>>>(w, h, in[0], in[1], in[2], sigma, out[0], out[1], out[2]);
  else if (step == 1)
    k_epf<1><<<grid, 256, 0, s>>>(w, h, in[0], in[1], in[2], sigma, out[0], out[1], out[2]);
  else
    k_epf<2><<<grid, 256, 0, s>>>(w, h, in[0], in[1], in[2], sigma, out[0], out[1], out[2]);
}

}  // namespace jpxlg
