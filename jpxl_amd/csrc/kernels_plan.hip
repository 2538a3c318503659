// gfx950 kernels for the plan structure: the hierarchical cover (AC strategy)
// and the chroma-from-luma estimate.
//
// Cover (jpxl-encode-policy/src/lib.rs:4909-5302): every square candidate
// (DCT8 at every block, DCT16 at every aligned 2x2, DCT32 at every aligned
// 4x4 that fits the block grid) is forward transformed in LDS, quantized
// with the nearest rule and priced as
//   J = bits * rate_scale + sum lambda_c * side^2 * w(cell) * (recon - t)^2
// with the sum taken in the reference's order (Y cells, then X, then B, each
// row-major over the non-LLF cells) so the f64 value is the reference's own.
// One thread per 32x32 region then walks the quadtree: a merge wins only when
// strictly cheaper than its split.
#include "internal.h"

namespace jpxlg {

namespace {

constexpr int kThreads = 192;

inline unsigned div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

__device__ __constant__ float kBias[3] = {1.0f - 0.054650073f, 1.0f - 0.0700545f,
                                          1.0f - 0.049935103f};

// CoverRateModel::Calibrated (request.rs:585-636).
__host__ __device__ constexpr double rate_scale(int side) { return side == 16 ? 1.63 : (side == 32 ? 1.88 : 1.56); }
__host__ __device__ constexpr double fixed_bits(int side) { return side == 16 ? 8.4 + 4.0 : (side == 32 ? 6.5 + 4.0 : 5.5); }

__device__ __forceinline__ int32_t choose_nearest(float target, float step, float bias, int* bad) {
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

// lib.rs:3794-3800 residual_bits.
__device__ __forceinline__ uint32_t residual_bits(int32_t q) {
  return q == 0 ? 0u : (uint32_t)(32 - __clz((unsigned)abs(q))) + 1u;
}

template <int N>
__device__ __forceinline__ void fwd_line_pass1(float* tile, int col, const DctTables& t) {
  constexpr int P = N + 1;
  float v[N];
#pragma unroll
  for (int r = 0; r < N; r++) v[r] = tile[r * P + col];
  dct_ii<N>(v, t);
  const float f = fwd_factor<N>();
#pragma unroll
  for (int r = 0; r < N; r++) tile[r * P + col] = v[r] * f;
}

template <int N>
__global__ __launch_bounds__(kThreads) void k_cover_cost(Geom g, const ConstTables* __restrict__ tab,
                                                         const ProbeParams* __restrict__ pp,
                                                         const float* __restrict__ X,
                                                         const float* __restrict__ Y,
                                                         const float* __restrict__ B,
                                                         double* __restrict__ cost,
                                                         uint32_t* __restrict__ err) {
  constexpr int VB = 64 / N, P = N + 1, NN = N * N, n = N / 8, si = size_index(N), cb = kCellBase[si];
  __shared__ float tile[VB * 3 * N * P];
  __shared__ float dyh[VB][NN];
  __shared__ double terms[VB][NN];
  __shared__ uint32_t bits[VB];
  __shared__ double sse[VB];
  __shared__ int vbx[VB], vby[VB], valid[VB];
  const int ncx = g.bw / n, ncy = g.bh / n, count = ncx * ncy;
  const int base = blockIdx.x * VB;
  if (threadIdx.x < VB) {
    int i = base + threadIdx.x;
    valid[threadIdx.x] = i < count;
    vbx[threadIdx.x] = i < count ? (i % ncx) * n : 0;
    vby[threadIdx.x] = i < count ? (i / ncx) * n : 0;
    bits[threadIdx.x] = 0;
    sse[threadIdx.x] = 0.0;
  }
  __syncthreads();
  const float* planes[3] = {X, Y, B};
  for (int e = threadIdx.x; e < VB * 3 * NN; e += kThreads) {
    int j = e / (3 * NN), rem = e % (3 * NN), c = rem / NN, k = rem % NN;
    int dy = k / N, dx = k % N;
    int x = min(8 * vbx[j] + dx, g.w - 1), y = min(8 * vby[j] + dy, g.h - 1);
    tile[((j * 3 + c) * N + dy) * P + dx] = planes[c][(size_t)y * g.w + x];
  }
  __syncthreads();
  const DctTables& t = tab->dct;
  {
    const int j = threadIdx.x / (3 * N), c = (threadIdx.x / N) % 3, line = threadIdx.x % N;
    float* tl = tile + (j * 3 + c) * N * P;
    fwd_line_pass1<N>(tl, line, t);
    __syncthreads();
    float out[N];
#pragma unroll
    for (int k = 0; k < N; k++) out[k] = tl[line * P + k];
    dct_ii<N>(out, t);
    const float f = fwd_factor<N>();
    __syncthreads();
    // coefficient cell k*N + line -> tile row k, column line
#pragma unroll
    for (int k = 0; k < N; k++) tl[k * P + line] = out[k] * f;
  }
  __syncthreads();
  int bad = 0;
  const double tsd = (double)(N * N) * 1.0;
  for (int ci = 0; ci < 3; ci++) {
    const int c = ci == 0 ? 1 : (ci == 1 ? 0 : 2);  // Y, then X, then B
    const float sc = pp->sc[c], bias = kBias[c];
    const double lw = pp->lambda[c] * tsd;
    for (int e = threadIdx.x; e < VB * NN; e += kThreads) {
      int j = e / NN, k = e % NN;
      if ((k % N) < n && (k / N) < n) continue;
      float coef = tile[((j * 3 + c) * N + k / N) * P + k % N];
      float target;
      if (c == 1) target = coef;
      else if (c == 0) target = coef - 0.0f * dyh[j][k];
      else target = coef - 1.0f * dyh[j][k];
      float step = sc * tab->matrix[c][cb + k];
      int32_t q = choose_nearest(target, step, bias, &bad);
      float recon = bias_adjust(q, bias) * step;
      if (c == 1) dyh[j][k] = recon;
      double d = (double)(recon - target);
      terms[j][k] = lw * (double)tab->donor[cb + k] * (d * d);
      uint32_t rb = residual_bits(q);
      if (rb) atomicAdd(&bits[j], rb);
    }
    __syncthreads();
    // Sequential f64 sum in cell order, one thread per candidate.
    if (threadIdx.x < VB) {
      const int j = threadIdx.x;
      double s = sse[j];
      for (int k = 0; k < NN; k++) {
        if ((k % N) < n && (k / N) < n) continue;
        s += terms[j][k];
      }
      sse[j] = s;
    }
    __syncthreads();
  }
  if (threadIdx.x < VB && valid[threadIdx.x]) {
    const int j = threadIdx.x;
    cost[(size_t)vby[j] * g.bw + vbx[j]] = (double)bits[j] * rate_scale(N) + sse[j];
  }
  if (bad) atomicOr(err, kErrQuantRange);
}

// The quadtree of tile_region_with over one 32x32 region (4x4 blocks).
__global__ __launch_bounds__(256) void k_cover_decide(Geom g, const double* __restrict__ c8,
                                                      const double* __restrict__ c16,
                                                      const double* __restrict__ c32,
                                                      uint8_t* __restrict__ bmap) {
  const int rx = blockIdx.x * blockDim.x + threadIdx.x, ry = blockIdx.y;
  const int nrx = (g.bw + 3) / 4;
  if (rx >= nrx) return;
  const int bx0 = 4 * rx, by0 = 4 * ry;
  const int bw = g.bw, bh = g.bh;
  uint8_t dec[4][4];
#pragma unroll
  for (int y = 0; y < 4; y++)
#pragma unroll
    for (int x = 0; x < 4; x++) dec[y][x] = 0;
  double split4 = 0.0;
  bool q16[2][2] = {{false, false}, {false, false}};
  const int qo[4][2] = {{0, 0}, {2, 0}, {0, 2}, {2, 2}};
  for (int qi = 0; qi < 4; qi++) {
    const int qx = bx0 + qo[qi][0], qy = by0 + qo[qi][1];
    if (qx >= bw || qy >= bh) continue;  // contributes 0
    double split2 = 0.0;
    const int so[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
    for (int si = 0; si < 4; si++) {
      const int sx = qx + so[si][0], sy = qy + so[si][1];
      if (sx >= bw || sy >= bh) continue;
      split2 += (c8[(size_t)sy * bw + sx] + fixed_bits(8)) + 0.0;
    }
    double best2 = split2;
    if (qx + 2 <= bw && qy + 2 <= bh) {
      const double fixed = fixed_bits(16) + 0.0;
      if (!(fixed >= split2)) {
        double single = c16[(size_t)qy * bw + qx] + fixed;
        if (single < split2) {
          best2 = single;
          q16[qo[qi][1] / 2][qo[qi][0] / 2] = true;
        }
      }
    }
    split4 += best2;
  }
  bool use32 = false;
  if (bx0 + 4 <= bw && by0 + 4 <= bh) {
    const double fixed = fixed_bits(32) + 0.0;
    if (!(fixed >= split4)) {
      double single = c32[(size_t)by0 * bw + bx0] + fixed;
      if (single < split4) use32 = true;
    }
  }
  if (use32) {
    dec[0][0] = 32;
  } else {
    for (int qy = 0; qy < 2; qy++)
      for (int qx = 0; qx < 2; qx++) {
        if (q16[qy][qx]) {
          dec[2 * qy][2 * qx] = 16;
        } else {
          dec[2 * qy][2 * qx] = 8;
          dec[2 * qy][2 * qx + 1] = 8;
          dec[2 * qy + 1][2 * qx] = 8;
          dec[2 * qy + 1][2 * qx + 1] = 8;
        }
      }
  }
  for (int y = 0; y < 4; y++)
    for (int x = 0; x < 4; x++)
      if (bx0 + x < bw && by0 + y < bh) bmap[(size_t)(by0 + y) * bw + bx0 + x] = dec[y][x];
}

}  // namespace

void launch_cover(hipStream_t s, const Geom& g, const ConstTables* tab, const ProbeParams* pp,
                  float* const xyb[3], double* cost8, double* cost16, double* cost32, uint8_t* bmap,
                  uint32_t* err) {
  int c8 = g.bw * g.bh, c16 = (g.bw / 2) * (g.bh / 2), c32 = (g.bw / 4) * (g.bh / 4);
  k_cover_cost<8><<<div_up(c8, 8), kThreads, 0, s>>>(g, tab, pp, xyb[0], xyb[1], xyb[2], cost8, err);
  if (c16 > 0)
    k_cover_cost<16><<<div_up(c16, 4), kThreads, 0, s>>>(g, tab, pp, xyb[0], xyb[1], xyb[2], cost16, err);
  if (c32 > 0)
    k_cover_cost<32><<<div_up(c32, 2), kThreads, 0, s>>>(g, tab, pp, xyb[0], xyb[1], xyb[2], cost32, err);
  dim3 grid(div_up((g.bw + 3) / 4, 256), (g.bh + 3) / 4);
  k_cover_decide<<<grid, 256, 0, s>>>(g, cost8, cost16, cost32, bmap);
}

}  // namespace jpxlg
