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
#include <algorithm>
#include <climits>
#include <cstddef>

#include "internal.h"

namespace jpxlg {

namespace {

constexpr int kThreads = 192;

inline unsigned div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

__device__ __constant__ float kBias[3] = {1.0f - 0.054650073f, 1.0f - 0.0700545f,
                                          1.0f - 0.049935103f};

// The cover's rate model (CoverRateModel, request.rs:585-636) and cell
// weights, from the probe parameters: calibrated + quantizer-donor for the
// quality and target-rate requests, legacy + flat for fixed quantizers.
__device__ __forceinline__ double rate_scale(const ProbeParams* pp, int side) {
  return pp->cover_rate[side == 8 ? 0 : (side == 16 ? 1 : 2)];
}
__device__ __forceinline__ double fixed_bits(const ProbeParams* pp, int side) {
  return pp->cover_fixed[side == 8 ? 0 : (side == 16 ? 1 : 2)];
}

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

struct DD {
  double hi, lo;
};

__device__ __forceinline__ DD dd_add(DD a, DD b) {
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  const double err = (a.hi - (s - bb)) + (b.hi - bb);
  const double e = err + (a.lo + b.lo);
  const double hi = s + e;
  return DD{hi, e - (hi - s)};
}

__device__ __forceinline__ DD dd_of(double v) { return DD{v, 0.0}; }

// The cost J of up to VB square candidates of side N whose top-left blocks
// are vbx/vby (LDS, set by the caller; valid[j] = 0 for an unused slot).
// EXACT: the f64 distortion sum is sequential in the reference's order (Y,
// X, B, cells in raster order; one lane per candidate).  Otherwise the sum is
// formed exactly as a double-double in any order and rad[] receives a bound
// on |sequential - exact| (the terms are non-negative, so gamma_(n) * S),
// for decisions that hold for every sum within it.
template <int N, bool EXACT>
__device__ void cover_cost_group(Geom g, const ConstTables* __restrict__ tab, const ProbeParams* __restrict__ pp,
                                 const float* __restrict__ X, const float* __restrict__ Y,
                                 const float* __restrict__ B, const int* vbx, const int* vby, const int* valid,
                                 double* __restrict__ cost, double* __restrict__ rad, uint32_t* __restrict__ err) {
  constexpr int VB = 64 / N, P = N + 1, NN = N * N, n = N / 8, si = size_index(N), cb = kCellBase[si];
  constexpr int G = kThreads / VB;  // threads per candidate for the double-double sums
  __shared__ float tile[VB * 3 * N * P];
  // the exact form keeps every term for its sequential sums; the
  // double-double form needs only the reduction of the lanes' partial sums
  constexpr int TN = EXACT ? NN : 1;
  __shared__ float dyh[VB][TN];
  __shared__ double terms[VB][TN];
  __shared__ DD red[EXACT ? 1 : kThreads];
  __shared__ uint32_t bits[VB];
  __shared__ double sse[VB];
  __shared__ float vsc[VB][3];  // the candidate's step scales at its footprint's HfMul
  if (threadIdx.x < VB) {
    bits[threadIdx.x] = 0;
    sse[threadIdx.x] = 0.0;
    varblock_scales(pp, g.bw, g.bh, vbx[threadIdx.x], vby[threadIdx.x], n, vsc[threadIdx.x], nullptr);
  }
  __syncthreads();
  const float* planes[3] = {X, Y, B};
  // every sample of this thread is requested before any is stored
  constexpr int GL = VB * 3 * NN / kThreads;
  static_assert(GL * kThreads == VB * 3 * NN, "gather split");
  float gv[GL];
#pragma unroll
  for (int i = 0; i < GL; i++) {
    const int e = threadIdx.x + kThreads * i;
    const int j = e / (3 * NN), rem = e % (3 * NN), c = rem / NN, k = rem % NN;
    const int x = min(8 * vbx[j] + k % N, g.w - 1), y = min(8 * vby[j] + k / N, g.h - 1);
    gv[i] = ldg32(planes[c], (uint32_t)(y * g.w + x));
  }
#pragma unroll
  for (int i = 0; i < GL; i++) {
    const int e = threadIdx.x + kThreads * i;
    const int j = e / (3 * NN), rem = e % (3 * NN), c = rem / NN, k = rem % NN;
    tile[((j * 3 + c) * N + k / N) * P + k % N] = gv[i];
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
  if constexpr (EXACT) {
    for (int ci = 0; ci < 3; ci++) {
      const int c = ci == 0 ? 1 : (ci == 1 ? 0 : 2);  // Y, then X, then B
      const float bias = kBias[c];
      const double lw = pp->lambda[c] * tsd;
      for (int e = threadIdx.x; e < VB * NN; e += kThreads) {
        int j = e / NN, k = e % NN;
        if ((k % N) < n && (k / N) < n) {
          terms[j][k] = 0.0;
          continue;
        }
        float coef = tile[((j * 3 + c) * N + k / N) * P + k % N];
        float target;
        if (c == 1) target = coef;
        else if (c == 0) target = coef - 0.0f * dyh[j][k];
        else target = coef - 1.0f * dyh[j][k];
        float step = vsc[j][c] * tab->matrix[c][cb + k];
        int32_t q = choose_nearest(target, step, bias, &bad);
        float recon = bias_adjust(q, bias) * step;
        if (c == 1) dyh[j][k] = recon;
        double d = (double)(recon - target);
        terms[j][k] = lw * (double)(pp->cover_flat ? 1.0f : tab->donor[cb + k]) * (d * d);
        uint32_t rb = residual_bits(q);
        if (rb) atomicAdd(&bits[j], rb);
      }
      __syncthreads();
      // Sequential f64 sum in cell order, one thread per candidate.  The LLF
      // cells hold +0.0, which leaves the (non-negative) sum unchanged.
      if (threadIdx.x < VB) sse[threadIdx.x] = sequential_sum(terms[threadIdx.x], NN, sse[threadIdx.x]);
      __syncthreads();
    }
    if (threadIdx.x < VB && valid[threadIdx.x]) {
      const int j = threadIdx.x;
      cost[(size_t)vby[j] * g.bw + vbx[j]] = (double)bits[j] * rate_scale(pp, N) + sse[j];
    }
  } else {
    // Lane r of candidate j's G lanes takes its cells k = r + m G through
    // all three channels (Y first: X and B subtract its reconstruction), and
    // adds their terms to its own double-double: the exact sum does not
    // depend on the order.  The tables of its cells are requested first.
    constexpr int MK = (NN + G - 1) / G;
    const int j = threadIdx.x / G, r = threadIdx.x % G;
    float mt[3][MK], dn[MK];
#pragma unroll
    for (int m = 0; m < MK; m++) {
      const int k = min(r + m * G, NN - 1);
#pragma unroll
      for (int c = 0; c < 3; c++) mt[c][m] = tab->matrix[c][cb + k];
      dn[m] = pp->cover_flat ? 1.0f : tab->donor[cb + k];
    }
    const double lw[3] = {pp->lambda[0] * tsd, pp->lambda[1] * tsd, pp->lambda[2] * tsd};
    DD part{0.0, 0.0};
    uint32_t nbits = 0;
#pragma unroll
    for (int m = 0; m < MK; m++) {
      const int k = r + m * G;
      if (k >= NN || ((k % N) < n && (k / N) < n)) continue;
      float dy = 0.0f;
#pragma unroll
      for (int ci = 0; ci < 3; ci++) {
        const int c = ci == 0 ? 1 : (ci == 1 ? 0 : 2);
        const float bias = kBias[c];
        const float coef = tile[((j * 3 + c) * N + k / N) * P + k % N];
        float target;
        if (c == 1) target = coef;
        else if (c == 0) target = coef - 0.0f * dy;
        else target = coef - 1.0f * dy;
        const float step = vsc[j][c] * mt[c][m];
        const int32_t q = choose_nearest(target, step, bias, &bad);
        const float recon = bias_adjust(q, bias) * step;
        if (c == 1) dy = recon;
        const double d = (double)(recon - target);
        part = dd_add(part, dd_of(lw[c] * (double)dn[m] * (d * d)));
        nbits += residual_bits(q);
      }
    }
    if (nbits) atomicAdd(&bits[j], nbits);
    red[threadIdx.x] = part;
    __syncthreads();
    if (threadIdx.x % G == 0) {
      DD tot = red[threadIdx.x];
      for (int rr = 1; rr < G; rr++) tot = dd_add(tot, red[threadIdx.x + rr]);
      if (valid[j]) {
        const double S = tot.hi + tot.lo;
        const double nterms = 3.0 * NN + 4.0, u = 1.1102230246251565e-16;
        const double gam = 2.0 * (nterms * u / (1.0 - nterms * u));
        const double cst = (double)bits[j] * rate_scale(pp, N) + S;
        const size_t o = (size_t)vby[j] * g.bw + vbx[j];
        cost[o] = cst;
        // the sequential sum's distance from S, plus the roundings of S and
        // of the cost
        rad[o] = gam * S + 4.0 * u * fabs(cst) + 1e-300;
      }
    }
  }
  if (bad) atomicOr(err, kErrQuantRange);
}

template <int N>
__global__ __launch_bounds__(kThreads) void k_cover_cost(Geom g, const ConstTables* __restrict__ tab,
                                                         const ProbeParams* __restrict__ pp,
                                                         const float* __restrict__ X,
                                                         const float* __restrict__ Y,
                                                         const float* __restrict__ B,
                                                         double* __restrict__ cost, double* __restrict__ rad,
                                                         uint32_t* __restrict__ err) {
  constexpr int VB = 64 / N, n = N / 8;
  __shared__ int vbx[VB], vby[VB], valid[VB];
  const int ncx = g.bw / n, ncy = g.bh / n, count = ncx * ncy;
  const int base = blockIdx.x * VB;
  if (threadIdx.x < VB) {
    int i = base + threadIdx.x;
    valid[threadIdx.x] = i < count;
    vbx[threadIdx.x] = i < count ? (i % ncx) * n : 0;
    vby[threadIdx.x] = i < count ? (i / ncx) * n : 0;
  }
  __syncthreads();
  cover_cost_group<N, false>(g, tab, pp, X, Y, B, vbx, vby, valid, cost, rad, err);
}

// An interval [mid - rad, mid + rad] that holds the value the reference's
// f64 arithmetic computes.
struct Ival {
  double mid, rad;
};
__device__ __forceinline__ Ival iv_add(Ival a, Ival b) {
  const double m = a.mid + b.mid;
  const double r = a.rad + b.rad;
  return Ival{m, r + 2.3e-16 * (fabs(m) + r)};
}
__device__ __forceinline__ Ival iv_const(double v) { return Ival{v, 0.0}; }

// The quadtree of tile_region_with over one 32x32 region (4x4 blocks) at
// (bx0, by0).  With rad arrays (certified costs) it returns false, writing
// nothing, when some comparison could go either way within the costs'
// bounds; with rad null (exact costs) it always decides.
__device__ __forceinline__ bool decide_region(Geom g, const ProbeParams* __restrict__ pp, const double* __restrict__ c8,
                              const double* __restrict__ c16, const double* __restrict__ c32,
                              const double* __restrict__ r8, const double* __restrict__ r16,
                              const double* __restrict__ r32, uint8_t* __restrict__ bmap, int bx0, int by0) {
  const int bw = g.bw, bh = g.bh;
  const bool cert = r8 != nullptr;
  uint8_t dec[4][4];
#pragma unroll
  for (int y = 0; y < 4; y++)
#pragma unroll
    for (int x = 0; x < 4; x++) dec[y][x] = 0;
  Ival split4 = iv_const(0.0);
  bool q16[2][2] = {{false, false}, {false, false}};
  const int qo[4][2] = {{0, 0}, {2, 0}, {0, 2}, {2, 2}};
  for (int qi = 0; qi < 4; qi++) {
    const int qx = bx0 + qo[qi][0], qy = by0 + qo[qi][1];
    if (qx >= bw || qy >= bh) continue;  // contributes 0
    Ival split2 = iv_const(0.0);
    const int so[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
    for (int s4 = 0; s4 < 4; s4++) {
      const int sx = qx + so[s4][0], sy = qy + so[s4][1];
      if (sx >= bw || sy >= bh) continue;
      const size_t o = (size_t)sy * bw + sx;
      const Ival c = Ival{c8[o], cert ? r8[o] : 0.0};
      split2 = iv_add(split2, iv_add(iv_add(c, iv_const(fixed_bits(pp, 8))), iv_const(aq_signal_bits(pp, bw, bh, sx, sy, 1))));
    }
    Ival best2 = split2;
    if (qx + 2 <= bw && qy + 2 <= bh) {
      const double fixed = fixed_bits(pp, 16) + aq_signal_bits(pp, bw, bh, qx, qy, 2);
      if (cert && fabs(fixed - split2.mid) <= split2.rad) return false;
      if (!(fixed >= split2.mid)) {
        const size_t o = (size_t)qy * bw + qx;
        const Ival single = iv_add(Ival{c16[o], cert ? r16[o] : 0.0}, iv_const(fixed));
        if (cert && fabs(single.mid - split2.mid) <= single.rad + split2.rad) return false;
        if (single.mid < split2.mid) {
          best2 = single;
          q16[qo[qi][1] / 2][qo[qi][0] / 2] = true;
        }
      }
    }
    split4 = iv_add(split4, best2);
  }
  bool use32 = false;
  if (bx0 + 4 <= bw && by0 + 4 <= bh) {
    const double fixed = fixed_bits(pp, 32) + aq_signal_bits(pp, bw, bh, bx0, by0, 4);
    if (cert && fabs(fixed - split4.mid) <= split4.rad) return false;
    if (!(fixed >= split4.mid)) {
      const size_t o = (size_t)by0 * bw + bx0;
      const Ival single = iv_add(Ival{c32[o], cert ? r32[o] : 0.0}, iv_const(fixed));
      if (cert && fabs(single.mid - split4.mid) <= single.rad + split4.rad) return false;
      if (single.mid < split4.mid) use32 = true;
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
  return true;
}

// One thread per 32x32 region; regions whose decision the certified costs
// cannot settle are queued for the exact pass.
__global__ __launch_bounds__(256) void k_cover_decide(Geom g, const ProbeParams* __restrict__ pp,
                                                      const double* __restrict__ c8,
                                                      const double* __restrict__ c16,
                                                      const double* __restrict__ c32,
                                                      const double* __restrict__ r8,
                                                      const double* __restrict__ r16,
                                                      const double* __restrict__ r32,
                                                      uint8_t* __restrict__ bmap, uint32_t* __restrict__ queue,
                                                      int all_exact) {
  const int rx = blockIdx.x * blockDim.x + threadIdx.x, ry = blockIdx.y;
  const int nrx = (g.bw + 3) / 4;
  if (rx >= nrx) return;
  if (all_exact || !decide_region(g, pp, c8, c16, c32, r8, r16, r32, bmap, 4 * rx, 4 * ry)) {
    const uint32_t slot = atomicAdd(&queue[0], 1u);
    queue[1 + slot] = ((uint32_t)ry << 16) | (uint32_t)rx;
  }
}

// The exact pass over the queued regions: every candidate cost of the
// region with the reference's sequential sums, then its decision.  A fixed
// grid walks the queue (usually empty).
__global__ __launch_bounds__(kThreads) void k_cover_exact(Geom g, const ConstTables* __restrict__ tab,
                                                          const ProbeParams* __restrict__ pp,
                                                          const float* __restrict__ X, const float* __restrict__ Y,
                                                          const float* __restrict__ B, double* __restrict__ c8,
                                                          double* __restrict__ c16, double* __restrict__ c32,
                                                          uint8_t* __restrict__ bmap,
                                                          const uint32_t* __restrict__ queue,
                                                          uint32_t* __restrict__ err) {
  __shared__ int vbx[8], vby[8], valid[8];
  const uint32_t nq = queue[0];
  for (uint32_t qi = blockIdx.x; qi < nq; qi += gridDim.x) {
    const uint32_t e = queue[1 + qi];
    const int bx0 = 4 * (int)(e & 0xffffu), by0 = 4 * (int)(e >> 16);
    // DCT8 candidates: two groups of eight blocks
    for (int half = 0; half < 2; half++) {
      __syncthreads();
      if (threadIdx.x < 8) {
        const int bx = bx0 + threadIdx.x % 4, by = by0 + 2 * half + threadIdx.x / 4;
        valid[threadIdx.x] = bx < g.bw && by < g.bh;
        vbx[threadIdx.x] = min(bx, g.bw - 1);
        vby[threadIdx.x] = min(by, g.bh - 1);
      }
      __syncthreads();
      cover_cost_group<8, true>(g, tab, pp, X, Y, B, vbx, vby, valid, c8, nullptr, err);
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      const int bx = bx0 + 2 * (threadIdx.x % 2), by = by0 + 2 * (threadIdx.x / 2);
      valid[threadIdx.x] = bx + 2 <= g.bw && by + 2 <= g.bh;
      vbx[threadIdx.x] = valid[threadIdx.x] ? bx : 0;
      vby[threadIdx.x] = valid[threadIdx.x] ? by : 0;
    }
    __syncthreads();
    cover_cost_group<16, true>(g, tab, pp, X, Y, B, vbx, vby, valid, c16, nullptr, err);
    __syncthreads();
    if (threadIdx.x < 2) {
      valid[threadIdx.x] = threadIdx.x == 0 && bx0 + 4 <= g.bw && by0 + 4 <= g.bh;
      vbx[threadIdx.x] = valid[threadIdx.x] ? bx0 : 0;
      vby[threadIdx.x] = valid[threadIdx.x] ? by0 : 0;
    }
    __syncthreads();
    cover_cost_group<32, true>(g, tab, pp, X, Y, B, vbx, vby, valid, c32, nullptr, err);
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence_block();
      decide_region(g, pp, c8, c16, c32, nullptr, nullptr, nullptr, bmap, bx0, by0);
    }
  }
}

// ----------------------------------------------------------------------------
// Chroma-from-luma estimate (jpxl-encode-policy/src/lib.rs:3437-4014,
// quantize.rs:1599-1717).  Samples are gathered per varblock into slots the
// host laid out in the reference's order (LF group, tile row, varblock raster,
// cell); the f64 regressions then run sequentially in that order (one thread
// per tile, one for the frame's LF), so every sum is the reference's own, and
// the integer residual-bit prices of the few candidate factors are
// order-free reductions.
// ----------------------------------------------------------------------------

__device__ __forceinline__ float cfl_k(float base, int factor) { return base + (float)factor / 84.0f; }

__device__ int best_factor_dev(double s_yy, double s_yc, double s_cc, float base) {
  if (s_yy <= 0.0) return 0;
  double k_ls = s_yc / s_yy;
  double star = (k_ls - (double)base) * 84.0;
  double r = round(star);
  long long centre = r >= 9.2e18 ? LLONG_MAX : (r <= -9.2e18 ? LLONG_MIN : (long long)r);
  int best = 0;
  double best_e = 0.0;
  bool have = false;
  for (int i = 0; i < 6; i++) {
    int cand;
    if (i < 5) {
      long long c = centre + (i - 2);
      c = c < -128 ? -128 : (c > 127 ? 127 : c);
      cand = (int)c;
    } else {
      cand = 0;
    }
    double k = (double)cfl_k(base, cand);
    double e = s_cc - 2.0 * k * s_yc + k * k * s_yy;
    if (!have || e < best_e || (e == best_e && abs(cand) < abs(best))) {
      have = true;
      best = cand;
      best_e = e;
    }
  }
  return best;
}

__device__ __forceinline__ uint32_t factor_bits(int f) {
  return f == 0 ? 0u : (uint32_t)(32 - __clz((unsigned)abs(f))) + 2u;
}

// seed-1..seed+1 clamped to [-128, 127], deduplicated, plus 0 (lib.rs:3771).
__device__ int factor_candidates(int seed, int* out) {
  int n = 0;
  for (int d = -1; d <= 1; d++) {
    int c = seed + d;
    c = c < -128 ? -128 : (c > 127 ? 127 : c);
    bool dup = false;
    for (int i = 0; i < n; i++) dup |= out[i] == c;
    if (!dup) out[n++] = c;
  }
  bool has0 = false;
  for (int i = 0; i < n; i++) has0 |= out[i] == 0;
  if (!has0) out[n++] = 0;
  return n;
}

template <int N>
__global__ __launch_bounds__(64) void k_cfl_collect(Geom g, const uint32_t* __restrict__ list,
                                                    const StructInfo* __restrict__ si,
                                                    const uint32_t* __restrict__ lf_off_b,
                                                    const uint32_t* __restrict__ hf_local,
                                                    const uint32_t* __restrict__ tile_base,
                                                    const ConstTables* __restrict__ tab,
                                                    const ProbeParams* __restrict__ pp,
                                                    const float* __restrict__ CX, const float* __restrict__ CY,
                                                    const float* __restrict__ CB, float4* __restrict__ lf_s,
                                                    float4* __restrict__ hf_s, uint16_t* __restrict__ hf_cell,
                                                    uint32_t* __restrict__ err) {
  constexpr int NN = N * N, n = N / 8, six = size_index(N), cb = kCellBase[six];
  const int i = blockIdx.x;
  if (i >= si->count[six]) return;
  const uint32_t e = list[i];
  const int bx = (int)(e & 0xffffu), by = (int)(e >> 16);
  const size_t blk = (size_t)by * g.bw + bx;
  const uint32_t lf_off = lf_off_b[blk], hf_off = tile_base[(by / 8) * g.tw + bx / 8] + hf_local[blk];
  int bad = 0;
  const DctTables& t = tab->dct;
  if (threadIdx.x == 0) {
    const float* sf = tab->scale_f[six];
    float lf[3][n * n];
    const float* cp[3] = {CX, CY, CB};
    for (int c = 0; c < 3; c++) {
#pragma unroll
      for (int y = 0; y < n; y++)
#pragma unroll
        for (int x = 0; x < n; x++) {
          float v = cp[c][(size_t)(8 * by + y) * g.cstride + 8 * bx + x];
          lf[c][y * n + x] = n == 1 ? v : v / (sf[y] * sf[x]);
        }
      if constexpr (n == 2) {
        float w2[4] = {lf[c][0], lf[c][1], lf[c][2], lf[c][3]};
        (void)w2;
      }
    }
    // small inverse DCTs (forward.rs:779-830)
    for (int c = 0; c < 3; c++) {
      if constexpr (n > 1) {
        const float f = inv_factor<n>();
        float v[n], s[n * n];
        for (int k = 0; k < n * n; k++) s[k] = lf[c][k];
        for (int r = 0; r < n; r++) {
          for (int k = 0; k < n; k++) v[k] = s[k * n + r];
          dct_iii<n>(v, t);
          for (int k = 0; k < n; k++) s[k * n + r] = v[k] * f;
        }
        for (int r = 0; r < n; r++)
          for (int c2 = 0; c2 < n; c2++) lf[c][r * n + c2] = s[c2 * n + r];
        for (int c2 = 0; c2 < n; c2++) {
          for (int r = 0; r < n; r++) v[r] = lf[c][r * n + c2];
          dct_iii<n>(v, t);
          for (int r = 0; r < n; r++) lf[c][r * n + c2] = v[r] * f;
        }
      }
    }
    const float mdc1 = pp->mdc[1];
    for (int k = 0; k < n * n; k++) {
      float y = lf[1][k];
      int32_t q = clamp_round(y * 1.0f / mdc1, &bad);
      float dy = mdc1 * (float)q / 1.0f;
      lf_s[lf_off + k] = make_float4(lf[0][k], y, lf[2][k], dy);
    }
  }
  float vsc[3];
  varblock_scales(pp, g.bw, g.bh, bx, by, n, vsc, nullptr);  // lib.rs:3582 the varblock's (transform, HfMul)
  const float sc = vsc[1], bias = kBias[1];
  for (int k = threadIdx.x; k < NN; k += 64) {
    const int row = k / N, col = k % N;
    if (row < n && col < n) continue;
    const int slot = row < n ? row * (N - n) + (col - n) : n * (N - n) + (row - n) * N + col;
    const size_t o = (size_t)(8 * by + row) * g.cstride + 8 * bx + col;
    const float y = CY[o];
    const float step = sc * tab->matrix[1][cb + k];
    const int32_t q = choose_nearest(y, step, bias, &bad);
    const float dy = bias_adjust(q, bias) * step;
    constexpr int fold = N / 8;
    const int cell8 = (k / N / fold) * 8 + (k % N / fold);
    hf_s[hf_off + slot] = make_float4(CX[o], y, CB[o], dy);
    hf_cell[hf_off + slot] = (uint16_t)cell8;
  }
  if (bad) atomicOr(err, kErrQuantRange);
}

// Sums of y*y, y*x, x*x, y*b and b*b over samples [b0, b1) in their order,
// each accumulated sequentially in f64 as the reference's CflAccumulator
// does (quantize.rs:1620-1640).  A product of two f32 values is exact in
// f64, so the workgroup first forms a chunk's products in LDS in parallel;
// then five waves each run one of the five sums, one lane per sum, in sample
// order.  The sums are valid in thread 0 after the call; every thread of the
// (kCflThreads-wide) workgroup must call it.
constexpr int kCflThreads = 320;
constexpr int kCflStage = 1024;
// The per-tile kernel's staging: 5 x 1000 doubles (the fallback's chunks and
// the fast path's 320 x 7 double-double tree fit), so that four 320-thread
// workgroups share a CU's LDS.
constexpr int kCflTileStage = 1000;

struct CflSums {
  double yy, ycx, ccx, ycb, ccb;
};

template <int kStage>
__device__ CflSums cfl_sequential_sums(const float4* __restrict__ s, uint32_t b0, uint32_t b1,
                                       double (*prod)[kStage]) {
  __shared__ double total[5];
  constexpr int kPer = (kStage + kCflThreads - 1) / kCflThreads;  // samples per thread per chunk
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double acc = 0.0;
  // Each thread prefetches its share of the next chunk into registers while
  // the five summing lanes work through the current one.
  float4 pre[kPer];
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const uint32_t i = b0 + threadIdx.x + k * kCflThreads;
    pre[k] = i < b1 && threadIdx.x + k * kCflThreads < kStage ? s[i] : make_float4(0, 0, 0, 0);
  }
  for (uint32_t c0 = b0; c0 < b1; c0 += kStage) {
    const uint32_t m = min((uint32_t)kStage, b1 - c0);
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t i = threadIdx.x + k * kCflThreads;
      if (i < m) {
        const double y = (double)pre[k].y, x = (double)pre[k].x, b = (double)pre[k].z;
        prod[0][i] = y * y;
        prod[1][i] = y * x;
        prod[2][i] = x * x;
        prod[3][i] = y * b;
        prod[4][i] = b * b;
      }
    }
    __syncthreads();
    const uint32_t n0 = c0 + kStage;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t i = n0 + threadIdx.x + k * kCflThreads;
      if (n0 < b1 && i < b1 && threadIdx.x + k * kCflThreads < kStage) pre[k] = s[i];
    }
    if (lane == 0 && wave < 5) acc = sequential_sum(prod[wave], (int)m, acc);
    __syncthreads();
  }
  if (lane == 0 && wave < 5) total[wave] = acc;
  __syncthreads();
  return CflSums{total[0], total[1], total[2], total[3], total[4]};
}


// best_factor_dev's seed from the exact sums, or INT_MIN when some sum within
// `gam` (relative to its absolute sum) of the exact one could change it.
__device__ int certified_seed(double syy, double syc, double scc, double ayc, float base, double gam) {
  if (syy <= 0.0) return 0;  // non-negative terms: zero exactly when every term is zero
  const double u = 1.1102230246251565e-16;
  const double dyy = gam * syy, dyc = gam * ayc, dcc = gam * scc;
  if (dyy >= 0.5 * syy) return INT_MIN;
  const double k = syc / syy;
  const double dk = (dyc + fabs(k) * dyy) / (syy - dyy) + 8.0 * u * fabs(k) + 1e-300;
  const double star = (k - (double)base) * 84.0;
  const double dstar = 84.0 * dk + 8.0 * u * (fabs(star) + 84.0 * fabs((double)base)) + 1e-300;
  const double fr = fabs(star) - floor(fabs(star));
  if (fabs(fr - 0.5) <= 2.0 * dstar) return INT_MIN;
  const int seed = best_factor_dev(syy, syc, scc, base);
  // the argmin among the candidates must hold under the perturbation
  const double r = round(star);
  const long long centre = r >= 9.2e18 ? LLONG_MAX : (r <= -9.2e18 ? LLONG_MIN : (long long)r);
  int cand[6];
  for (int i = 0; i < 6; i++) {
    if (i < 5) {
      long long c = centre + (i - 2);
      cand[i] = (int)(c < -128 ? -128 : (c > 127 ? 127 : c));
    } else {
      cand[i] = 0;
    }
  }
  const double kb = (double)cfl_k(base, seed);
  const double eb = scc - 2.0 * kb * syc + kb * kb * syy;
  const double bb = dcc + 2.0 * fabs(kb) * dyc + kb * kb * dyy + 12.0 * u * (scc + 2.0 * fabs(kb) * fabs(syc) + kb * kb * syy);
  for (int i = 0; i < 6; i++) {
    if (cand[i] == seed) continue;
    const double kj = (double)cfl_k(base, cand[i]);
    const double ej = scc - 2.0 * kj * syc + kj * kj * syy;
    const double bj = dcc + 2.0 * fabs(kj) * dyc + kj * kj * dyy + 12.0 * u * (scc + 2.0 * fabs(kj) * fabs(syc) + kj * kj * syy);
    if (!(ej - eb > 2.0 * (bj + bb))) return INT_MIN;
  }
  return seed;
}

// One 64x64 tile's HF factors: sequential regression seeds, then the
// residual prices of the candidate factors against the DCT8 baseline
// quantizer (lib.rs:3899-4014).
__device__ void cfl_tile(int tile, const uint32_t* __restrict__ tile_base,
                         const ConstTables* __restrict__ tab, const ProbeParams* __restrict__ pp,
                         const float4* __restrict__ hf_s, const uint16_t* __restrict__ hf_cell,
                         int8_t* __restrict__ cfl_x, int8_t* __restrict__ cfl_b,
                         double (*stage)[kCflTileStage], const int* __restrict__ gray) {
  __shared__ int cands[2][4];
  __shared__ int ncand[2];
  __shared__ unsigned long long cost[2][4];
  if (*gray) {  // lib.rs:708: no CfL for a grayscale frame
    if (threadIdx.x == 0) cfl_x[tile] = cfl_b[tile] = 0;
    return;
  }
  const uint32_t b0 = tile_base[tile], b1 = tile_base[tile + 1];
  // Certified fast path: the exact sums (double-double, any order) decide the
  // seeds when every sum within the recursive-summation bound of them gives
  // the same seeds as the sequential f64 sums would; otherwise the sequential
  // sums run.
  __shared__ int seeds[3];
  {
    DD acc[7];
#pragma unroll
    for (int k = 0; k < 7; k++) acc[k] = DD{0.0, 0.0};
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
      const float4 v = hf_s[i];
      const double y = (double)v.y, x = (double)v.x, b = (double)v.z;
      const double t[7] = {y * y, y * x, x * x, y * b, b * b, fabs(y * x), fabs(y * b)};  // exact products
#pragma unroll
      for (int k = 0; k < 7; k++) acc[k] = dd_add(acc[k], dd_of(t[k]));
    }
    static_assert(5 * kCflTileStage * sizeof(double) >= kCflThreads * 7 * sizeof(DD), "tree fits the staging");
    DD (*red)[7] = reinterpret_cast<DD (*)[7]>(&stage[0][0]);  // 320 x 7 DD fit the staging
#pragma unroll
    for (int k = 0; k < 7; k++) red[threadIdx.x][k] = acc[k];
    __syncthreads();
    if (threadIdx.x < (unsigned)(kCflThreads - 256))
#pragma unroll
      for (int k = 0; k < 7; k++) red[threadIdx.x][k] = dd_add(red[threadIdx.x][k], red[threadIdx.x + 256][k]);
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
      if ((int)threadIdx.x < stride)
#pragma unroll
        for (int k = 0; k < 7; k++) red[threadIdx.x][k] = dd_add(red[threadIdx.x][k], red[threadIdx.x + stride][k]);
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      double S[7];
      for (int k = 0; k < 7; k++) S[k] = red[0][k].hi + red[0][k].lo;
      const double n = (double)(b1 - b0) + 4.0, u = 1.1102230246251565e-16;
      const double gam = 2.0 * (n * u / (1.0 - n * u)) + 1e-28;
      const int sx = b1 > b0 ? certified_seed(S[0], S[1], S[2], S[5], 0.0f, gam) : 0;
      const int sb = b1 > b0 ? certified_seed(S[0], S[3], S[4], S[6], 1.0f, gam) : 0;
      seeds[0] = sx;
      seeds[1] = sb;
      seeds[2] = sx != INT_MIN && sb != INT_MIN;
    }
    __syncthreads();
  }
  if (!seeds[2]) {
    const CflSums a = cfl_sequential_sums<kCflTileStage>(hf_s, b0, b1, stage);
    if (threadIdx.x == 0) {
      // both channels share s_yy (same sample sequence)
      seeds[0] = b1 > b0 ? best_factor_dev(a.yy, a.ycx, a.ccx, 0.0f) : 0;
      seeds[1] = b1 > b0 ? best_factor_dev(a.yy, a.ycb, a.ccb, 1.0f) : 0;
    }
  }
  if (threadIdx.x == 0) {
    ncand[0] = seeds[0] == 0 ? 0 : factor_candidates(seeds[0], cands[0]);
    ncand[1] = seeds[1] == 0 ? 0 : factor_candidates(seeds[1], cands[1]);
    for (int ch = 0; ch < 2; ch++)
      for (int i = 0; i < 4; i++) cost[ch][i] = 0;
  }
  __syncthreads();
  const float sx = pp->sc[0], sb = pp->sc[2];
  unsigned long long local[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
  int dummy = 0;
  for (uint32_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    const float4 s = hf_s[i];
    const int cell = hf_cell[i];
    for (int ch = 0; ch < 2; ch++) {
      const int c = ch == 0 ? 0 : 2;
      const float base = ch == 0 ? 0.0f : 1.0f;
      const float src = ch == 0 ? s.x : s.z;
      const float step = (ch == 0 ? sx : sb) * tab->matrix[c][cell];
      for (int j = 0; j < ncand[ch]; j++) {
        const float t = src - cfl_k(base, cands[ch][j]) * s.w;
        local[ch][j] += residual_bits(choose_nearest(t, step, kBias[c], &dummy));
      }
    }
  }
  for (int ch = 0; ch < 2; ch++)
    for (int j = 0; j < 4; j++)
      if (local[ch][j]) atomicAdd(&cost[ch][j], local[ch][j]);
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int ch = 0; ch < 2; ch++) {
      int best = 0;
      if (ncand[ch] > 0) {
        unsigned long long best_cost = 0;
        for (int j = 0; j < ncand[ch]; j++)
          if (cands[ch][j] == 0) best_cost = cost[ch][j];
        for (int j = 0; j < ncand[ch]; j++) {
          const int f = cands[ch][j];
          if (f == 0) continue;
          unsigned long long cst = cost[ch][j] + factor_bits(f);
          if (cst < best_cost || (cst == best_cost && abs(f) < abs(best))) {
            best_cost = cst;
            best = f;
          }
        }
      }
      (ch == 0 ? cfl_x : cfl_b)[tile] = (int8_t)best;
    }
  }
}

// The frame-wide LF factors (lib.rs:3831-3894 refine_lf_factors), in three
// launches: the regression seeds (one workgroup, sequential sums), the
// residual prices of every candidate (an order-free integer reduction over
// the whole grid), and the joint selection.
// --- certified fast path of the frame-LF regression ------------------------
// The reference sums ~(blocks) products sequentially in f64.  The exact sums
// (double-double, any order) differ from the sequential ones by at most
// gamma_(n-1) * sum|t| (the recursive-summation bound, gamma_k = k u/(1-k u));
// when the seeds best_factor derives are provably the same for every sum
// within that bound of the exact ones, they are the sequential path's seeds
// and the long chain is skipped.  Otherwise the chain runs.

__global__ __launch_bounds__(256) void k_cfl_lf_fast_partial(const StructInfo* __restrict__ si, const float4* __restrict__ lf_s,
                                                             CflLfScratch* __restrict__ sc) {
  const uint32_t count = si->lf_total;
  __shared__ DD red[256][7];
  DD acc[7];
#pragma unroll
  for (int k = 0; k < 7; k++) acc[k] = DD{0.0, 0.0};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const float4 v = lf_s[i];
    const double y = (double)v.y, x = (double)v.x, b = (double)v.z;
    const double t[7] = {y * y, y * x, x * x, y * b, b * b, fabs(y * x), fabs(y * b)};  // exact products
#pragma unroll
    for (int k = 0; k < 7; k++) acc[k] = dd_add(acc[k], dd_of(t[k]));
  }
#pragma unroll
  for (int k = 0; k < 7; k++) red[threadIdx.x][k] = acc[k];
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if ((int)threadIdx.x < stride)
#pragma unroll
      for (int k = 0; k < 7; k++) red[threadIdx.x][k] = dd_add(red[threadIdx.x][k], red[threadIdx.x + stride][k]);
    __syncthreads();
  }
  if (threadIdx.x < 7) sc->part[blockIdx.x][threadIdx.x] = make_double2(red[0][threadIdx.x].hi, red[0][threadIdx.x].lo);
}


__global__ __launch_bounds__(256) void k_cfl_lf_fast_final(const StructInfo* __restrict__ si, int groups,
                                                            CflLfScratch* __restrict__ sc) {
  const uint32_t count = si->lf_total;
  // the per-group double-double partials, reduced by a tree (any order: the
  // bound below covers the double-double's own error)
  __shared__ DD red[256][7];
  for (int k = 0; k < 7; k++)
    red[threadIdx.x][k] = (int)threadIdx.x < groups ? DD{sc->part[threadIdx.x][k].x, sc->part[threadIdx.x][k].y}
                                                    : DD{0.0, 0.0};
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if ((int)threadIdx.x < stride)
      for (int k = 0; k < 7; k++) red[threadIdx.x][k] = dd_add(red[threadIdx.x][k], red[threadIdx.x + stride][k]);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  DD tot[7];
  for (int k = 0; k < 7; k++) tot[k] = red[0][k];
  double S[7];
  for (int k = 0; k < 7; k++) S[k] = tot[k].hi + tot[k].lo;
  // recursive-summation bound for n terms, with room for the double-double
  // and final rounding
  const double n = (double)count + 4.0;
  const double u = 1.1102230246251565e-16;
  const double gam = 2.0 * (n * u / (1.0 - n * u)) + 1e-28;
  const int sx = certified_seed(S[0], S[1], S[2], S[5], 0.0f, gam);
  const int sb = certified_seed(S[0], S[3], S[4], S[6], 1.0f, gam);
  const int ok = sx != INT_MIN && sb != INT_MIN;
  if (ok) {
    sc->ncand[0] = factor_candidates(sx, sc->cands[0]);
    sc->ncand[1] = factor_candidates(sb, sc->cands[1]);
    for (int ch = 0; ch < 2; ch++)
      for (int j = 0; j < 4; j++) sc->cost[ch][j] = 0;
  }
  sc->certified = ok;
}

__global__ __launch_bounds__(kCflThreads) void k_cfl_lf_seed(const StructInfo* __restrict__ si, const float4* __restrict__ lf_s,
                                                             CflLfScratch* __restrict__ sc) {
  const uint32_t count = si->lf_total;
  __shared__ double stage[5][kCflStage];
  if (sc->certified) return;  // the fast path's seeds stand
  const CflSums a = cfl_sequential_sums<kCflStage>(lf_s, 0, count, stage);
  if (threadIdx.x == 0) {
    sc->ncand[0] = factor_candidates(best_factor_dev(a.yy, a.ycx, a.ccx, 0.0f), sc->cands[0]);
    sc->ncand[1] = factor_candidates(best_factor_dev(a.yy, a.ycb, a.ccb, 1.0f), sc->cands[1]);
    for (int ch = 0; ch < 2; ch++)
      for (int j = 0; j < 4; j++) sc->cost[ch][j] = 0;
  }
}

__global__ __launch_bounds__(256) void k_cfl_lf_price(const StructInfo* __restrict__ si, const ProbeParams* __restrict__ pp,
                                                      const float4* __restrict__ lf_s,
                                                      CflLfScratch* __restrict__ sc) {
  const uint32_t count = si->lf_total;
  __shared__ unsigned long long cost[2][4];
  __shared__ int cands[2][4], ncand[2];
  if (threadIdx.x < 8) {
    cost[threadIdx.x / 4][threadIdx.x % 4] = 0;
    cands[threadIdx.x / 4][threadIdx.x % 4] = sc->cands[threadIdx.x / 4][threadIdx.x % 4];
  }
  if (threadIdx.x < 2) ncand[threadIdx.x] = sc->ncand[threadIdx.x];
  __syncthreads();
  const float mdc[2] = {pp->mdc[0], pp->mdc[2]};
  unsigned long long local[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
  int dummy = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const float4 s = lf_s[i];
    for (int ch = 0; ch < 2; ch++) {
      const float base = ch == 0 ? 0.0f : 1.0f, src = ch == 0 ? s.x : s.z;
      for (int j = 0; j < ncand[ch]; j++) {
        const float t = src - cfl_k(base, cands[ch][j]) * s.w;
        local[ch][j] += residual_bits(clamp_round(t * 1.0f / mdc[ch], &dummy));
      }
    }
  }
  for (int ch = 0; ch < 2; ch++)
    for (int j = 0; j < 4; j++)
      if (local[ch][j]) atomicAdd(&cost[ch][j], local[ch][j]);
  __syncthreads();
  if (threadIdx.x < 8 && cost[threadIdx.x / 4][threadIdx.x % 4])
    atomicAdd(&sc->cost[threadIdx.x / 4][threadIdx.x % 4], cost[threadIdx.x / 4][threadIdx.x % 4]);
}

__global__ void k_cfl_lf_select(const CflLfScratch* __restrict__ sc, CflParams* __restrict__ out,
                                const int* __restrict__ gray) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (*gray) {
    out->x_factor_lf = out->b_factor_lf = 0;
    return;
  }
  unsigned long long nx = 0, nb = 0;
  for (int j = 0; j < sc->ncand[0]; j++) if (sc->cands[0][j] == 0) nx = sc->cost[0][j];
  for (int j = 0; j < sc->ncand[1]; j++) if (sc->cands[1][j] == 0) nb = sc->cost[1][j];
  int bx = 0, bb = 0;
  unsigned long long best = nx + nb;
  for (int i = 0; i < sc->ncand[0]; i++)
    for (int j = 0; j < sc->ncand[1]; j++) {
      const int fx = sc->cands[0][i], fb = sc->cands[1][j];
      unsigned long long c = sc->cost[0][i] + sc->cost[1][j];
      if (fx != 0 || fb != 0) c += 50ull + factor_bits(fx) + factor_bits(fb);
      const unsigned mag = (unsigned)abs(fx) + (unsigned)abs(fb);
      const unsigned bmag = (unsigned)abs(bx) + (unsigned)abs(bb);
      if (c < best || (c == best && mag < bmag)) {
        bx = fx;
        bb = fb;
        best = c;
      }
    }
  out->x_factor_lf = bx;
  out->b_factor_lf = bb;
}

__global__ __launch_bounds__(kCflThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_cfl_tiles(const uint32_t* __restrict__ tile_base, int tiles,
                                                           const ConstTables* __restrict__ tab,
                                                           const ProbeParams* __restrict__ pp,
                                                           const float4* __restrict__ hf_s,
                                                           const uint16_t* __restrict__ hf_cell,
                                                           int8_t* __restrict__ cfl_x, int8_t* __restrict__ cfl_b,
                                                           const int* __restrict__ gray) {
  __shared__ double stage[5][kCflTileStage];
  if ((int)blockIdx.x < tiles)
    cfl_tile((int)blockIdx.x, tile_base, tab, pp, hf_s, hf_cell, cfl_x, cfl_b, stage, gray);
}

// ----------------------------------------------------------------------------
// Adaptive quantization field (jpxl-encode-policy/src/analysis.rs:111-158,
// field.rs:200-320).  One thread per 8x8 atom for the atlas and the field;
// the frame mean of the activities is the reference's sequential f64 sum,
// reproduced by a certified double-double sum (the narrowed f32 mean is the
// same for every sum within the recursive-summation bound of the exact
// one) with the sequential chain as the fallback.
// ----------------------------------------------------------------------------

// The activity's f32 log2: log2 in f64, one rounding to f32.
__device__ __forceinline__ float aq_log2f(float x) { return (float)log2((double)x); }

__global__ __launch_bounds__(256) void k_aq_activity(Geom g, const float* __restrict__ X,
                                                     const float* __restrict__ Y, const float* __restrict__ B,
                                                     float* __restrict__ act) {
  const int atom = blockIdx.x * blockDim.x + threadIdx.x;
  if (atom >= g.bw * g.bh) return;
  const int ax = atom % g.bw, ay = atom / g.bw;
  const int x0 = ax * 8, y0 = ay * 8, x1 = min(x0 + 8, g.w), y1 = min(y0 + 8, g.h);
  const double count = (double)(x1 - x0) * (double)(y1 - y0);
  const float* planes[3] = {X, Y, B};
  float var[3];
  for (int c = 0; c < 3; c++) {
    double sum = 0.0, sum_sq = 0.0;
    for (int y = y0; y < y1; y++)
      for (int x = x0; x < x1; x++) {
        const double v = (double)planes[c][(size_t)y * g.w + x];
        sum += v;
        sum_sq += v * v;
      }
    const double mean = sum / count;
    double variance = sum_sq / count - mean * mean;
    if (!(variance > 0.0)) variance = 0.0;
    var[c] = (float)variance;
  }
  const float k8 = 255.0f * 255.0f;
  const float luma = aq_log2f(1.0f + var[1] * k8);
  const float chroma = aq_log2f(1.0f + (var[0] + var[2]) * k8);
  act[atom] = luma + 0.35f * chroma;
}

// Exact order statistics of the per-atom variances (non-negative f32, so
// their bit patterns sort as the values do) by radix selection over three
// digits (bits 31-21, 20-10, 9-0): each pass counts the digit of the
// elements that share the prefix found so far, in LDS per workgroup (hot
// bins stay on-chip) and then once per non-empty bin in global memory; the
// select step finds the digit holding each query's rank.  Integer counts:
// the same answer in any order.
constexpr int kSelectBlocks = 128;

__global__ void k_zero_u32(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

template <int PASS>
__global__ __launch_bounds__(256) void k_select_hist(const float* __restrict__ luma, const float* __restrict__ chroma,
                                                     int n, float flat_below, SelectRanks r,
                                                     SelectScratch* __restrict__ s) {
  constexpr int kShift = PASS == 0 ? 21 : (PASS == 1 ? 10 : 0);
  constexpr int kBins = PASS == 2 ? 1024 : 2048;
  constexpr int kSets = PASS == 0 ? 2 : kSelectQueries;
  __shared__ uint32_t hist[kSets][kBins];
  __shared__ uint32_t flat;
  for (int i = threadIdx.x; i < kSets * kBins; i += 256) (&hist[0][0])[i] = 0u;
  if (threadIdx.x == 0) flat = 0u;
  uint32_t prefix[kSelectQueries];
  if constexpr (PASS > 0)
    for (int q = 0; q < kSelectQueries; q++) prefix[q] = s->prefix[q];
  __syncthreads();
  uint32_t my_flat = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kSelectBlocks * 256) {
    const uint32_t ul = __float_as_uint(luma[i]), uc = __float_as_uint(chroma[i]);
    if constexpr (PASS == 0) {
      atomicAdd(&hist[0][ul >> kShift], 1u);
      atomicAdd(&hist[1][uc >> kShift], 1u);
      my_flat += luma[i] < flat_below ? 1u : 0u;
    } else {
#pragma unroll
      for (int q = 0; q < kSelectQueries; q++) {
        const uint32_t u = r.set[q] == 0 ? ul : uc;
        if ((u >> (kShift + (PASS == 1 ? 11 : 10))) == prefix[q]) atomicAdd(&hist[q][(u >> kShift) & (kBins - 1)], 1u);
      }
    }
  }
  if constexpr (PASS == 0)
    if (my_flat) atomicAdd(&flat, my_flat);
  __syncthreads();
  uint32_t* out = PASS == 0 ? &s->h1[0][0] : (PASS == 1 ? &s->h2[0][0] : &s->h3[0][0]);
  for (int i = threadIdx.x; i < kSets * kBins; i += 256) {
    const uint32_t c = (&hist[0][0])[i];
    if (c) atomicAdd(&out[i], c);
  }
  if (PASS == 0 && threadIdx.x == 0 && flat) atomicAdd(&s->flat, flat);
}

// One workgroup per query: the digit whose cumulative count first exceeds
// the query's remaining rank; the prefix gains the digit, the rank loses the
// counts below it.  On the last pass the prefix is the value.
template <int PASS>
__global__ __launch_bounds__(256) void k_select_digit(SelectRanks r, SelectScratch* __restrict__ s) {
  constexpr int kBins = PASS == 2 ? 1024 : 2048, kPer = kBins / 256;
  constexpr int kDigitBits = PASS == 2 ? 10 : 11;
  __shared__ uint32_t part[256];
  const int q = blockIdx.x, t = threadIdx.x;
  const uint32_t* hist = PASS == 0 ? s->h1[r.set[q]] : (PASS == 1 ? s->h2[q] : s->h3[q]);
  const uint32_t rank = PASS == 0 ? r.rank[q] : s->rank[q];
  uint32_t sum = 0;
  for (int b = 0; b < kPer; b++) sum += hist[t * kPer + b];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const uint32_t v = t >= o ? part[t - o] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  const uint32_t before = part[t] - sum;
  if (rank >= before && rank < part[t]) {
    uint32_t cum = before;
    for (int b = 0; b < kPer; b++) {
      const uint32_t c = hist[t * kPer + b];
      if (rank < cum + c) {
        const uint32_t digit = (uint32_t)(t * kPer + b);
        const uint32_t prefix = (PASS == 0 ? 0u : s->prefix[q] << kDigitBits) | digit;
        s->prefix[q] = prefix;
        s->rank[q] = rank - cum;
        if (PASS == 2) s->value[q] = __uint_as_float(prefix);
        break;
      }
      cum += c;
    }
  }
}

// preanalysis_features (preanalysis.rs:277-484): one workgroup per whole
// 32x32 cell.  The samples' top eight bits give integer BT.601 planes; the
// 3x3 Sobel reads luma clamped to the frame; the cell's integer sums are
// exact in any order, so waves reduce them freely.
template <bool U8>
__global__ __launch_bounds__(256) void k_preanalysis_cells(int w, int h, int shift, const void* __restrict__ rgb,
                                                           int cells_x, CellStats* __restrict__ out) {
  __shared__ int lum[34][35];
  __shared__ int red[4][14];
  const int x0 = (int)(blockIdx.x % cells_x) * 32, y0 = (int)(blockIdx.x / cells_x) * 32;
  auto sample = [&](int x, int y, int ch) -> int {
    const size_t i = ((size_t)y * w + x) * 3 + ch;
    if constexpr (U8) return (int)((const uint8_t*)rgb)[i];
    return min((int)(((const uint16_t*)rgb)[i] >> shift), 255);
  };
  // the tile's samples (34 x 34 = 1156 < 5 x 256) and the lane's four
  // interior pixels' are requested before any is used
  int tr[5][3], px3[4][3];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const int e = min((int)threadIdx.x + 256 * k, 34 * 34 - 1);
    const int ly = e / 34, lx = e % 34;
    const int x = min(max(x0 - 1 + lx, 0), w - 1), y = min(max(y0 - 1 + ly, 0), h - 1);
#pragma unroll
    for (int ch = 0; ch < 3; ch++) tr[k][ch] = sample(x, y, ch);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int p = (int)threadIdx.x + 256 * k;
#pragma unroll
    for (int ch = 0; ch < 3; ch++) px3[k][ch] = sample(x0 + (p & 31), y0 + (p >> 5), ch);
  }
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const int e = (int)threadIdx.x + 256 * k;
    if (e < 34 * 34) lum[e / 34][e % 34] = (77 * tr[k][0] + 150 * tr[k][1] + 29 * tr[k][2] + 128) >> 8;
  }
  __syncthreads();
  // sum, sum_sq, edge, flat, noise_sum, weak, dir[4], cb_sum, cb_sq, cr_sum, cr_sq
  int a[14] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int p = (int)threadIdx.x + 256 * k, px = p & 31, py = p >> 5;
    const int lx = px + 1, ly = py + 1;
    const int c = lum[ly][lx];
    const int tl = lum[ly - 1][lx - 1], tc = lum[ly - 1][lx], tr = lum[ly - 1][lx + 1];
    const int ml = lum[ly][lx - 1], mr = lum[ly][lx + 1];
    const int bl = lum[ly + 1][lx - 1], bc = lum[ly + 1][lx], br = lum[ly + 1][lx + 1];
    a[0] += c;
    a[1] += c * c;
    const int gx = (tr + 2 * mr + br) - (tl + 2 * ml + bl);
    const int gy = (bl + 2 * bc + br) - (tl + 2 * tc + tr);
    const int ax = abs(gx), ay = abs(gy), grad = ax + ay;
    if (grad > 48) {
      a[2]++;
      const int bin = ax >= 2 * ay ? 0 : (ay >= 2 * ax ? 1 : ((gx > 0) == (gy > 0) ? 2 : 3));
      a[6 + bin]++;
    }
    if (grad < 12) a[3]++;
    if (grad < 24) {
      const int box = (tl + tc + tr + ml + c + mr + bl + bc + br) / 9;
      a[4] += abs(c - box);
      a[5]++;
    }
    const int r = px3[k][0], gg = px3[k][1], b = px3[k][2];
    const int cb = min(max(((-43 * r - 85 * gg + 128 * b + 128) >> 8) + 128, 0), 255);
    const int cr = min(max(((128 * r - 107 * gg - 21 * b + 128) >> 8) + 128, 0), 255);
    a[10] += cb;
    a[11] += cb * cb;
    a[12] += cr;
    a[13] += cr * cr;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < 14; i++) {
    int v = a[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wv][i] = v;
  }
  __syncthreads();
  if (threadIdx.x < 14) {
    const int i = threadIdx.x;
    const int v = red[0][i] + red[1][i] + red[2][i] + red[3][i];
    int* o = (int*)&out[blockIdx.x];
    o[i] = v;
  }
}

// field.rs:200-219: the minimum over the clamped 3x3 neighbourhood.
__global__ __launch_bounds__(256) void k_aq_erode(int bw, int bh, const float* __restrict__ a,
                                                  float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= bw * bh) return;
  const int x = i % bw, y = i / bw;
  float m = INFINITY;
  for (int ny = max(y - 1, 0); ny < min(y + 2, bh); ny++)
    for (int nx = max(x - 1, 0); nx < min(x + 2, bw); nx++) m = fminf(m, a[ny * bw + nx]);
  out[i] = isfinite(m) ? m : 0.0f;
}

// field.rs:286-295: mean = (sum_i a_i in f64, raster order) / count, as f32.
// The activities are non-negative, so the sequential sum lies within
// gamma_(n-1) * S of the exact sum S.
__global__ __launch_bounds__(1024) void k_aq_mean(const float* __restrict__ a, int n, AqScratch* __restrict__ sc) {
  __shared__ DD red[1024];
  DD acc{0.0, 0.0};
  for (int i = threadIdx.x; i < n; i += 1024) acc = dd_add(acc, dd_of((double)a[i]));
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int stride = 512; stride > 0; stride >>= 1) {
    if ((int)threadIdx.x < stride) red[threadIdx.x] = dd_add(red[threadIdx.x], red[threadIdx.x + stride]);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double S = red[0].hi + red[0].lo;
  const double u = 1.1102230246251565e-16, k = (double)n + 4.0;
  const double E = 2.0 * (k * u / (1.0 - k * u)) * S + 1e-300;
  const double lo = S - E, hi = S + E;
  const float m_lo = (float)(lo / (double)n), m_hi = (float)(hi / (double)n);
  if (m_lo == m_hi) {
    sc->mean = m_lo;
    sc->certified = 1;
    return;
  }
  double sum = 0.0;
  for (int i = 0; i < n; i++) sum += (double)a[i];
  sc->mean = (float)(sum / (double)n);
  sc->certified = 0;
}

// field.rs:262-300 the per-atom adjustment, and field.rs:312-320 whether any
// atom leaves the baseline (the field is then not neutral).
__global__ __launch_bounds__(256) void k_aq_adj(int count, int mode, const float* __restrict__ act,
                                                const float* __restrict__ ero, AqScratch* __restrict__ sc,
                                                float* __restrict__ adj) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const bool uniform = mode == kAqUniform || mode == kAqFineUniform || mode == kAqEdgeRefine;
  const bool fine = mode >= kAqFineMasking;
  const float k = (uniform ? -1.0f : 1.0f) * 0.25f;
  float a;
  if (mode == kAqEdgeRefine) {
    const float d = act[i] - ero[i] - 6.0f;
    a = k * (d > 0.0f ? d : 0.0f);
  } else {
    const float v = k * ((fine ? ero[i] : act[i]) - sc->mean);
    a = v < -1.0f ? -1.0f : (v > 1.0f ? 1.0f : v);
  }
  adj[i] = a;
  const bool moved = fine ? aq_mul_of(16u, a) != 16u : roundf(a * 2.0f) != 0.0f;
  if (moved) atomicOr(&sc->active, 1);
}

// The HfMul of every varblock of the lists, at its top-left block.
template <int N>
__global__ __launch_bounds__(256) void k_aq_block_muls(Geom g, const ProbeParams* __restrict__ pp,
                                                       const uint32_t* __restrict__ list, int count,
                                                       uint32_t* __restrict__ muls) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int bx = (int)(list[i] & 0xffffu), by = (int)(list[i] >> 16);
  muls[(size_t)by * g.bw + bx] =
      (pp->aq_adj || pp->vb_mul) ? aq_footprint_mul(pp, g.bw, g.bh, bx, by, N / 8) : pp->aq_baseline;
}

// ----------------------------------------------------------------------------
// The structure's work lists and CfL sample slots, built on the device from
// the block map.  Varblocks are visited in the reference's CfL order: LF
// groups of 256x256 blocks in raster order, blocks in raster order inside a
// group (lib.rs:3437-3600).  Each size's list takes its varblocks in that
// order; a varblock's LF samples start at the running sum of n*n; its HF
// samples start at its tile's base plus the running sum of side^2 - n^2 over
// the tile's earlier varblocks (tiles are 8x8 blocks, inside one group).
// ----------------------------------------------------------------------------

// Exclusive scan of four counters over a 1024-thread workgroup.
__device__ __forceinline__ void scan4_1024(uint32_t v[4], uint32_t tot[4]) {
  __shared__ uint32_t part[16][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint32_t x = v[k];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    inc[k] = x;
  }
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < 4; k++) part[wave][k] = inc[k];
  __syncthreads();
  if (threadIdx.x < 4) {
    uint32_t run = 0;
    for (int w = 0; w < 16; w++) {
      const uint32_t t = part[w][threadIdx.x];
      part[w][threadIdx.x] = run;
      run += t;
    }
    tot[threadIdx.x] = run;  // tot is shared: read after the next barrier
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = part[wave][k] + inc[k] - v[k];
  __syncthreads();
}

// The varblock lists over the whole chip: a chunk is one block row of one
// 256x256-block LF group (chunk id = group * 256 + row, groups in raster
// order; ids past a partial group's rows are empty), so the reference's
// list order (groups, then rows, then columns) is chunk order, then column.
// k_rows_count counts each chunk's varblocks per size and LF samples,
// k_rows_scan turns the counts into each chunk's first list slots, and
// k_rows_write places every varblock with a workgroup scan of its row.
__device__ __forceinline__ void scan4_256(uint32_t v[4], uint32_t tot[4]) {
  __shared__ uint32_t part[4][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint32_t x = v[k];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    inc[k] = x;
  }
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < 4; k++) part[wave][k] = inc[k];
  __syncthreads();
  if (threadIdx.x < 4) {
    uint32_t run = 0;
    for (int w = 0; w < 4; w++) {
      const uint32_t t = part[w][threadIdx.x];
      part[w][threadIdx.x] = run;
      run += t;
    }
    tot[threadIdx.x] = run;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = part[wave][k] + inc[k] - v[k];  // exclusive
  __syncthreads();
}

// This thread's block of chunk blockIdx.x: side 0 when outside the frame.
__device__ __forceinline__ int chunk_block(const Geom& g, const uint8_t* __restrict__ bmap, int* bx, int* by) {
  const int gw = (g.bw + 255) / 256;
  const int grp = blockIdx.x / 256, dy = blockIdx.x % 256;
  *by = (grp / gw) * 256 + dy;
  *bx = (grp % gw) * 256 + (int)threadIdx.x;
  if (*by >= g.bh || *bx >= g.bw) return 0;
  return bmap[(size_t)*by * g.bw + *bx];
}

__global__ __launch_bounds__(256) void k_rows_count(Geom g, const uint8_t* __restrict__ bmap, uint4* __restrict__ cnt) {
  __shared__ uint32_t tot[4];
  int bx, by;
  const int side = chunk_block(g, bmap, &bx, &by);
  uint32_t v[4] = {0, 0, 0, 0};
  if (side) {
    v[size_index(side)] = 1;
    v[3] = (uint32_t)(side / 8) * (side / 8);
  }
  scan4_256(v, tot);
  if (threadIdx.x == 0) cnt[blockIdx.x] = make_uint4(tot[0], tot[1], tot[2], tot[3]);
}

__global__ __launch_bounds__(1024) void k_rows_scan(int nch, const uint4* __restrict__ cnt, uint4* __restrict__ base,
                                                    StructInfo* __restrict__ si) {
  __shared__ uint32_t tot[4];
  const int run = (nch + 1023) / 1024, j0 = min(nch, (int)threadIdx.x * run), j1 = min(nch, j0 + run);
  uint32_t v[4] = {0, 0, 0, 0};
  for (int j = j0; j < j1; j++) {
    const uint4 c = cnt[j];
    v[0] += c.x, v[1] += c.y, v[2] += c.z, v[3] += c.w;
  }
  scan4_1024(v, tot);
  for (int j = j0; j < j1; j++) {
    base[j] = make_uint4(v[0], v[1], v[2], v[3]);
    const uint4 c = cnt[j];
    v[0] += c.x, v[1] += c.y, v[2] += c.z, v[3] += c.w;
  }
  if (threadIdx.x == 0) {
    si->count[0] = (int)tot[0];
    si->count[1] = (int)tot[1];
    si->count[2] = (int)tot[2];
    si->lf_total = tot[3];
  }
}

__global__ __launch_bounds__(256) void k_rows_write(Geom g, const uint8_t* __restrict__ bmap,
                                                    const uint4* __restrict__ base, uint32_t* __restrict__ l8,
                                                    uint32_t* __restrict__ l16, uint32_t* __restrict__ l32,
                                                    uint32_t* __restrict__ lf_off) {
  __shared__ uint32_t tot[4];
  int bx, by;
  const int side = chunk_block(g, bmap, &bx, &by);
  uint32_t v[4] = {0, 0, 0, 0};
  if (side) {
    v[size_index(side)] = 1;
    v[3] = (uint32_t)(side / 8) * (side / 8);
  }
  scan4_256(v, tot);
  if (!side) return;
  const uint4 b = base[blockIdx.x];
  const int k = size_index(side);
  uint32_t* list = k == 0 ? l8 : (k == 1 ? l16 : l32);
  const uint32_t slot = (k == 0 ? b.x : (k == 1 ? b.y : b.z)) + v[k];
  list[slot] = ((uint32_t)by << 16) | (uint32_t)bx;
  lf_off[(size_t)by * g.bw + bx] = b.w + v[3];
}

// One 64-thread workgroup per 8x8-block tile: each block's HF sample offset
// inside its tile (raster order) and the tile's total.
__global__ __launch_bounds__(64) void k_struct_tiles(Geom g, const uint8_t* __restrict__ bmap,
                                                     uint32_t* __restrict__ hf_local,
                                                     uint32_t* __restrict__ tile_count) {
  const int t = blockIdx.x, tx = t % g.tw, ty = t / g.tw;
  const int bx = tx * 8 + (threadIdx.x & 7), by = ty * 8 + (threadIdx.x >> 3);
  const bool in = bx < g.bw && by < g.bh;
  const int side = in ? bmap[(size_t)by * g.bw + bx] : 0;
  const uint32_t cells = side ? (uint32_t)(side * side - (side / 8) * (side / 8)) : 0u;
  uint32_t x = cells;
  const int lane = threadIdx.x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (in) hf_local[(size_t)by * g.bw + bx] = x - cells;
  if (lane == 63) tile_count[t] = x;
}

// tile_base = exclusive prefix of the tile totals (tiles + 1 entries).
__global__ __launch_bounds__(1024) void k_struct_tile_base(int tiles, const uint32_t* __restrict__ tile_count,
                                                           uint32_t* __restrict__ tile_base,
                                                           StructInfo* __restrict__ si) {
  __shared__ uint32_t tot[4];
  const int run = (tiles + 1023) / 1024, j0 = min(tiles, (int)threadIdx.x * run), j1 = min(tiles, j0 + run);
  uint32_t v[4] = {0, 0, 0, 0};
  for (int j = j0; j < j1; j++) v[0] += tile_count[j];
  scan4_1024(v, tot);
  uint32_t acc = v[0];
  for (int j = j0; j < j1; j++) {
    tile_base[j] = acc;
    acc += tile_count[j];
  }
  if (threadIdx.x == 0) {
    tile_base[tiles] = tot[0];
    si->hf_total = tot[0];
  }
}

// The structure's readback in one contiguous block: counts, LF factors, the
// error word, the tile factors and the block map (layout of
// structure_readback_bytes).
__global__ __launch_bounds__(256) void k_gather_structure(const StructInfo* __restrict__ si,
                                                          const CflParams* __restrict__ cp,
                                                          const uint32_t* __restrict__ err,
                                                          const int8_t* __restrict__ cx,
                                                          const int8_t* __restrict__ cb,
                                                          const uint8_t* __restrict__ bmap, int tiles, int bpx,
                                                          char* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, step = gridDim.x * blockDim.x;
  if (i == 0) {
    *(StructInfo*)out = *si;
    *(CflParams*)(out + 32) = *cp;
    *(uint32_t*)(out + 48) = *err;
  }
  for (int k = i; k < tiles; k += step) {
    out[64 + k] = (char)cx[k];
    out[64 + tiles + k] = (char)cb[k];
  }
  for (int k = i; k < bpx; k += step) out[64 + 2 * tiles + k] = (char)bmap[k];
}

}  // namespace

void launch_cover(hipStream_t s, const Geom& g, const ConstTables* tab, const ProbeParams* pp,
                  float* const xyb[3], double* cost8, double* cost16, double* cost32, double* rad8,
                  double* rad16, double* rad32, uint32_t* queue, uint8_t* bmap, uint32_t* err, int all_exact) {
  int c8 = g.bw * g.bh, c16 = (g.bw / 2) * (g.bh / 2), c32 = (g.bw / 4) * (g.bh / 4);
  k_cover_cost<8><<<div_up(c8, 8), kThreads, 0, s>>>(g, tab, pp, xyb[0], xyb[1], xyb[2], cost8, rad8, err);
  if (c16 > 0)
    k_cover_cost<16><<<div_up(c16, 4), kThreads, 0, s>>>(g, tab, pp, xyb[0], xyb[1], xyb[2], cost16, rad16, err);
  if (c32 > 0)
    k_cover_cost<32><<<div_up(c32, 2), kThreads, 0, s>>>(g, tab, pp, xyb[0], xyb[1], xyb[2], cost32, rad32, err);
  dim3 grid(div_up((g.bw + 3) / 4, 256), (g.bh + 3) / 4);
  k_cover_decide<<<grid, 256, 0, s>>>(g, pp, cost8, cost16, cost32, rad8, rad16, rad32, bmap, queue, all_exact);
  k_cover_exact<<<64, kThreads, 0, s>>>(g, tab, pp, xyb[0], xyb[1], xyb[2], cost8, cost16, cost32, bmap, queue, err);
}

}  // namespace jpxlg

namespace jpxlg {

void launch_cfl_collect(hipStream_t s, const Geom& g, const ConstTables* tab, const ProbeParams* pp,
                        uint32_t* const list[3], const StructInfo* si, const uint32_t* lf_off,
                        const uint32_t* hf_local, const uint32_t* tile_base, float* const coeff[3], float4* lf_s,
                        float4* hf_s, uint16_t* hf_cell, uint32_t* err) {
  // upper bounds of each size's varblock count; workgroups past the count exit
  const int ub[3] = {g.bw * g.bh, (g.bw / 2) * (g.bh / 2), (g.bw / 4) * (g.bh / 4)};
  if (ub[0]) k_cfl_collect<8><<<ub[0], 64, 0, s>>>(g, list[0], si, lf_off, hf_local, tile_base, tab, pp, coeff[0], coeff[1], coeff[2], lf_s, hf_s, hf_cell, err);
  if (ub[1]) k_cfl_collect<16><<<ub[1], 64, 0, s>>>(g, list[1], si, lf_off, hf_local, tile_base, tab, pp, coeff[0], coeff[1], coeff[2], lf_s, hf_s, hf_cell, err);
  if (ub[2]) k_cfl_collect<32><<<ub[2], 64, 0, s>>>(g, list[2], si, lf_off, hf_local, tile_base, tab, pp, coeff[0], coeff[1], coeff[2], lf_s, hf_s, hf_cell, err);
}

int structure_rows(const Geom& g) { return ((g.bw + 255) / 256) * ((g.bh + 255) / 256) * 256; }

void launch_structure_lists(hipStream_t s, const Geom& g, const uint8_t* bmap, uint32_t* const list[3],
                            uint32_t* lf_off, uint32_t* hf_local, uint32_t* tile_count, uint32_t* tile_base,
                            StructInfo* si, uint4* rows) {
  const int nch = structure_rows(g);
  k_rows_count<<<nch, 256, 0, s>>>(g, bmap, rows);
  k_rows_scan<<<1, 1024, 0, s>>>(nch, rows, rows + nch, si);
  k_rows_write<<<nch, 256, 0, s>>>(g, bmap, rows + nch, list[0], list[1], list[2], lf_off);
  const int tiles = g.tw * g.th;
  k_struct_tiles<<<tiles, 64, 0, s>>>(g, bmap, hf_local, tile_count);
  k_struct_tile_base<<<1, 1024, 0, s>>>(tiles, tile_count, tile_base, si);
}

void launch_cfl_tiles(hipStream_t s, const Geom& g, const ConstTables* tab, const ProbeParams* pp,
                      const uint32_t* tile_base, const float4* hf_s, const uint16_t* hf_cell,
                      int8_t* cfl_x, int8_t* cfl_b, const int* gray) {
  const int tiles = g.tw * g.th;
  k_cfl_tiles<<<tiles, kCflThreads, 0, s>>>(tile_base, tiles, tab, pp, hf_s, hf_cell, cfl_x, cfl_b, gray);
}

void launch_cfl_lf(hipStream_t s, const ProbeParams* pp, const StructInfo* si, const float4* lf_s,
                   CflLfScratch* scratch, CflParams* cflp, const int* gray) {
  // fixed grids; every kernel reads the sample count from si
  k_cfl_lf_fast_partial<<<kCflLfFastGroups, 256, 0, s>>>(si, lf_s, scratch);
  k_cfl_lf_fast_final<<<1, 256, 0, s>>>(si, kCflLfFastGroups, scratch);
  k_cfl_lf_seed<<<1, kCflThreads, 0, s>>>(si, lf_s, scratch);
  k_cfl_lf_price<<<256, 256, 0, s>>>(si, pp, lf_s, scratch);
  k_cfl_lf_select<<<1, 64, 0, s>>>(scratch, cflp, gray);
}

void launch_aq_field(hipStream_t s, const Geom& g, int mode, float* const xyb[3], float* act, float* ero,
                     AqScratch* scratch, float* adj) {
  const int count = g.bw * g.bh;
  k_aq_activity<<<div_up(count, 256), 256, 0, s>>>(g, xyb[0], xyb[1], xyb[2], act);
  if (mode >= kAqFineMasking) k_aq_erode<<<div_up(count, 256), 256, 0, s>>>(g.bw, g.bh, act, ero);
  if (mode != kAqEdgeRefine) k_aq_mean<<<1, 1024, 0, s>>>(mode >= kAqFineMasking ? ero : act, count, scratch);
  k_aq_adj<<<div_up(count, 256), 256, 0, s>>>(count, mode, act, ero, scratch, adj);
}

// The QuantizedFrameIr hand-off (jpxl-encode/src/vardct/plan.rs:519-636):
// each varblock's three channels, each in its coefficient array's own
// raster order (cell k of a side-s square at footprint row 8 by + k / s,
// column 8 bx + k % s), stored contiguously at 3 * 64 * lf_off[origin] in
// the frame's arena (lf_off is the varblock's position in LF-group, then
// block-raster order, counted in 8x8 blocks); one workgroup per varblock.
__global__ __launch_bounds__(256) void k_gather_ir(Geom g, const uint32_t* __restrict__ list, int count, int side,
                                                   const uint32_t* __restrict__ lf_off,
                                                   const int32_t* __restrict__ QX, const int32_t* __restrict__ QY,
                                                   const int32_t* __restrict__ QB, int32_t* __restrict__ arena) {
  const int v = blockIdx.x;
  if (v >= count) return;
  const uint32_t e = list[v];
  const int bx = (int)(e & 0xffffu), by = (int)(e >> 16);
  const size_t base = (size_t)192 * lf_off[(size_t)by * g.bw + bx];
  const int nn = side * side;
  const int32_t* q[3] = {QX, QY, QB};
  for (int i = threadIdx.x; i < 3 * nn; i += blockDim.x) {
    const int c = i / nn, k = i % nn;
    arena[base + i] = q[c][(size_t)(8 * by + k / side) * g.cstride + 8 * bx + k % side];
  }
}

// Per LF group (raster order), the group's three LF planes in raster order:
// LfQuantPlanes of every QuantizedLfGroup, concatenated.  A group of the
// frame's block grid is 256 x 256 blocks, clipped at the right and bottom.
__global__ __launch_bounds__(256) void k_gather_lf_groups(Geom g, const int32_t* __restrict__ LX,
                                                          const int32_t* __restrict__ LY,
                                                          const int32_t* __restrict__ LB,
                                                          int32_t* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)g.bw * g.bh) return;
  const int bx = (int)(i % g.bw), by = (int)(i / g.bw);
  const int gx = bx / 256, gy = by / 256;
  const int gw = min(256, g.bw - 256 * gx), gh = min(256, g.bh - 256 * gy);
  const size_t off = 3 * ((size_t)256 * gy * g.bw + (size_t)256 * gx * gh);
  const size_t in_g = (size_t)(by % 256) * gw + bx % 256, plane = (size_t)gw * gh;
  out[off + in_g] = LX[i];
  out[off + plane + in_g] = LY[i];
  out[off + 2 * plane + in_g] = LB[i];
}

void launch_gather_ir(hipStream_t s, const Geom& g, uint32_t* const list[3], const int counts[3],
                      const uint32_t* lf_off, int32_t* const quant[3], int32_t* const lf[3], int32_t* arena,
                      int32_t* lf_out) {
  const int sides[3] = {8, 16, 32};
  if (arena)
    for (int i = 0; i < 3; i++)
      if (counts[i])
        k_gather_ir<<<counts[i], 256, 0, s>>>(g, list[i], counts[i], sides[i], lf_off, quant[0], quant[1], quant[2],
                                              arena);
  const size_t bpx = (size_t)g.bw * g.bh;
  if (lf_out) k_gather_lf_groups<<<div_up(bpx, 256), 256, 0, s>>>(g, lf[0], lf[1], lf[2], lf_out);
}

// Token-ready runs of the quantized plan.  One wavefront per (varblock,
// channel), four per workgroup; the varblock's entries live at its block
// offset in the IR (lf_off), so the tables are indexed 3 * offset + channel
// and entries of blocks that are not a varblock's top-left stay zero.
//
// Count: the HF non-zeros (natural-order positions nblocks .. size - 1) and
// the run length up to the last of them, packed count | nonzeros << 16.
__global__ __launch_bounds__(256) void k_runs_count(Geom g, const uint32_t* __restrict__ list, int count, int side,
                                                    const uint32_t* __restrict__ lf_off,
                                                    const int32_t* __restrict__ QX, const int32_t* __restrict__ QY,
                                                    const int32_t* __restrict__ QB,
                                                    const uint16_t* __restrict__ order,
                                                    uint32_t* __restrict__ len_nz) {
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = (int)(threadIdx.x & 63);
  const int v = wave / 3, c = wave % 3;
  if (v >= count) return;  // the whole wavefront leaves together
  const uint32_t e = list[v];
  const int bx = (int)(e & 0xffffu), by = (int)(e >> 16);
  const int nb1 = side / 8, nblocks = nb1 * nb1, size = side * side;
  const int32_t* q = c == 0 ? QX : (c == 1 ? QY : QB);
  const size_t base = (size_t)(8 * by) * g.cstride + 8 * bx;
  uint32_t nz = 0, last = 0;  // last: natural position + 1 of this lane's last non-zero
  for (int k = nblocks + lane; k < size; k += 64) {
    const int cell = order[k];
    if (q[base + (size_t)(cell / side) * g.cstride + cell % side] != 0) {
      nz++;
      last = (uint32_t)k + 1;
    }
  }
  for (int m = 32; m >= 1; m >>= 1) {
    nz += __shfl_xor(nz, m, 64);
    last = max(last, (uint32_t)__shfl_xor(last, m, 64));
  }
  if (lane == 0) {
    const uint32_t off = lf_off[(size_t)by * g.bw + bx];
    len_nz[3 * (size_t)off + c] = (last ? last - (uint32_t)nblocks : 0u) | nz << 16;
  }
}

// Exclusive prefix sum of the run lengths in three passes: each workgroup
// scans 1024 entries (4 per thread, in order) and leaves its total in part;
// one workgroup scans the totals (start[n] receives the sum); every
// workgroup then adds its offset.
constexpr int kScanTile = 1024;

__global__ __launch_bounds__(256) void k_runs_scan_local(const uint32_t* __restrict__ len_nz, int n,
                                                         uint32_t* __restrict__ start, uint32_t* __restrict__ part) {
  __shared__ uint32_t sh[256];
  const int t = (int)threadIdx.x, base = (int)blockIdx.x * kScanTile + 4 * t;
  uint32_t v[4], sum = 0;
  for (int i = 0; i < 4; i++) {
    v[i] = base + i < n ? (len_nz[base + i] & 0xffffu) : 0u;
    sum += v[i];
  }
  sh[t] = sum;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    const uint32_t add = t >= d ? sh[t - d] : 0u;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  uint32_t run = t ? sh[t - 1] : 0u;
  for (int i = 0; i < 4; i++) {
    if (base + i < n) start[base + i] = run;
    run += v[i];
  }
  if (t == 255) part[blockIdx.x] = sh[255];
}

__global__ __launch_bounds__(1024) void k_runs_scan_parts(uint32_t* __restrict__ part, int nb,
                                                          uint32_t* __restrict__ total) {
  __shared__ uint32_t sh[1024];
  const int t = (int)threadIdx.x;
  const int per = (nb + 1023) / 1024, b = min(nb, t * per), e = min(nb, b + per);
  uint32_t sum = 0;
  for (int i = b; i < e; i++) sum += part[i];
  sh[t] = sum;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const uint32_t add = t >= d ? sh[t - d] : 0u;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  uint32_t run = t ? sh[t - 1] : 0u;
  for (int i = b; i < e; i++) {
    const uint32_t x = part[i];
    part[i] = run;
    run += x;
  }
  if (t == 1023) *total = sh[1023];
}

__global__ __launch_bounds__(256) void k_runs_scan_add(uint32_t* __restrict__ start, int n,
                                                       const uint32_t* __restrict__ part) {
  const int base = (int)blockIdx.x * kScanTile + 4 * (int)threadIdx.x;
  const uint32_t off = part[blockIdx.x];
  for (int i = 0; i < 4; i++)
    if (base + i < n) start[base + i] += off;
}

// Write: each run's coefficients in natural order at its start.
__global__ __launch_bounds__(256) void k_runs_write(Geom g, const uint32_t* __restrict__ list, int count, int side,
                                                    const uint32_t* __restrict__ lf_off,
                                                    const int32_t* __restrict__ QX, const int32_t* __restrict__ QY,
                                                    const int32_t* __restrict__ QB,
                                                    const uint16_t* __restrict__ order,
                                                    const uint32_t* __restrict__ len_nz,
                                                    const uint32_t* __restrict__ start, int32_t* __restrict__ arena) {
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = (int)(threadIdx.x & 63);
  const int v = wave / 3, c = wave % 3;
  if (v >= count) return;
  const uint32_t e = list[v];
  const int bx = (int)(e & 0xffffu), by = (int)(e >> 16);
  const int nb1 = side / 8, nblocks = nb1 * nb1;
  const int32_t* q = c == 0 ? QX : (c == 1 ? QY : QB);
  const size_t base = (size_t)(8 * by) * g.cstride + 8 * bx;
  const size_t at = 3 * (size_t)lf_off[(size_t)by * g.bw + bx] + c;
  const uint32_t n = len_nz[at] & 0xffffu;
  int32_t* out = arena + start[at];
  for (uint32_t j = (uint32_t)lane; j < n; j += 64) {
    const int cell = order[nblocks + (int)j];
    out[j] = q[base + (size_t)(cell / side) * g.cstride + cell % side];
  }
}

void launch_runs(hipStream_t s, const Geom& g, uint32_t* const list[3], const int counts[3], const uint32_t* lf_off,
                 int32_t* const quant[3], const ConstTables* tab, uint32_t* len_nz, uint32_t* start,
                 int32_t* arena) {
  const int sides[3] = {8, 16, 32}, order_at[3] = {0, 64, 64 + 256};
  const int n = 3 * g.bw * g.bh;
  k_zero_u32<<<div_up(n, 256), 256, 0, s>>>(len_nz, n);
  for (int i = 0; i < 3; i++)
    if (counts[i])
      k_runs_count<<<div_up(3 * counts[i], 4), 256, 0, s>>>(g, list[i], counts[i], sides[i], lf_off, quant[0],
                                                             quant[1], quant[2], tab->order + order_at[i], len_nz);
  const int nb = div_up(n, kScanTile);
  uint32_t* part = start + n + 1;  // nb words after the total
  k_runs_scan_local<<<nb, 256, 0, s>>>(len_nz, n, start, part);
  k_runs_scan_parts<<<1, 1024, 0, s>>>(part, nb, start + n);
  k_runs_scan_add<<<nb, 256, 0, s>>>(start, n, part);
  for (int i = 0; i < 3; i++)
    if (counts[i])
      k_runs_write<<<div_up(3 * counts[i], 4), 256, 0, s>>>(g, list[i], counts[i], sides[i], lf_off, quant[0],
                                                             quant[1], quant[2], tab->order + order_at[i], len_nz,
                                                             start, arena);
}


void launch_select(hipStream_t s, const float* luma, const float* chroma, int n, float flat_below,
                   const SelectRanks& r, SelectScratch* sc) {
  const int words = (int)(offsetof(SelectScratch, prefix) / 4);
  k_zero_u32<<<div_up(words, 256), 256, 0, s>>>((uint32_t*)sc, words);
  k_select_hist<0><<<kSelectBlocks, 256, 0, s>>>(luma, chroma, n, flat_below, r, sc);
  k_select_digit<0><<<kSelectQueries, 256, 0, s>>>(r, sc);
  k_select_hist<1><<<kSelectBlocks, 256, 0, s>>>(luma, chroma, n, flat_below, r, sc);
  k_select_digit<1><<<kSelectQueries, 256, 0, s>>>(r, sc);
  k_select_hist<2><<<kSelectBlocks, 256, 0, s>>>(luma, chroma, n, flat_below, r, sc);
  k_select_digit<2><<<kSelectQueries, 256, 0, s>>>(r, sc);
}

int preanalysis_cells(int w, int h) { return (w / 32) * (h / 32); }

void launch_preanalysis_cells(hipStream_t s, int w, int h, int bits, const void* rgb, CellStats* out) {
  const int n = preanalysis_cells(w, h);
  if (n == 0) return;
  if (bits == 8)
    k_preanalysis_cells<true><<<n, 256, 0, s>>>(w, h, 0, rgb, w / 32, out);
  else
    k_preanalysis_cells<false><<<n, 256, 0, s>>>(w, h, bits - 8, rgb, w / 32, out);
}

void launch_aq_block_muls(hipStream_t s, const Geom& g, const ProbeParams* pp, uint32_t* const list[3],
                          const int counts[3], uint32_t* muls) {
  if (counts[0]) k_aq_block_muls<8><<<div_up(counts[0], 256), 256, 0, s>>>(g, pp, list[0], counts[0], muls);
  if (counts[1]) k_aq_block_muls<16><<<div_up(counts[1], 256), 256, 0, s>>>(g, pp, list[1], counts[1], muls);
  if (counts[2]) k_aq_block_muls<32><<<div_up(counts[2], 256), 256, 0, s>>>(g, pp, list[2], counts[2], muls);
}

size_t structure_readback_bytes(int tiles, int bpx) { return 64 + 2 * (size_t)tiles + (size_t)bpx; }

void launch_gather_structure(hipStream_t s, const StructInfo* si, const CflParams* cp, const uint32_t* err,
                             const int8_t* cx, const int8_t* cb, const uint8_t* bmap, int tiles, int bpx,
                             char* out) {
  const int n = std::max(tiles, bpx);
  k_gather_structure<<<std::min((int)div_up(n, 256), 1024), 256, 0, s>>>(si, cp, err, cx, cb, bmap, tiles, bpx, out);
}

}  // namespace jpxlg
