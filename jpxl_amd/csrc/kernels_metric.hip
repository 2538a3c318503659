// gfx950 kernels for the colour stage and SSIMULACRA2: XYB^-1 + depth
// quantizer, positive XYB, the 2:1 pyramid, the sigma-1.5 Gaussian blur and
// the fused SSIM / artifact / detail-lost pooling with f64 partial sums.
//
// The blur is the 9-tap FIR that the reference's recursive truncated-cosine
// filter equals in exact arithmetic (jpxl-perceptual/src/blur.rs:1-51): the
// three poles sit on the unit circle at 2cos(w_k) with w_k = k*pi/10, k odd,
// so the response to an impulse vanishes beyond |t| = 4.  The taps are the
// recursion's own impulse response, computed on the host in f64.  Zero padding
// at the plane border is kept, exactly as the metric defines it.
#include "internal.h"

namespace jpxlg {

namespace {

inline unsigned div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// x^(1/2.4) for the depth quantizer's first guess: the hardware log2 and
// exp2 (a few f32 ulps; the library pow costs ~160 instructions).  The guess
// only picks where the exact threshold search starts.
__device__ __forceinline__ float guess_pow(float x) {
  return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 2.4f));
}

// Depth tables of up to 256 levels are staged in LDS (TAB_LDS), larger ones
// are read from global memory; the accessor tells the compiler which.
template <bool TAB_LDS>
__device__ __forceinline__ float tab_at(const float* p, int i) {
  if constexpr (TAB_LDS) return ((const __attribute__((address_space(3))) float*)p)[i];
  else return ldg32(p, (uint32_t)i);
}

// jpxl-plan-render/src/lib.rs:341-347 level_of and :421-454 linear_at_depth:
// the level's linear value for one sample, `lut[#thresholds <= x]`.
template <bool TAB_LDS>
__device__ __forceinline__ float depth_value(float x, const float* thr, const float* lut,
                                             int levels) {
  if (!isfinite(x)) return tab_at<TAB_LDS>(lut, 0);
  // A first guess from the device's own sRGB encode (within a level of the
  // truth), then corrected against the exact thresholds, so the result is
  // the count of thresholds <= x whatever the guess.
  const float max = (float)(levels - 1);
  float e;
  if (x <= 0.0031308f) e = 12.92f * x;
  else e = 1.055f * guess_pow(x) - 0.055f;
  float gl = rintf(e * max);
  const int g = gl <= 0.0f ? 0 : (gl >= max ? levels - 1 : (int)gl);
  // The two thresholds either side of the guess, loaded together; the walks
  // run only when the guess is off.
  const float tl = g > 0 ? tab_at<TAB_LDS>(thr, g - 1) : -INFINITY;
  const float th = g < levels - 1 ? tab_at<TAB_LDS>(thr, g) : INFINITY;
  int lo = g;
  if (tl > x) {
    lo = g - 1;
    while (lo > 0 && tab_at<TAB_LDS>(thr, lo - 1) > x) lo--;
  } else if (th <= x) {
    lo = g + 1;
    while (lo < levels - 1 && tab_at<TAB_LDS>(thr, lo) <= x) lo++;
  }
  return tab_at<TAB_LDS>(lut, lo);
}

// depth_value over M samples at once: every guess first, then all their
// threshold pairs (clamped indices, no branches), the rare corrective walks,
// then all the level values -- two rounds of table latency for the M samples
// instead of two per sample.  Same results as depth_value.
template <bool TAB_LDS, int M>
__device__ __forceinline__ void depth_values(const float (&x)[M], float (&out)[M], const float* thr,
                                             const float* lut, int levels) {
  const float lmax = (float)(levels - 1);
  const int top = max(levels - 2, 0);  // last threshold index
  int g[M];
#pragma unroll
  for (int i = 0; i < M; i++) {
    float e;
    if (x[i] <= 0.0031308f) e = 12.92f * x[i];
    else e = 1.055f * guess_pow(x[i]) - 0.055f;
    const float gl = rintf(e * lmax);
    g[i] = !(gl > 0.0f) ? 0 : (gl >= lmax ? levels - 1 : (int)gl);  // NaN -> 0 (unused)
  }
  float tl[M], th[M];
#pragma unroll
  for (int i = 0; i < M; i++) {
    tl[i] = levels > 1 ? tab_at<TAB_LDS>(thr, min(max(g[i] - 1, 0), top)) : 0.0f;
    th[i] = levels > 1 ? tab_at<TAB_LDS>(thr, min(g[i], top)) : 0.0f;
  }
  int lo[M];
#pragma unroll
  for (int i = 0; i < M; i++) {
    const float xi = x[i];
    const float l = g[i] > 0 ? tl[i] : -INFINITY;
    const float hgh = g[i] < levels - 1 ? th[i] : INFINITY;
    int v = g[i];
    if (!isfinite(xi)) {
      v = 0;
    } else if (l > xi) {
      v = g[i] - 1;
      while (v > 0 && tab_at<TAB_LDS>(thr, v - 1) > xi) v--;
    } else if (hgh <= xi) {
      v = g[i] + 1;
      while (v < levels - 1 && tab_at<TAB_LDS>(thr, v) <= xi) v++;
    }
    lo[i] = v;
  }
#pragma unroll
  for (int i = 0; i < M; i++) out[i] = tab_at<TAB_LDS>(lut, lo[i]);
}

// jpxl-perceptual/src/color.rs:66-81 linear_srgb_to_positive_xyb.
__device__ __forceinline__ void positive_xyb(float r, float g, float b, float* ox, float* oy,
                                             float* ob) {
  const float bias = 0.0037930734f, bias_cbrt = 0.1559542f;
  float lm = 0.3f * r + 0.622f * g + 0.078f * b + bias;
  float mm = 0.23f * r + 0.692f * g + 0.078f * b + bias;
  float sm = 0.2434227f * r + 0.2047674f * g + 0.5518099f * b + bias;
  float lg = cbrt_metric(lm) - bias_cbrt;
  float mg = cbrt_metric(mm) - bias_cbrt;
  float sg = cbrt_metric(sm) - bias_cbrt;
  float x = 0.5f * (lg - mg);
  float y = 0.5f * (lg + mg);
  *ox = x * 14.0f + 0.42f;
  *oy = y + 0.01f;
  *ob = (sg - y) + 0.55f;
}

// jpxl-core/src/color.rs:673-700 OpsinInverse::convert with the Table L.1
// defaults of jpxl-plan-render/src/lib.rs:55-68 (itscale = 1).
__device__ __forceinline__ void opsin_inverse(float x, float y, float b, float root, float* o) {
  const float bias = -0.0037930733f;
  const float m[9] = {11.031567f, -9.866944f, -0.16462299f, -3.2541474f, 4.4187705f,
                      -0.16462299f, -3.6588513f, 2.712923f, 1.9459282f};
  float gam[3] = {y + x, y - x, b};
  float mix[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    float v = gam[c] - root;
    mix[c] = (v * v * v + bias) * 1.0f;
  }
#pragma unroll
  for (int r = 0; r < 3; r++) {
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; c++) acc = fmaf(m[r * 3 + c], mix[c], acc);
    o[r] = acc;
  }
}

// One thread per 2x2 quad: converts up to four samples of this scale and
// writes the next scale's box-filtered linear RGB (pyramid.rs:19-44).
template <int MODE, bool TAB_LDS>
__device__ __forceinline__ void colour_quad(int w, int h, const float* __restrict__ in0,
                                            const float* __restrict__ in1, const float* __restrict__ in2,
                                            const float* thr, const float* lut, int levels, float* __restrict__ p0,
                                            float* __restrict__ p1, float* __restrict__ p2, float* __restrict__ d0,
                                            float* __restrict__ d1, float* __restrict__ d2, int ox, int oy) {
  const int w2 = (w + 1) / 2, h2 = (h + 1) / 2;
  if (ox >= w2 || oy >= h2) return;
  const float root = MODE == 0 ? cbrt_core(-0.0037930733f) : 0.0f;
  float lin[2][2][3];
  float dv[12];
  if constexpr (MODE == 0) {
    float o[12];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int x = min(2 * ox + (q & 1), w - 1), y = min(2 * oy + (q >> 1), h - 1);
      const size_t i = (size_t)y * w + x;
      opsin_inverse(in0[i], in1[i], in2[i], root, o + 3 * q);
    }
    depth_values<TAB_LDS, 12>(o, dv, thr, lut, levels);
  }
#pragma unroll
  for (int dy = 0; dy < 2; dy++)
#pragma unroll
    for (int dx = 0; dx < 2; dx++) {
      const int x = min(2 * ox + dx, w - 1), y = min(2 * oy + dy, h - 1);
      const size_t i = (size_t)y * w + x;
      float v[3];
      if constexpr (MODE == 0) {
#pragma unroll
        for (int c = 0; c < 3; c++) v[c] = dv[3 * (2 * dy + dx) + c];
      } else {
        v[0] = in0[i]; v[1] = in1[i]; v[2] = in2[i];
      }
      lin[dy][dx][0] = v[0]; lin[dy][dx][1] = v[1]; lin[dy][dx][2] = v[2];
      if (2 * ox + dx < w && 2 * oy + dy < h) {
        float a, b, c;
        positive_xyb(v[0], v[1], v[2], &a, &b, &c);
        p0[i] = a; p1[i] = b; p2[i] = c;
      }
    }
  if (d0) {
    const size_t o = (size_t)oy * w2 + ox;
    float* d[3] = {d0, d1, d2};
#pragma unroll
    for (int c = 0; c < 3; c++)
      d[c][o] = (lin[0][0][c] + lin[0][1][c] + lin[1][0][c] + lin[1][1][c]) * 0.25f;
  }
}

// Stages depth thresholds of up to 256 levels in LDS (the workgroup's image
// is uniform); larger tables are read from global memory.
__device__ __forceinline__ void stage_depth(const float* thr_g, const float* lut_g, int levels, float* s_thr,
                                            float* s_lut, const float** thr, const float** lut) {
  *thr = thr_g;
  *lut = lut_g;
  if (levels <= 256) {
    for (int i = threadIdx.x; i < levels; i += blockDim.x) {
      s_lut[i] = lut_g[i];
      if (i < levels - 1) s_thr[i] = thr_g[i];
    }
    *thr = s_thr;
    *lut = s_lut;
  }
  __syncthreads();
}

template <int MODE>
__global__ __launch_bounds__(256) void k_colour_metric(
    int w, int h, const float* __restrict__ in0, const float* __restrict__ in1,
    const float* __restrict__ in2, const float* __restrict__ thr_g, const float* __restrict__ lut_g,
    int levels, float* __restrict__ p0, float* __restrict__ p1, float* __restrict__ p2,
    float* __restrict__ d0, float* __restrict__ d1, float* __restrict__ d2) {
  __shared__ float s_thr[256], s_lut[256];
  const float* thr = thr_g;
  const float* lut = lut_g;
  if constexpr (MODE == 0) stage_depth(thr_g, lut_g, levels, s_thr, s_lut, &thr, &lut);
  if (MODE == 0 && levels <= 256)
    colour_quad<MODE, true>(w, h, in0, in1, in2, thr, lut, levels, p0, p1, p2, d0, d1, d2,
                            blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
  else
    colour_quad<MODE, false>(w, h, in0, in1, in2, thr, lut, levels, p0, p1, p2, d0, d1, d2,
                             blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// Batched colour pass of scale s: workgroup (quad row, 256-quad block) of
// each desc.  Scale 0 reads the reconstructed XYB (render planes, src 0, or
// their ping-pong, src 1) through the depth quantizer; later scales read the
// previous scale's downsampled linear RGB.
template <int MODE>
__global__ __launch_bounds__(256) void k_colour_b(const ProbeDesc* __restrict__ D, const int* __restrict__ off,
                                                  int nd, int s, int src) {
  __shared__ float s_thr[256], s_lut[256];
  const int d = find_desc(off, nd, blockIdx.x);
  DescSpace& P = desc_at(D, d);
  const int w = P.sw[s], h = P.sh[s];
  const float* thr = P.thr;
  const float* lut = P.lut;
  if constexpr (MODE == 0) stage_depth(P.thr, P.lut, P.levels, s_thr, s_lut, &thr, &lut);
  const int lb = blockIdx.x - off[d], xb = ((w + 1) / 2 + 255) / 256;
  const float* in[3];
  float* dn[3] = {nullptr, nullptr, nullptr};
  for (int c = 0; c < 3; c++) {
    in[c] = MODE == 0 ? (src ? P.aux[c] : P.rx[c]) : P.cand_rgb[s][c];
    if (s + 1 < P.nscales) dn[c] = P.cand_rgb[s + 1][c];
  }
  if (MODE == 0 && P.levels <= 256)
    colour_quad<MODE, true>(w, h, in[0], in[1], in[2], thr, lut, P.levels, P.cand_x[s][0], P.cand_x[s][1],
                            P.cand_x[s][2], dn[0], dn[1], dn[2], (lb % xb) * 256 + threadIdx.x, lb / xb);
  else
    colour_quad<MODE, false>(w, h, in[0], in[1], in[2], thr, lut, P.levels, P.cand_x[s][0], P.cand_x[s][1],
                             P.cand_x[s][2], dn[0], dn[1], dn[2], (lb % xb) * 256 + threadIdx.x, lb / xb);
}

// The colour pass of every scale in one workgroup: a 32x32 tile of scale 0
// and the tiles it covers at the coarser scales (16x16, 8x8, ... 1x1; six
// scales at most).  Scale 0 converts its samples by 2x2 quads (XYB^-1 +
// depth for MODE 0, the linear input for MODE 1), writes their positive XYB
// and box-filters each quad into scale 1; every coarser scale, held in LDS
// (a and b alternate), writes its pixels' positive XYB and box-filters its
// quads into the next.  The per-sample arithmetic, clamps and order are
// colour_quad's.
constexpr int kPyrTile = 32;
// Sizes and output planes are read where they lie (a descriptor in constant
// memory or the kernel's parameter block): IntP and OutP are their pointer
// types.
template <int MODE, bool TAB_LDS, class IntP, class OutP>
__device__ void colour_pyramid_tile(int ns, IntP sw, IntP sh, const float* const* in, OutP out,
                                    const float* thr, const float* lut, int levels,
                                    int tx0, int ty0, float (*a)[256], float (*b)[64]) {
  const float root = MODE == 0 ? cbrt_core(-0.0037930733f) : 0.0f;
  {
    const int w = sw[0], h = sh[0], w2 = (w + 1) / 2, h2 = (h + 1) / 2;
    for (int q = threadIdx.x; q < 16 * 16; q += blockDim.x) {
      const int qx = q % 16, qy = q / 16, ox = tx0 / 2 + qx, oy = ty0 / 2 + qy;
      if (ox >= w2 || oy >= h2) continue;
      float lin[2][2][3];
      float dv[12];
      if constexpr (MODE == 0) {
        float o[12];
#pragma unroll
        for (int q4 = 0; q4 < 4; q4++) {
          const int x = min(2 * ox + (q4 & 1), w - 1), y = min(2 * oy + (q4 >> 1), h - 1);
          const size_t i = (size_t)y * w + x;
          opsin_inverse(in[0][i], in[1][i], in[2][i], root, o + 3 * q4);
        }
        depth_values<TAB_LDS, 12>(o, dv, thr, lut, levels);
      }
#pragma unroll
      for (int dy = 0; dy < 2; dy++)
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
          const int x = min(2 * ox + dx, w - 1), y = min(2 * oy + dy, h - 1);
          const size_t i = (size_t)y * w + x;
          float v[3];
          if constexpr (MODE == 0) {
#pragma unroll
            for (int c = 0; c < 3; c++) v[c] = dv[3 * (2 * dy + dx) + c];
          } else {
            v[0] = in[0][i]; v[1] = in[1][i]; v[2] = in[2][i];
          }
          lin[dy][dx][0] = v[0]; lin[dy][dx][1] = v[1]; lin[dy][dx][2] = v[2];
          if (2 * ox + dx < w && 2 * oy + dy < h) {
            float p0, p1, p2;
            positive_xyb(v[0], v[1], v[2], &p0, &p1, &p2);
            out[0][0][i] = p0; out[0][1][i] = p1; out[0][2][i] = p2;
          }
        }
      if (ns > 1)
#pragma unroll
        for (int c = 0; c < 3; c++)
          a[c][qy * 16 + qx] = (lin[0][0][c] + lin[0][1][c] + lin[1][0][c] + lin[1][1][c]) * 0.25f;
    }
  }
  for (int s = 1; s < ns; s++) {
    __syncthreads();
    const int T = kPyrTile >> s, w = sw[s], h = sh[s], X = tx0 >> s, Y = ty0 >> s;
    const float* cur[3] = {(s & 1) ? a[0] : b[0], (s & 1) ? a[1] : b[1], (s & 1) ? a[2] : b[2]};
    float* nxt[3] = {(s & 1) ? b[0] : a[0], (s & 1) ? b[1] : a[1], (s & 1) ? b[2] : a[2]};
    if (X >= w || Y >= h) continue;
    // this scale's pixels
    for (int p = threadIdx.x; p < T * T; p += blockDim.x) {
      const int x = X + p % T, y = Y + p / T;
      if (x >= w || y >= h) continue;
      float p0, p1, p2;
      positive_xyb(cur[0][p], cur[1][p], cur[2][p], &p0, &p1, &p2);
      const uint32_t i = (uint32_t)(y * w + x);
      stg32(out[s][0], i, p0);
      stg32(out[s][1], i, p1);
      stg32(out[s][2], i, p2);
    }
    if (s + 1 >= ns) continue;
    // the next scale's tile (T/2 x T/2) from this scale's quads
    const int w2 = (w + 1) / 2, h2 = (h + 1) / 2, hq = T / 2;
    for (int q = threadIdx.x; q < hq * hq; q += blockDim.x) {
      const int qx = q % hq, qy = q / hq, ox = X / 2 + qx, oy = Y / 2 + qy;
      if (ox >= w2 || oy >= h2) continue;
      const int x0 = 2 * ox - X, y0 = 2 * oy - Y;
      const int x1 = min(2 * ox + 1, w - 1) - X, y1 = min(2 * oy + 1, h - 1) - Y;
#pragma unroll
      for (int c = 0; c < 3; c++)
        nxt[c][qy * hq + qx] = (cur[c][y0 * T + x0] + cur[c][y0 * T + x1] + cur[c][y1 * T + x0] + cur[c][y1 * T + x1]) *
                               0.25f;
    }
  }
}

// Batched colour pyramid: workgroups enumerate (desc, 64x64 tile of scale 0)
// through off[]; the candidate's scale 0 is its restored planes (src 0) or
// their ping-pong (src 1).
__global__ __launch_bounds__(256) void k_colour_pyr_b(const ProbeDesc* __restrict__ D, const int* __restrict__ off,
                                                      int nd, int src) {
  __shared__ float s_thr[256], s_lut[256];
  __shared__ float a[3][256], b[3][64];
  const int d = find_desc(off, nd, blockIdx.x);
  DescSpace& P = desc_at(D, d);
  const float* thr = P.thr;
  const float* lut = P.lut;
  stage_depth(P.thr, P.lut, P.levels, s_thr, s_lut, &thr, &lut);
  const int lb = blockIdx.x - off[d], ntx = (P.sw[0] + kPyrTile - 1) / kPyrTile;
  const float* in[3] = {src ? P.aux[0] : P.rx[0], src ? P.aux[1] : P.rx[1], src ? P.aux[2] : P.rx[2]};
  const int ns = min(P.nscales, kMaxPyr);
  if (P.levels <= 256)
    colour_pyramid_tile<0, true>(ns, P.sw, P.sh, in, P.cand_x, thr, lut, P.levels, (lb % ntx) * kPyrTile,
                                 (lb / ntx) * kPyrTile, a, b);
  else
    colour_pyramid_tile<0, false>(ns, P.sw, P.sh, in, P.cand_x, thr, lut, P.levels, (lb % ntx) * kPyrTile,
                                  (lb / ntx) * kPyrTile, a, b);
}

// The reference's pyramid of one image from its linear RGB.
struct PyrPlanes {
  int ns, sw[kMaxPyr], sh[kMaxPyr];
  const float* in[3];
  float* out[kMaxPyr][3];
};
__global__ __launch_bounds__(256) void k_colour_pyr_ref(PyrPlanes A) {
  __shared__ float a[3][256], b[3][64];
  const int ntx = (A.sw[0] + kPyrTile - 1) / kPyrTile;
  colour_pyramid_tile<1, false>(A.ns, A.sw, A.sh, A.in, A.out, nullptr, nullptr, 0, (blockIdx.x % ntx) * kPyrTile,
                         (blockIdx.x / ntx) * kPyrTile, a, b);
}

__global__ __launch_bounds__(256) void k_linear_at_depth(int n, const float* __restrict__ x,
                                                         const float* __restrict__ y,
                                                         const float* __restrict__ b,
                                                         const float* __restrict__ thr,
                                                         const float* __restrict__ lut, int levels,
                                                         float* __restrict__ r, float* __restrict__ g,
                                                         float* __restrict__ bl) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float root = cbrt_core(-0.0037930733f);
  float o[3];
  opsin_inverse(x[i], y[i], b[i], root, o);
  r[i] = depth_value<false>(o[0], thr, lut, levels);
  g[i] = depth_value<false>(o[1], thr, lut, levels);
  bl[i] = depth_value<false>(o[2], thr, lut, levels);
}

// ----------------------------------------------------------------------------
// Blur (FIR form of blur.rs) and pooling (pool.rs:97-278)
// ----------------------------------------------------------------------------

constexpr int kTW = 64, kTH = 16, kHalo = 4;
constexpr int kLW = kTW + 2 * kHalo, kLH = kTH + 2 * kHalo;

// Accumulated in f64 with one rounding to f32 per pass, as the reference's
// f64-state recursion rounds each pass's output once.
__device__ __forceinline__ float fir9(const float* v, int stride, const double* hf) {
  // v points at the centre sample.
  double acc = hf[0] * (double)v[0];
#pragma unroll
  for (int t = 1; t <= 4; t++) acc = fma(hf[t], (double)v[-t * stride] + (double)v[t * stride], acc);
  return (float)acc;
}

__device__ __forceinline__ void blur_moments_tile(int w, int h, const ConstTables* __restrict__ tab,
                                                  const float* __restrict__ x, float* __restrict__ mu,
                                                  float* __restrict__ s11, int x0, int y0) {
  __shared__ float tin[kLH][kLW];
  __shared__ float hb[2][kLH][kTW];
  double hf[5];
#pragma unroll
  for (int t = 0; t < 5; t++) hf[t] = tab->fir[t];
  for (int e = threadIdx.x; e < kLH * kLW; e += 256) {
    int ly = e / kLW, lx = e % kLW, gx = x0 + lx - kHalo, gy = y0 + ly - kHalo;
    tin[ly][lx] = (gx >= 0 && gx < w && gy >= 0 && gy < h) ? x[(size_t)gy * w + gx] : 0.0f;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kLH * kTW; e += 256) {
    int ly = e / kTW, lx = e % kTW;
    const float* c = &tin[ly][lx + kHalo];
    // the squared input is rounded to f32 first (blur.rs:97-113)
    double a = hf[0] * (double)c[0], b = hf[0] * (double)(c[0] * c[0]);
#pragma unroll
    for (int t = 1; t <= 4; t++) {
      a = fma(hf[t], (double)c[-t] + (double)c[t], a);
      b = fma(hf[t], (double)(c[-t] * c[-t]) + (double)(c[t] * c[t]), b);
    }
    hb[0][ly][lx] = (float)a;
    hb[1][ly][lx] = (float)b;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kTH * kTW; e += 256) {
    int ly = e / kTW, lx = e % kTW, gx = x0 + lx, gy = y0 + ly;
    if (gx >= w || gy >= h) continue;
    mu[(size_t)gy * w + gx] = fir9(&hb[0][ly + kHalo][lx], kTW, hf);
    s11[(size_t)gy * w + gx] = fir9(&hb[1][ly + kHalo][lx], kTW, hf);
  }
}

__global__ __launch_bounds__(256) void k_blur_moments(int w, int h, const ConstTables* __restrict__ tab,
                                                      const float* __restrict__ x, float* __restrict__ mu,
                                                      float* __restrict__ s11) {
  blur_moments_tile(w, h, tab, x, mu, s11, blockIdx.x * kTW, blockIdx.y * kTH);
}

// Every scale and channel of a reference pyramid in one launch: workgroup
// (tile of the flattened per-scale tile ranges, channel blockIdx.y).
struct MomentPlanes {
  int ns, sw[kMaxPyr], sh[kMaxPyr], off[kMaxPyr + 1];
  const float* x[kMaxPyr][3];
  float* mu[kMaxPyr][3];
  float* s11[kMaxPyr][3];
};
__global__ __launch_bounds__(256) void k_blur_moments_all(const ConstTables* __restrict__ tab, MomentPlanes A) {
  int s = 0;
  while (s + 1 < A.ns && (int)blockIdx.x >= A.off[s + 1]) s++;
  const int lt = blockIdx.x - A.off[s], tx = (A.sw[s] + kTW - 1) / kTW, c = blockIdx.y;
  blur_moments_tile(A.sw[s], A.sh[s], tab, A.x[s][c], A.mu[s][c], A.s11[s][c], (lt % tx) * kTW, (lt / tx) * kTH);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// The candidate pass of one scale and channel over a 64x24 tile: the three
// candidate blurs (mu2 = G*x2, s22 = G*x2^2, s12 = G*x1x2), the SSIM and
// edge maps, and the tile's f64 partial sums (pool.rs:97-278).  Each lane
// slides its 9-tap window over a run of outputs held in registers, so every
// input sample is widened to f64 once per run instead of once per tap; the
// taps, their pairing and the one f32 rounding per pass are those of fir9.
constexpr int kPW = 64, kPH = 24;
constexpr int kPIW = kPW + 2 * kHalo, kPIH = kPH + 2 * kHalo;  // 72 x 32
constexpr int kHRun = 8;                                      // horizontal outputs per lane
__host__ __device__ constexpr int pool_tiles_dev(int w, int h) { return ((w + kPW - 1) / kPW) * ((h + kPH - 1) / kPH); }
// Streaming pass: 64-column strips cut into segments of kSH output rows.
constexpr int kSH = 64;
__host__ __device__ constexpr int pool_units_dev(int w, int h) { return ((w + kPW - 1) / kPW) * ((h + kSH - 1) / kSH); }
constexpr int kVRun = 6;                                      // vertical outputs per lane

template <int R>
__device__ __forceinline__ void fir_run(const double (&d)[R + 8], const double* hf, double (&o)[R]) {
#pragma unroll
  for (int i = 0; i < R; i++) {
    double acc = hf[0] * d[i + 4];
#pragma unroll
    for (int t = 1; t <= 4; t++) acc = fma(hf[t], d[i + 4 - t] + d[i + 4 + t], acc);
    o[i] = acc;
  }
}

// One 64x24 tile of channel c at one scale; writes the tile's 6 partial
// sums of that channel to out6[0..5].
__device__ __forceinline__ void blur_pool_tile(const ConstTables* __restrict__ tab, int w, int h, int x0, int y0,
                                               const float* __restrict__ X1, const float* __restrict__ MU1,
                                               const float* __restrict__ S11, const float* __restrict__ X2,
                                               double* __restrict__ out6) {
  __shared__ float t1[kPIH][kPIW];
  __shared__ float t2[kPIH][kPIW];
  __shared__ float hb[kPIH][kPW];
  __shared__ double red[4][6];
  double hf[5];
#pragma unroll
  for (int t = 0; t < 5; t++) hf[t] = tab->fir[t];
  // The pooling inputs do not depend on the blurs: issue their loads first.
  const int plx = threadIdx.x & 63, ply0 = (threadIdx.x >> 6) * kVRun;
  float pmu[kVRun], ps11[kVRun];
#pragma unroll
  for (int i = 0; i < kVRun; i++) {
    const int gx = x0 + plx, gy = y0 + ply0 + i;
    const bool in = gx < w && gy < h;
    pmu[i] = in ? MU1[(size_t)gy * w + gx] : 0.0f;
    ps11[i] = in ? S11[(size_t)gy * w + gx] : 0.0f;
  }
  for (int e = threadIdx.x; e < kPIH * kPIW; e += 256) {
    const int ly = e / kPIW, lx = e % kPIW, gx = x0 + lx - kHalo, gy = y0 + ly - kHalo;
    const bool in = gx >= 0 && gx < w && gy >= 0 && gy < h;
    const size_t o = (size_t)gy * w + gx;
    t1[ly][lx] = in ? X1[o] : 0.0f;
    t2[ly][lx] = in ? X2[o] : 0.0f;
  }
  __syncthreads();
  // The three moments one after another through one LDS plane (a smaller
  // footprint keeps more workgroups resident): horizontal 32 rows x 8 runs
  // of 8 outputs, then vertical 64 columns x 4 runs of 6 rows (wave k takes
  // rows 6k..6k+5).
  const int hly = threadIdx.x >> 3, hlx0 = (threadIdx.x & 7) * kHRun;
  const int lx = threadIdx.x & 63, ly0 = (threadIdx.x >> 6) * kVRun;
  const int gx = x0 + lx;
  float vm[3][kVRun];
#pragma unroll
  for (int m = 0; m < 3; m++) {
    {
      double d[kHRun + 8], o[kHRun];
#pragma unroll
      for (int i = 0; i < kHRun + 8; i++) {
        const float a = t1[hly][hlx0 + i], b = t2[hly][hlx0 + i];
        // the squared and product inputs are rounded to f32 first (blur.rs:97-113)
        d[i] = m == 0 ? (double)b : (m == 1 ? (double)(b * b) : (double)(a * b));
      }
      fir_run<kHRun>(d, hf, o);
#pragma unroll
      for (int i = 0; i < kHRun; i++) hb[hly][hlx0 + i] = (float)o[i];
    }
    __syncthreads();
    {
      double d[kVRun + 8], o[kVRun];
#pragma unroll
      for (int i = 0; i < kVRun + 8; i++) d[i] = (double)hb[ly0 + i][lx];
      fir_run<kVRun>(d, hf, o);
#pragma unroll
      for (int i = 0; i < kVRun; i++) vm[m][i] = (float)o[i];
    }
    __syncthreads();
  }
  float* mu2 = vm[0];
  float* s22 = vm[1];
  float* s12 = vm[2];
  const float C2 = 0.0009f;
  double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < kVRun; i++) {
    const int gy = y0 + ly0 + i;
    if (gx >= w || gy >= h) continue;
    const float i1 = t1[ly0 + i + kHalo][lx + kHalo], i2 = t2[ly0 + i + kHalo][lx + kHalo];
    const float mu1 = pmu[i], s11 = ps11[i];
    const float mu11 = mu1 * mu1, mu22 = mu2[i] * mu2[i], mu12 = mu1 * mu2[i];
    const float md = mu1 - mu2[i];
    const float num_m = 1.0f - md * md;
    const float num_s = 2.0f * (s12[i] - mu12) + C2;
    const float den_s = (s11 - mu11) + (s22[i] - mu22) + C2;
    double d = 1.0 - (double)((num_m * num_s) / den_s);
    d = d > 0.0 ? d : 0.0;
    acc[0] += d;
    acc[1] += d * d * d * d;
    double d1 = (1.0 + (double)fabsf(i2 - mu2[i])) / (1.0 + (double)fabsf(i1 - mu1)) - 1.0;
    double art = d1 > 0.0 ? d1 : 0.0;
    double lost = -d1 > 0.0 ? -d1 : 0.0;
    acc[2] += art;
    acc[3] += art * art * art * art;
    acc[4] += lost;
    acc[5] += lost * lost * lost * lost;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    double v = wave_sum(acc[k]);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double v = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
    out6[threadIdx.x] = v;
  }
}

__global__ __launch_bounds__(256) void k_blur_pool(const ConstTables* __restrict__ tab, PoolScales P,
                                                   double* __restrict__ partials) {
  const int c = blockIdx.z;
  const int tile = blockIdx.x;
  int sc = 0;
#pragma unroll
  for (int i = 1; i < kMaxPoolScales; i++)
    if (i < P.L.nscales && tile >= P.L.off[i]) sc = i;
  const int lt = tile - P.L.off[sc];
  blur_pool_tile(tab, P.w[sc], P.h[sc], (lt % P.tx[sc]) * kPW, (lt / P.tx[sc]) * kPH, P.x1[sc][c], P.mu1[sc][c],
                 P.s11[sc][c], P.x2[sc][c], partials + (size_t)tile * 18 + c * 6);
}

// ---------------------------------------------------------------------------
// Streaming form of the candidate pass: one wave walks a 64-column strip of
// one channel down a segment of kSH output rows.  Each input row is loaded
// once (plus a 4-column halo either side), widened to f64 once per sample
// into the wave's LDS row, filtered horizontally (one output column per
// lane) and pushed into a 9-row ring held in the lane's registers; the
// vertical filter of the row 4 behind runs from that ring, so the vertical
// halo costs 8 rows per segment instead of a recomputed band per tile.  The
// taps, pairing and roundings are those of blur_pool_tile, so the sums are
// the same per pixel.  The row loop is unrolled by the ring size so every
// ring index is a register.
// ---------------------------------------------------------------------------
struct StreamRing {
  double m[9][3];  // horizontally filtered moments (mu2, s22, s12) per ring row, f32 values
};

// The global samples one step of the walk reads: input row yi (a: column
// gx-4; b: the right halo on lanes 0..7) and, at column gx of output row
// yo = yi - 4, the reference's x1, mu1, s11 and the candidate's x2.
struct StreamIn {
  float a1, a2, b1, b2, mu1, s11, i1, i2;
};

__device__ __forceinline__ StreamIn stream_load(int t, int y0, int y1, int w, int h, int gx, int lane,
                                                const float* __restrict__ X1, const float* __restrict__ X2,
                                                const float* __restrict__ MU1,
                                                const float* __restrict__ S11) {
  const int yi = y0 - 4 + t, yo = yi - 4;
  const int xa = gx - 4, xb = gx + 60;
  // every address clamped into the plane, so the eight loads issue together
  // without branches; samples outside are then zeroed
  // (row offsets are wave-uniform: the row pointers stay in scalar registers
  // and each lane adds its fixed column)
  const size_t ri = (size_t)min(max(yi, 0), h - 1) * w, ro = (size_t)min(max(yo, 0), h - 1) * w;
  const uint32_t ca = min(max(xa, 0), w - 1), cbx = min(xb, w - 1), cg = min(gx, w - 1);
  StreamIn v;
  v.a1 = ldg32(X1 + ri, ca);
  v.a2 = ldg32(X2 + ri, ca);
  v.b1 = ldg32(X1 + ri, cbx);
  v.b2 = ldg32(X2 + ri, cbx);
  v.mu1 = ldg32(MU1 + ro, cg);
  v.s11 = ldg32(S11 + ro, cg);
  v.i1 = ldg32(X1 + ro, cg);
  v.i2 = ldg32(X2 + ro, cg);
  return v;
}

template <int K>
__device__ __forceinline__ void stream_row(int t, int y1, int y0, int w, int h, int gx, int lane, const StreamIn& in,
                                           const double* hf, double (*row)[72], StreamRing& R,
                                           double* acc) {
  const int yi = y0 - 4 + t, yo = yi - 4;
  const bool out = t >= 8 && yo < y1 && gx < w;
  // the loads read clamped addresses; samples outside the plane count as 0
  const bool rin = yi >= 0 && yi < h;
  const int xa = gx - 4, xb = gx + 60;
  const bool ina = rin && xa >= 0 && xa < w;
  const bool inb = rin && xb < w;
  const float a1 = ina ? in.a1 : 0.0f, a2 = ina ? in.a2 : 0.0f;
  // widened samples into the wave's LDS row: x2, x2^2, x1*x2 (the squared and
  // product inputs rounded to f32 first, blur.rs:97-113)
  row[0][lane] = (double)a2;
  row[1][lane] = (double)(a2 * a2);
  row[2][lane] = (double)(a1 * a2);
  if (lane < 8) {
    const float b1 = inb ? in.b1 : 0.0f, b2 = inb ? in.b2 : 0.0f;
    row[0][64 + lane] = (double)b2;
    row[1][64 + lane] = (double)(b2 * b2);
    row[2][64 + lane] = (double)(b1 * b2);
  }
  // the row's writes by every lane before any lane reads its neighbours
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int m = 0; m < 3; m++) {
    const double* v = row[m] + lane + 4;  // centre of this lane's window
    double acc_h = hf[0] * v[0];
#pragma unroll
    for (int k = 1; k <= 4; k++) acc_h = fma(hf[k], v[-k] + v[k], acc_h);
    R.m[K][m] = (double)(float)acc_h;
  }
  // every lane's reads of this row before the next row overwrites it
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (!out) return;
  // vertical filter over ring rows yo-4 .. yo+4: slots (K+1) .. K (mod 9)
  float vm[3];
#pragma unroll
  for (int m = 0; m < 3; m++) {
    double d[9];
#pragma unroll
    for (int k = 0; k < 9; k++) d[k] = R.m[(K + 1 + k) % 9][m];
    double a = hf[0] * d[4];
#pragma unroll
    for (int k = 1; k <= 4; k++) a = fma(hf[k], d[4 - k] + d[4 + k], a);
    vm[m] = (float)a;
  }
  const float mu1 = in.mu1, s11 = in.s11, i1 = in.i1, i2 = in.i2;
  const float mu2 = vm[0], s22 = vm[1], s12 = vm[2];
  const float C2 = 0.0009f;
  const float mu11 = mu1 * mu1, mu22 = mu2 * mu2, mu12 = mu1 * mu2;
  const float md = mu1 - mu2;
  const float num_m = 1.0f - md * md;
  const float num_s = 2.0f * (s12 - mu12) + C2;
  const float den_s = (s11 - mu11) + (s22 - mu22) + C2;
  double dd = 1.0 - (double)((num_m * num_s) / den_s);
  dd = dd > 0.0 ? dd : 0.0;
  acc[0] += dd;
  acc[1] += dd * dd * dd * dd;
  double d1 = (1.0 + (double)fabsf(i2 - mu2)) / (1.0 + (double)fabsf(i1 - mu1)) - 1.0;
  double art = d1 > 0.0 ? d1 : 0.0;
  double lost = -d1 > 0.0 ? -d1 : 0.0;
  acc[2] += art;
  acc[3] += art * art * art * art;
  acc[4] += lost;
  acc[5] += lost * lost * lost * lost;
}

// One wave: strip x0 .. x0+63 of channel planes, output rows y0 .. y1-1; the
// 6 sums of this (strip, segment, channel) go to out6 (lane 0).  The global
// samples of step t+1 are requested before step t is computed.
__device__ void stream_unit(const ConstTables* __restrict__ tab, int w, int h, int x0, int y0,
                            const float* __restrict__ X1, const float* __restrict__ MU1,
                            const float* __restrict__ S11, const float* __restrict__ X2,
                            double (*row)[72], double* __restrict__ out6) {
  const int lane = threadIdx.x & 63;
  const int gx = x0 + lane;
  const int y1 = min(y0 + kSH, h);
  double hf[5];
#pragma unroll
  for (int t = 0; t < 5; t++) hf[t] = tab->fir[t];
  StreamRing R;
#pragma unroll
  for (int k = 0; k < 9; k++) R.m[k][0] = R.m[k][1] = R.m[k][2] = 0.0;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int rows = (y1 - y0) + 8;  // input rows y0-4 .. y1+3 (rows >= 9)
  // three load buffers: step K computes from buf[K % 3] and requests step
  // K + 2 into buf[(K + 2) % 3] (the unroll of 9 keeps every index constant)
  StreamIn buf[3];
  buf[0] = stream_load(0, y0, y1, w, h, gx, lane, X1, X2, MU1, S11);
  buf[1] = stream_load(1, y0, y1, w, h, gx, lane, X1, X2, MU1, S11);
  for (int t = 0; t < rows; t += 9) {
// step t + K writes ring slot K (compile-time over the unroll of 9)
#define JPXLG_STEP(K)                                                                  \
    if (t + K < rows) {                                                                \
      buf[(K + 2) % 3] = stream_load(t + K + 2, y0, y1, w, h, gx, lane, X1, X2, MU1, S11); \
      stream_row<K>(t + K, y1, y0, w, h, gx, lane, buf[K % 3], hf, row, R, acc);                \
    }
    JPXLG_STEP(0) JPXLG_STEP(1) JPXLG_STEP(2) JPXLG_STEP(3) JPXLG_STEP(4)
    JPXLG_STEP(5) JPXLG_STEP(6) JPXLG_STEP(7) JPXLG_STEP(8)
#undef JPXLG_STEP
  }
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) out6[k] = v;
  }
}

// Batched streaming launch: blockDim.x / 64 (tile, channel) units per
// workgroup, one per wave; units enumerate (desc, scale) through off[] in
// units (tiles * 3).  With `swizzle`, the workgroups the dispatcher deals
// round-robin to the eight XCDs take contiguous runs of units, so the strips
// that share halo rows and columns meet in one XCD's L2.  The launch shape
// changes no sum.
__global__ __launch_bounds__(256) void k_blur_pool_stream(const ConstTables* __restrict__ tab,
                                                          const ProbeDesc* __restrict__ D,
                                                          const int* __restrict__ off, int nds, int total,
                                                          int swizzle) {
  __shared__ double rows[4][3][72];
  // the wave index is wave-uniform: say so, so that everything derived from
  // it (unit, plane pointers, row offsets) lives in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), upw = blockDim.x >> 6;
  const int nb = gridDim.x;
  int b = blockIdx.x;
  if (swizzle && nb >= 64) {
    // workgroup b runs on XCD b % 8: give that XCD the b / 8-th block of its
    // contiguous share (the last nb % 8 workgroups keep their own index)
    const int per = nb / 8;
    if (b < per * 8) b = (b % 8) * per + b / 8;
  }
  const int unit = b * upw + wave;
  if (unit >= total) return;
  const int ds = find_desc(off, nds, unit);
  const int d = ds / kDescScales, sc = ds % kDescScales;
  DescSpace& P = desc_at(D, d);
  const int w = P.sw[sc], h = P.sh[sc], tx = (w + kPW - 1) / kPW;
  const int tiles = P.punits[sc];
  const int lu = unit - off[ds], c = lu / tiles, lt = lu % tiles;
  stream_unit(tab, w, h, (lt % tx) * kPW, (lt / tx) * kSH, P.ref_x[sc][c], P.ref_mu[sc][c], P.ref_s11[sc][c],
              P.cand_x[sc][c], rows[wave], P.partials + (size_t)(P.ptile_off[sc] + lt) * 18 + c * 6);
}

// Fixed-order reduction of every scale's per-tile partial sums in one launch:
// workgroup (scale, term) sums that term over the scale's tiles.
__global__ __launch_bounds__(256) void k_reduce_scales(PoolLayout L, const double* __restrict__ partials,
                                                       double* __restrict__ sums) {
  __shared__ double red[4];
  const int s = blockIdx.x / 18, term = blockIdx.x % 18;
  const double* p = partials + (size_t)L.off[s] * 18;
  double acc = 0.0;
  for (int t = threadIdx.x; t < L.tiles[s]; t += 256) acc += p[(size_t)t * 18 + term];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) sums[s * 18 + term] = ((red[0] + red[1]) + red[2]) + red[3];
}

// Batched reduction: workgroup (desc, scale, term); scales past a desc's
// count do nothing.
__global__ __launch_bounds__(256) void k_reduce_b(const ProbeDesc* __restrict__ D) {
  __shared__ double red[4];
  const int d = blockIdx.x / (kDescScales * 18), rem = blockIdx.x % (kDescScales * 18);
  const int s = rem / 18, term = rem % 18;
  DescSpace& P = desc_at(D, d);
  if (s >= P.nscales) return;
  const int tiles = P.punits[s];
  const double* p = P.partials + (size_t)P.ptile_off[s] * 18;
  double acc = 0.0;
  for (int t = threadIdx.x; t < tiles; t += 256) acc += p[(size_t)t * 18 + term];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) P.sums[s * 18 + term] = ((red[0] + red[1]) + red[2]) + red[3];
}

}  // namespace

void launch_colour_metric(hipStream_t s, int mode, int w, int h, float* const in[3],
                          const float* thresholds, const float* depth_lut, int levels,
                          float* pxyb[3], float* down[3]) {
  const int w2 = (w + 1) / 2, h2 = (h + 1) / 2;
  dim3 grid(div_up(w2, 256), h2);
  float* d0 = down ? down[0] : nullptr;
  float* d1 = down ? down[1] : nullptr;
  float* d2 = down ? down[2] : nullptr;
  if (mode == 1)
    k_colour_metric<1><<<grid, 256, 0, s>>>(w, h, in[0], in[1], in[2], thresholds, depth_lut, levels, pxyb[0],
                                            pxyb[1], pxyb[2], d0, d1, d2);
  else
    k_colour_metric<0><<<grid, 256, 0, s>>>(w, h, in[0], in[1], in[2], thresholds, depth_lut, levels, pxyb[0],
                                            pxyb[1], pxyb[2], d0, d1, d2);
}

int colour_blocks(int w, int h) { return (int)(div_up((w + 1) / 2, 256) * ((h + 1) / 2)); }

void launch_colour_b(hipStream_t s, int scale, const ProbeDesc* D, const int* off, int nd, int nwg, int src) {
  if (nwg <= 0) return;
  if (scale == 0) k_colour_b<0><<<nwg, 256, 0, s>>>(D, off, nd, 0, src);
  else k_colour_b<1><<<nwg, 256, 0, s>>>(D, off, nd, scale, src);
}

int pyramid_tiles(int w, int h) { return ((w + kPyrTile - 1) / kPyrTile) * ((h + kPyrTile - 1) / kPyrTile); }

void launch_colour_pyr_b(hipStream_t s, const ProbeDesc* D, const int* off, int nd, int nwg, int src) {
  if (nwg <= 0) return;
  k_colour_pyr_b<<<nwg, 256, 0, s>>>(D, off, nd, src);
}

void launch_colour_pyr_ref(hipStream_t s, int ns, const int* sw, const int* sh, float* const in[3],
                           float* const (*out)[3]) {
  PyrPlanes A{};
  A.ns = ns;
  for (int i = 0; i < ns; i++) {
    A.sw[i] = sw[i];
    A.sh[i] = sh[i];
    for (int c = 0; c < 3; c++) A.out[i][c] = out[i][c];
  }
  for (int c = 0; c < 3; c++) A.in[c] = in[c];
  k_colour_pyr_ref<<<pyramid_tiles(sw[0], sh[0]), 256, 0, s>>>(A);
}

void launch_linear_at_depth(hipStream_t s, int w, int h, float* const in[3], const float* thresholds,
                            const float* depth_lut, int levels, float* out[3]) {
  int n = w * h;
  k_linear_at_depth<<<div_up(n, 256), 256, 0, s>>>(n, in[0], in[1], in[2], thresholds, depth_lut,
                                                   levels, out[0], out[1], out[2]);
}

void launch_blur_moments_all(hipStream_t s, const ConstTables* tab, int ns, const int* sw, const int* sh,
                             float* const (*x)[3], float* const (*mu)[3], float* const (*s11)[3]) {
  MomentPlanes A{};
  A.ns = ns;
  int total = 0;
  for (int i = 0; i < ns; i++) {
    A.sw[i] = sw[i];
    A.sh[i] = sh[i];
    A.off[i] = total;
    total += (int)(div_up(sw[i], kTW) * div_up(sh[i], kTH));
    for (int c = 0; c < 3; c++) {
      A.x[i][c] = x[i][c];
      A.mu[i][c] = mu[i][c];
      A.s11[i][c] = s11[i][c];
    }
  }
  A.off[ns] = total;
  k_blur_moments_all<<<dim3(total, 3), 256, 0, s>>>(tab, A);
}

void launch_blur_moments(hipStream_t s, int w, int h, const ConstTables* tab, float* const x1[3],
                         float* mu1[3], float* s11[3]) {
  dim3 grid(div_up(w, kTW), div_up(h, kTH));
  for (int c = 0; c < 3; c++) k_blur_moments<<<grid, 256, 0, s>>>(w, h, tab, x1[c], mu1[c], s11[c]);
}

int pool_tiles(int w, int h) { return pool_tiles_dev(w, h); }
int pool_units(int w, int h) { return pool_units_dev(w, h); }

void launch_blur_pool(hipStream_t s, const ConstTables* tab, PoolScales& P, double* partials) {
  int total = 0;
  for (int i = 0; i < P.L.nscales; i++) {
    P.tx[i] = (int)div_up(P.w[i], kPW);
    P.L.off[i] = total;
    P.L.tiles[i] = pool_tiles(P.w[i], P.h[i]);
    total += P.L.tiles[i];
  }
  dim3 grid(total, 1, 3);
  k_blur_pool<<<grid, 256, 0, s>>>(tab, P, partials);
}

void launch_reduce_scales(hipStream_t s, const PoolLayout& L, const double* partials, double* sums) {
  k_reduce_scales<<<18 * L.nscales, 256, 0, s>>>(L, partials, sums);
}

void launch_blur_pool_stream(hipStream_t s, const ConstTables* tab, const ProbeDesc* D, const int* off, int nds,
                             int units, int upw, int swizzle) {
  if (units <= 0) return;
  upw = upw == 1 || upw == 2 ? upw : 4;
  k_blur_pool_stream<<<div_up(units, upw), 64 * upw, 0, s>>>(tab, D, off, nds, units, swizzle);
}

void launch_reduce_b(hipStream_t s, const ProbeDesc* D, int nd) {
  if (nd <= 0) return;
  k_reduce_b<<<nd * kDescScales * 18, 256, 0, s>>>(D);
}

}  // namespace jpxlg
