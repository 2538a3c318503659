// gfx950 kernels for the colour stage and SSIMULACRA2: XYB^-1 + depth
// quantizer, positive XYB, the 2:1 pyramid, the metric's blur (the
// reference's recursion, jpxl-perceptual/src/blur.rs, evaluated exactly in a
// parallel form; see the blur section) and the SSIM / artifact / detail-lost
// pooling with f64 partial sums.  Zero padding at the plane border is kept,
// exactly as the metric defines it.
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
// Blur (blur.rs:97-174, evaluated in a parallel form) and pooling
// (pool.rs:97-278)
// ----------------------------------------------------------------------------
// The reference blurs a line u[0..L) with a recursion fed, at step p, the f32
// pair sum s_p = f32(u[p-10] + u[p]) (zero outside the line) through three
// f64 resonators, and emits output n = p - 4 rounded to f32.  The
// resonators' summed response g to one sample is anti-periodic
// (g[j+10] = -g[j]) and over one half-period it is the symmetric 9-tap blur
// (g[4+t] = g[4-t] = fir[|t|]; g[9] ~ 1e-16), so the recursion's value before
// its rounding is
//     Y[p] = Psi[p] - Y[p-10],   Psi[p] = sum_{|t|<=4} fir[|t|] s[p-4+t],
// Psi in f64 from the f32 pair sums, whose rounding is the reference's own.
// A line is cut into segments of ten positions, one per lane: a lane holds
// the ten classes p mod 10 of its segment, so each stride-10 chain runs
// across consecutive lanes, and with sigma = (-1)^lane the chain is
// Y = sigma * (prefix sum over lanes of sigma * Psi).  A line of up to 636
// outputs takes one group of lanes (a power of two; several lines per wave).
// A longer one is walked in blocks of 64 lanes; a block's first two lanes
// repeat the previous block's last two segments (the pair sums and 9-tap
// window of its third lane) and its chain continues from the previous
// block's last lane.  This f64 evaluation and the reference's f64 resonators
// differ by a few f64 ulps: about one output in 10^6 rounds to the other f32
// neighbour.
//
// The horizontal pass walks the rows of a plane (16 waves per workgroup)
// and writes its outputs transposed through an LDS tile, so that the
// vertical pass walks contiguous lines again (column x at x * h).  The
// reference's vertical pass stores mu1 and s11 (transposed); the
// candidate's forms the SSIM and edge maps and one six-sum partial per wave.
// ----------------------------------------------------------------------------

constexpr int kSeg = 10;         // positions per lane
constexpr int kBlockStep = 620;  // positions a continued block adds (62 lanes)
constexpr int kHWaves = 16;      // horizontal pass: waves (1024 threads) per workgroup
constexpr int kTile = 10752;     // horizontal pass: LDS tile of R rows x (10 G + 1), R = 1024 / G

using SetSpace = const __attribute__((address_space(4))) LineSet;

struct LineShape {
  int G;       // lanes per line: a power of two <= 64
  int lpw;     // lines per wave (64 / G)
  int blocks;  // blocks per line (more than one only with G == 64)
};

__host__ __device__ inline LineShape line_shape(int L) {
  const int need = (L + 4 + kSeg - 1) / kSeg;  // segments of positions 0 .. L + 3
  LineShape s;
  if (need <= 64) {
    int G = 2;
    while (G < need) G <<= 1;
    s.G = G;
    s.lpw = 64 / G;
    s.blocks = 1;
  } else {
    s.G = 64;
    s.lpw = 1;
    s.blocks = 1 + (L + 4 - 640 + kBlockStep - 1) / kBlockStep;
  }
  return s;
}

struct LineTaps {
  double fir[5];
};

__device__ __forceinline__ LineTaps line_taps(const ConstTables* __restrict__ tab) {
  LineTaps T;
#pragma unroll
  for (int t = 0; t < 5; t++) T.fir[t] = tab->fir[t];
  return T;
}

// Lane i receives lane i-1's v; lane 0 receives 0.
__device__ __forceinline__ float from_prev_lane(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138, 0xf, 0xf, true));
}

// A DPP move of an f64 value.  Row shifts (BC): lanes without a source
// receive 0.  Row broadcasts: rows outside ROWS receive 0.
template <int CTRL, int ROWS, bool BC>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int rlo = __builtin_amdgcn_update_dpp(BC ? lo : 0, lo, CTRL, ROWS, 0xf, BC);
  const int rhi = __builtin_amdgcn_update_dpp(BC ? hi : 0, hi, CTRL, ROWS, 0xf, BC);
  return __hiloint2double(rhi, rlo);
}

// One step for the ten values, five at a time (the moves of five values
// cover each other's DPP wait states).
template <int CTRL, int ROWS, bool BC>
__device__ __forceinline__ void prefix_step(double (&v)[kSeg]) {
#pragma unroll
  for (int h = 0; h < kSeg; h += 5) {
    double t[5];
#pragma unroll
    for (int j = 0; j < 5; j++) t[j] = dpp_f64<CTRL, ROWS, BC>(v[h + j]);
#pragma unroll
    for (int j = 0; j < 5; j++) v[h + j] += t[j];
  }
}

template <int CTRL, int D>
__device__ __forceinline__ void prefix_step_masked(double (&v)[kSeg], int gl) {
#pragma unroll
  for (int h = 0; h < kSeg; h += 5) {
    double t[5];
#pragma unroll
    for (int j = 0; j < 5; j++) t[j] = dpp_f64<CTRL, 0xf, true>(v[h + j]);
#pragma unroll
    for (int j = 0; j < 5; j++) v[h + j] += gl >= D ? t[j] : 0.0;
  }
}

// Inclusive prefix sums of the ten values over aligned groups of G lanes.
// G >= 16: row shifts by 1, 2, 4, 8 within each row of 16 lanes, then row
// 0's last lane into row 1 and row 2's into row 3, then lane 31 into rows 2
// and 3.  G < 16: row shifts masked at the group's start.
template <int G>
__device__ __forceinline__ void group_prefix(double (&v)[kSeg], int gl) {
  if constexpr (G >= 16) {
    prefix_step<0x111, 0xf, true>(v);
    prefix_step<0x112, 0xf, true>(v);
    prefix_step<0x114, 0xf, true>(v);
    prefix_step<0x118, 0xf, true>(v);
    if constexpr (G >= 32) prefix_step<0x142, 0xa, false>(v);
    if constexpr (G == 64) prefix_step<0x143, 0xc, false>(v);
  } else {
    prefix_step_masked<0x111, 1>(v, gl);
    if constexpr (G > 2) prefix_step_masked<0x112, 2>(v, gl);
    if constexpr (G > 4) prefix_step_masked<0x114, 4>(v, gl);
  }
}

// Lane `src`'s v, in every lane.
__device__ __forceinline__ double lane_value(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// Barrier of the workgroup that orders its LDS accesses only (global stores
// stay in flight).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The M moment values of an input pair (the reference's x1, x1^2 or the
// candidate's x2, x2^2, x1 x2; products rounded to f32 first, blur.rs:97-113).
template <int M>
__device__ __forceinline__ float moment_value(int m, float a, float b) {
  if constexpr (M == 2) return m == 0 ? a : a * a;
  else return m == 0 ? b : (m == 1 ? b * b : a * b);
}

// One block of one line and moment: from the lane's samples u[j] at
// positions P + j (P = block start + 10 gl) the recursion's values Y[j] at the
// same positions, before their f32 rounding, for a chain that starts at zero
// at lane `warm` (the first lanes of a continued block only seed their
// neighbours).  gl: lane within the line's group of G lanes.  A continued
// block's true values are Y[j] - (-1)^gl C[j], C[j] being the previous
// block's last-lane value of class j (a linear chain: its start only adds
// that alternating term).
template <int G, bool MULTI>
__device__ __forceinline__ void seg_block(const float (&u)[kSeg], int gl, int warm, const LineTaps& T,
                                          double (&Y)[kSeg]) {
  // a group start receives zeros (the line starts there); with G == 64 the
  // only group start is lane 0, which the shift already fills with 0
  const bool zero_in = G < 64 && gl == 0;
  float s[kSeg];
#pragma unroll
  for (int j = 0; j < kSeg; j++) {
    const float up = from_prev_lane(u[j]);  // u[P + j - 10]
    s[j] = (zero_in ? 0.0f : up) + u[j];
  }
  // sd[i + 9]: s at P + i for i in [-8, 9] (the previous lane's s[2..9] below P)
  double sd[2 * kSeg - 1];
  sd[0] = 0.0;
#pragma unroll
  for (int j = 2; j < kSeg; j++) {
    const float sp = from_prev_lane(s[j]);
    sd[j - 1] = (double)(zero_in ? 0.0f : sp);
  }
#pragma unroll
  for (int j = 0; j < kSeg; j++) sd[9 + j] = (double)s[j];
  // sigma = (-1)^gl as a sign bit on the high word
  const int sign = (gl & 1) << 31;
#pragma unroll
  for (int j = 0; j < kSeg; j++) {
    const double* c = &sd[9 + j - 4];
    double psi = T.fir[0] * c[0];
#pragma unroll
    for (int t = 1; t <= 4; t++) psi = fma(T.fir[t], c[-t] + c[t], psi);
    Y[j] = __hiloint2double(__double2hiint(psi) ^ sign, __double2loint(psi));
  }
  if constexpr (MULTI) {
#pragma unroll
    for (int j = 0; j < kSeg; j++) Y[j] = gl < warm ? 0.0 : Y[j];
  }
  group_prefix<G>(Y, gl);
#pragma unroll
  for (int j = 0; j < kSeg; j++) Y[j] = __hiloint2double(__double2hiint(Y[j]) ^ sign, __double2loint(Y[j]));
}

// Y - (-1)^gl c, the continued chain's values.
__device__ __forceinline__ double continue_chain(double y, double c, int gl) {
  return (gl & 1) ? y + c : y - c;
}

// Horizontal pass: workgroup (set, band of R = 1024 / G rows), each wave
// 64 / G rows, a long row's blocks one after another (the carry is the
// previous block's last lane, read from the wave's LDS slot).  The outputs
// (and the first moment's source plane) go out transposed.
template <int CAND, int G, bool MULTI>
__device__ __forceinline__ void blur_h_band(SetSpace& S, int band, const LineTaps& T, float* tile, double* carry) {
  constexpr int M = CAND ? 3 : 2;
  constexpr int LPW = 64 / G, R = kHWaves * LPW, STRIDE = kSeg * G + 1, NT = 64 * kHWaves;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int gl = lane & (G - 1), grp = lane / G;
  const int w = S.w, h = S.h, pitch = S.pitch;
  const int blocks = MULTI ? line_shape(w).blocks : 1;
  const int y0 = band * R;
  const int rl = wv * LPW + grp;  // the group's row in the band
  // rows past the plane repeat its last row; their outputs are not stored
  const int y = min(y0 + rl, h - 1);
  const float* r1 = S.x1 + (size_t)y * w;
  const float* r2 = CAND ? S.x2 + (size_t)y * w : r1;
  // this wave's carries per moment (MULTI: one row per wave)
  double* C = carry + wv * 3 * kSeg;
  if (MULTI && lane < M * kSeg) C[lane] = 0.0;
  // the lane's samples of block b: positions 620 b + 10 gl + j
  auto fetch = [&](int b, float (&a)[kSeg], float (&bv)[kSeg]) {
#pragma unroll
    for (int j = 0; j < kSeg; j++) {
      const int p = kBlockStep * b + kSeg * gl + j;
      const bool in = p < w;
      const uint32_t pc = min(p, w - 1);
      a[j] = in ? ldg32(r1, pc) : 0.0f;
      bv[j] = in ? ldg32(r2, pc) : 0.0f;
    }
  };
  float a[kSeg], bv[kSeg];
  fetch(0, a, bv);
  for (int b = 0; b < blocks; b++) {
    const int pb = kBlockStep * b, warm = b ? 2 : 0;
    // stage plane m (a moment's blur, or with m == M the source plane) in
    // tile half q
    auto stage = [&](auto mc, int q) {
      constexpr int m = decltype(mc)::value;
      float out[kSeg];
      if constexpr (m < M) {
        float u[kSeg];
#pragma unroll
        for (int j = 0; j < kSeg; j++) u[j] = moment_value<M>(m, a[j], bv[j]);
        double Y[kSeg];
        seg_block<G, MULTI>(u, gl, warm, T, Y);
        if constexpr (MULTI) {
#pragma unroll
          for (int j = 0; j < kSeg; j++) Y[j] = continue_chain(Y[j], C[m * kSeg + j], gl);
          // the next block continues from this block's last lane
          if (gl == 63)
#pragma unroll
            for (int j = 0; j < kSeg; j++) C[m * kSeg + j] = Y[j];
        }
#pragma unroll
        for (int j = 0; j < kSeg; j++) out[j] = (float)Y[j];
      } else {
#pragma unroll
        for (int j = 0; j < kSeg; j++) out[j] = CAND ? bv[j] : a[j];
      }
      float* t = tile + q * kTile;
#pragma unroll
      for (int j = 0; j < kSeg; j++) t[rl * STRIDE + kSeg * gl + j] = out[j];
    };
    // write plane m from tile half q: tile column col holds position
    // pb + col, i.e. output n = pb + col - 4 (or source sample n = pb + col);
    // R consecutive rows per column
    auto flush = [&](int m, int q) {
      float* dst = m < M ? S.hT[m] : S.xT;
      const float* t = tile + q * kTile;
      const int nb = m < M ? pb - 4 : pb;
#pragma unroll
      for (int it = 0; it < kSeg; it++) {
        const int e = it * NT + (int)threadIdx.x;
        const int col = e / R, row = e % R;
        const int n = nb + col, yy = y0 + row;
        if (col >= kSeg * warm && n >= 0 && n < w && yy < h)
          stg32(dst, (uint32_t)n * (uint32_t)pitch + (uint32_t)yy, t[row * STRIDE + col]);
      }
    };
    // one plane per round: the moments' blurs, then the source plane (the
    // next block's samples are requested before the last round goes out)
    stage(std::integral_constant<int, 0>{}, 0);
    lds_barrier();
    flush(0, 0);
    lds_barrier();
    stage(std::integral_constant<int, 1>{}, 0);
    lds_barrier();
    flush(1, 0);
    lds_barrier();
    stage(std::integral_constant<int, 2>{}, 0);
    lds_barrier();
    if constexpr (M == 2)
      if (MULTI && b + 1 < blocks) fetch(b + 1, a, bv);
    flush(2, 0);
    lds_barrier();
    if constexpr (M == 3) {
      stage(std::integral_constant<int, 3>{}, 0);
      lds_barrier();
      if (MULTI && b + 1 < blocks) fetch(b + 1, a, bv);
      flush(3, 0);
      lds_barrier();
    }
  }
}

template <int CAND>
__global__ __launch_bounds__(64 * kHWaves) void k_blur_h(const ConstTables* __restrict__ tab, const LineSet* __restrict__ sets,
                                                 const int* __restrict__ off, int nsets) {
  __shared__ float tile[kTile];
  __shared__ double carry[kHWaves * 3 * kSeg];
  const int k = find_desc(off, nsets, blockIdx.x);
  SetSpace& S = ((SetSpace*)sets)[k];
  const int band = blockIdx.x - off[k];
  const LineTaps T = line_taps(tab);
  const LineShape sh = line_shape(S.w);
  switch (sh.G) {
    case 64:
      if (sh.blocks > 1) blur_h_band<CAND, 64, true>(S, band, T, tile, carry);
      else blur_h_band<CAND, 64, false>(S, band, T, tile, carry);
      break;
    case 32: blur_h_band<CAND, 32, false>(S, band, T, tile, carry); break;
    case 16: blur_h_band<CAND, 16, false>(S, band, T, tile, carry); break;
    case 8: blur_h_band<CAND, 8, false>(S, band, T, tile, carry); break;
    case 4: blur_h_band<CAND, 4, false>(S, band, T, tile, carry); break;
    default: blur_h_band<CAND, 2, false>(S, band, T, tile, carry); break;
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// The SSIM and edge terms of one pixel (pool.rs:97-278) into acc.  The
// SSIM ratio is the reference's f32 division; the edge ratio (an f64
// division in the reference) is a reciprocal refined by two Newton steps and
// a corrected quotient, within an ulp of the quotient (it only enters f64
// sums).
__device__ __forceinline__ double edge_ratio(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(r, fma(-b, r, 1.0), r);
  r = fma(r, fma(-b, r, 1.0), r);
  const double q = a * r;
  return fma(r, fma(-b, q, a), q);
}

__device__ __forceinline__ void pool_pixel(float mu1, float s11, float i1, float i2, float mu2, float s22, float s12,
                                           double (&acc)[6]) {
  const float C2 = 0.0009f;
  const float mu11 = mu1 * mu1, mu22 = mu2 * mu2, mu12 = mu1 * mu2;
  const float md = mu1 - mu2;
  const float num_m = 1.0f - md * md;
  const float num_s = 2.0f * (s12 - mu12) + C2;
  const float den_s = (s11 - mu11) + (s22 - mu22) + C2;
  double d = 1.0 - (double)((num_m * num_s) / den_s);
  d = d > 0.0 ? d : 0.0;
  acc[0] += d;
  const double d2 = d * d;
  acc[1] += d2 * d2;
  const double d1 = edge_ratio(1.0 + (double)fabsf(i2 - mu2), 1.0 + (double)fabsf(i1 - mu1)) - 1.0;
  const double art = d1 > 0.0 ? d1 : 0.0;
  const double lost = -d1 > 0.0 ? -d1 : 0.0;
  const double a2 = art * art, l2 = lost * lost;
  acc[2] += art;
  acc[3] += a2 * a2;
  acc[4] += lost;
  acc[5] += l2 * l2;
}

// Pairs of samples of a transposed line: v[j] at position p0 + j (p0 even),
// zero before 0 and at and past h; the line's allocation ends at its even
// pitch.
__device__ __forceinline__ void load_line10(const float* __restrict__ line, int p0, int h, float (&v)[kSeg]) {
#pragma unroll
  for (int k = 0; k < kSeg; k += 2) {
    const int p = p0 + k;
    uint64_t t = 0;
    if (p >= 0 && p < h) t = ldg32((const uint64_t*)line, (uint32_t)(p >> 1));
    v[k] = __uint_as_float((uint32_t)t);
    v[k + 1] = p + 1 < h ? __uint_as_float((uint32_t)(t >> 32)) : 0.0f;
  }
}

// Vertical pass: workgroups of four waves.  A short line (one block) takes
// part of a wave (four waves: 4 * 64 / G lines); a long line takes a whole
// workgroup, its blocks four at a time, one per wave: each block's chain
// runs from zero, its last lane's values go to LDS, and after a barrier each
// block continues from the sum of the earlier blocks' last values (the
// running carry of the rounds before plus this round's earlier blocks); the
// chains' f64 values wait in registers for their carries.  Every wave leaves one
// partial row of six sums (zero when it pools nothing).  With `swizzle`, the
// workgroups the dispatcher deals round-robin to the eight XCDs take
// contiguous runs.  The launch shape changes no output.
constexpr int kVWaves = 4;

struct VShared {
  float ys[kVWaves][3 * kSeg * 64];   // rounded outputs per wave: [moment][j][lane]
  double ends[kVWaves][3 * kSeg];     // each round's blocks' last-lane values
  double carry[3 * kSeg];             // the running carry entering the round
};

// Outputs of one block held in ys (moment m, class j at ys[(m * 10 + j) * 64 +
// lane]) for positions n0 + j: the candidate's SSIM and edge sums, or the
// reference's mu1 and s11.
template <int CAND>
__device__ __forceinline__ void v_outputs(SetSpace& S, size_t base, int h, int n0, bool own, const float* ys,
                                          int lane, double (&acc)[6]) {
  if constexpr (CAND != 0) {
    // the reference's mu1, s11, x1 and the candidate's x2, a pair of
    // positions at a time
    const float* planes[4] = {S.mu1T + base, S.s11T + base, S.x1T + base, S.x2T + base};
#pragma unroll 3
    for (int k = 0; k < kSeg; k += 2) {
      const int n = n0 + k;
      if (!own || n < 0 || n >= h) continue;
      float q[4][2];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint64_t t = ldg32((const uint64_t*)planes[i], (uint32_t)(n >> 1));
        q[i][0] = __uint_as_float((uint32_t)t);
        q[i][1] = __uint_as_float((uint32_t)(t >> 32));
      }
      const float* y0 = ys + k * 64 + lane;
      pool_pixel(q[0][0], q[1][0], q[2][0], q[3][0], y0[0], y0[kSeg * 64], y0[2 * kSeg * 64], acc);
      if (n + 1 < h)
        pool_pixel(q[0][1], q[1][1], q[2][1], q[3][1], y0[64], y0[kSeg * 64 + 64], y0[2 * kSeg * 64 + 64], acc);
    }
  } else {
#pragma unroll
    for (int j = 0; j < kSeg; j++) {
      const int n = n0 + j;
      if (!own || n < 0 || n >= h) continue;
      stg32(S.out[0] + base, (uint32_t)n, ys[j * 64 + lane]);
      stg32(S.out[1] + base, (uint32_t)n, ys[(kSeg + j) * 64 + lane]);
    }
  }
}

// Short lines: wave wv of the workgroup's lw-th group of 4 * 64 / G lines.
template <int CAND, int G>
__device__ __forceinline__ void blur_v_short(SetSpace& S, int lw, const LineTaps& T, VShared& sh, double (&acc)[6]) {
  constexpr int M = CAND ? 3 : 2;
  constexpr int LPW = 64 / G;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gl = lane & (G - 1), grp = lane / G;
  const int w = S.w, h = S.h;
  const int xr = (lw * kVWaves + wv) * LPW + grp;
  const bool line_ok = xr < w;
  const size_t base = (size_t)min(xr, w - 1) * S.pitch;
  float* ys = sh.ys[wv];
  float u[M][kSeg];
#pragma unroll
  for (int m = 0; m < M; m++) load_line10(S.hT[m] + base, kSeg * gl, h, u[m]);
#pragma unroll
  for (int m = 0; m < M; m++) {
    double Y[kSeg];
    seg_block<G, false>(u[m], gl, 0, T, Y);
#pragma unroll
    for (int j = 0; j < kSeg; j++) ys[(m * kSeg + j) * 64 + lane] = (float)Y[j];
  }
  v_outputs<CAND>(S, base, h, kSeg * gl - 4, line_ok, ys, lane, acc);
}

// A long line x = lw: blocks in rounds of four, one per wave.
template <int CAND>
__device__ __forceinline__ void blur_v_long(SetSpace& S, int lw, const LineTaps& T, VShared& sh, double (&acc)[6]) {
  constexpr int M = CAND ? 3 : 2;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int h = S.h;
  const int blocks = line_shape(h).blocks;
  const size_t base = (size_t)lw * S.pitch;
  float* ys = sh.ys[wv];
  if (threadIdx.x < M * kSeg) sh.carry[threadIdx.x] = 0.0;
  for (int r = 0; r * kVWaves < blocks; r++) {
    const int b = r * kVWaves + wv;
    const bool active = b < blocks;  // wave-uniform
    const int pb = kBlockStep * b, warm = b ? 2 : 0;
    float u[M][kSeg];
    if (active)
#pragma unroll
      for (int m = 0; m < M; m++) load_line10(S.hT[m] + base, pb + kSeg * lane, h, u[m]);
    // one moment at a time, so only one moment's chain values wait (in
    // registers) for their carry across the barrier
#pragma unroll
    for (int m = 0; m < M; m++) {
      double Y[kSeg];
      if (active) {
        seg_block<64, true>(u[m], lane, warm, T, Y);
        if (lane == 63)
#pragma unroll
          for (int j = 0; j < kSeg; j++) sh.ends[wv][m * kSeg + j] = Y[j];
      }
      lds_barrier();
      if (active)
        // this block's carry: the running carry plus this round's earlier ends
#pragma unroll
        for (int j = 0; j < kSeg; j++) {
          double c = sh.carry[m * kSeg + j];
          for (int k = 0; k < wv; k++) c += sh.ends[k][m * kSeg + j];
          ys[(m * kSeg + j) * 64 + lane] = (float)continue_chain(Y[j], c, lane);
        }
    }
    if (active) v_outputs<CAND>(S, base, h, pb + kSeg * lane - 4, lane >= warm, ys, lane, acc);
    lds_barrier();
    // the carry entering the next round (the blocks' ends in order)
    if (threadIdx.x < M * kSeg) {
      double c = sh.carry[threadIdx.x];
      for (int k = 0; k < kVWaves && r * kVWaves + k < blocks; k++) c += sh.ends[k][threadIdx.x];
      sh.carry[threadIdx.x] = c;
    }
    lds_barrier();
  }
}

template <int CAND>
__global__ __launch_bounds__(64 * kVWaves) __attribute__((amdgpu_waves_per_eu(5))) void k_blur_v(const ConstTables* __restrict__ tab,
                                                         const LineSet* __restrict__ sets,
                                                         const int* __restrict__ off, int nsets, int total,
                                                         int swizzle) {
  __shared__ VShared sh;
  const int nb = gridDim.x;
  int bi = blockIdx.x;
  if (swizzle && nb >= 64) {
    const int per = nb / 8;
    if (bi < per * 8) bi = (bi % 8) * per + bi / 8;
  }
  if (bi >= total) return;
  const int k = find_desc(off, nsets, bi);
  SetSpace& S = ((SetSpace*)sets)[k];
  const int lw = bi - off[k];
  const LineTaps T = line_taps(tab);
  const LineShape shp = line_shape(S.h);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  switch (shp.G) {
    case 64:
      if (shp.blocks > 1) blur_v_long<CAND>(S, lw, T, sh, acc);
      else blur_v_short<CAND, 64>(S, lw, T, sh, acc);
      break;
    case 32: blur_v_short<CAND, 32>(S, lw, T, sh, acc); break;
    case 16: blur_v_short<CAND, 16>(S, lw, T, sh, acc); break;
    case 8: blur_v_short<CAND, 8>(S, lw, T, sh, acc); break;
    case 4: blur_v_short<CAND, 4>(S, lw, T, sh, acc); break;
    default: blur_v_short<CAND, 2>(S, lw, T, sh, acc); break;
  }
  if constexpr (CAND != 0) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int t = 0; t < 6; t++) {
      const double v = wave_sum(acc[t]);
      if (lane == 0) S.partials[(size_t)(lw * kVWaves + wv) * 18 + t] = v;
    }
  }
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

int blur_bands(int w, int h) { return (int)div_up(h, kHWaves * line_shape(w).lpw); }
// Workgroups of the vertical pass over a w x h plane.
int blur_columns(int w, int h) {
  const LineShape s = line_shape(h);
  return s.blocks > 1 ? w : (int)div_up(w, kVWaves * s.lpw);
}
int pool_units(int w, int h) { return kVWaves * blur_columns(w, h); }

void launch_blur_h(hipStream_t s, int cand, const ConstTables* tab, const LineSet* sets, const int* off, int nsets,
                   int bands) {
  if (bands <= 0) return;
  if (cand) k_blur_h<1><<<bands, 64 * kHWaves, 0, s>>>(tab, sets, off, nsets);
  else k_blur_h<0><<<bands, 64 * kHWaves, 0, s>>>(tab, sets, off, nsets);
}

void launch_blur_v(hipStream_t s, int cand, const ConstTables* tab, const LineSet* sets, const int* off, int nsets,
                   int groups, int swizzle) {
  if (groups <= 0) return;
  if (cand) k_blur_v<1><<<groups, 64 * kVWaves, 0, s>>>(tab, sets, off, nsets, groups, swizzle);
  else k_blur_v<0><<<groups, 64 * kVWaves, 0, s>>>(tab, sets, off, nsets, groups, swizzle);
}

void launch_reduce_scales(hipStream_t s, const PoolLayout& L, const double* partials, double* sums) {
  k_reduce_scales<<<18 * L.nscales, 256, 0, s>>>(L, partials, sums);
}

void launch_reduce_b(hipStream_t s, const ProbeDesc* D, int nd) {
  if (nd <= 0) return;
  k_reduce_b<<<nd * kDescScales * 18, 256, 0, s>>>(D);
}

}  // namespace jpxlg
