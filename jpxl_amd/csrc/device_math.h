// Device arithmetic shared by the pixel and metric kernels.
//
// Every routine here performs the same IEEE-754 operations, in the same
// order, as the CPU code of JPXL it restates, so results are bit-identical
// to it.  The translation units are compiled with -ffp-contract=off: a
// multiply-add is fused only where the reference calls mul_add (fmaf here).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jpxlg {

// Quantizer arena limits (quantize.rs MAX_QUANT = 2^20).
constexpr int32_t kMaxQuant = 1 << 20;

// 1/sqrt(2) rounded to f32 (core::f32::consts::FRAC_1_SQRT_2).
#define JPXLG_R2 0.70710677f

struct DctTables {
  float cos_pi_8, sin_pi_8;
  float iv4[16];
  float iv8[64];
  float iv16[256];
};

// jpxl-core/src/simd.rs:36-67 reproducible_cbrt (musl estimate, two f64
// Halley steps, one rounding).
__device__ __forceinline__ float cbrt_core(float value) {
  const uint32_t B1 = 709958130u, B2 = 642849266u;
  uint32_t bits = __float_as_uint(value);
  uint32_t mag = bits & 0x7fffffffu;
  if (mag == 0 || mag >= 0x7f800000u) return value;
  uint32_t est;
  if (mag < 0x00800000u) {
    float scaled = value * 16777216.0f;
    est = (__float_as_uint(scaled) & 0x7fffffffu) / 3u + B2;
  } else {
    est = mag / 3u + B1;
  }
  double x = (double)value;
  double t = (double)__uint_as_float((bits & 0x80000000u) | est);
#pragma unroll
  for (int i = 0; i < 2; i++) {
    double r = (t * t) * t;
    t = (t * ((x + x) + r)) / ((x + r) + r);
  }
  return (float)t;
}

// The metric's cube root (jpxl-perceptual/src/color.rs:46-61): the reference
// rounds to f32 the f64 root its bit-hack estimate reaches after four Newton
// steps, i.e. the f64 cube root to within an ulp or two.  Here one Halley
// step in f64 from the f32 library root (relative error < 2^-22, cubed by the
// step to far below f64 precision) reaches the same f64 root to within a few
// ulps, so the two f32 results can differ only when the true root lies within
// a few f64 ulps of an f32 rounding boundary (about one sample in 2^27),
// and then by one f32 ulp.  It spends one f64 division where the
// reference's recursion spends four.
__device__ __forceinline__ float cbrt_metric_halley(double a, float y) {
  const double yd = (double)y;
  const double y3 = yd * yd * yd;
  return (float)(yd * ((y3 + 2.0 * a) / (2.0 * y3 + a)));
}

// The same root without the division in all but ~1e-5 of the samples: one
// Newton step from the library root with an f32 reciprocal puts t within
// ~2^-43 (relative) of the root when the library root was within a few f32
// ulps (checked through the size of the correction).  Rounding t to f32 is
// then certain -- the root, the reference's f64 root and the Halley root all
// round alike -- unless t lies within 2^11 f64 units of an f32 rounding
// boundary (the low 29 mantissa bits near 2^28); those samples take the
// Halley step, so the result equals the Halley form's everywhere.
__device__ __forceinline__ float cbrt_metric(float value) {
  if (value == 0.0f || !isfinite(value)) return value;
  const float av = fabsf(value);
  const float y = cbrtf(av);
  const double a = (double)av, yd = (double)y, y2 = yd * yd;
  const double d = (y2 * yd - a) * (double)__builtin_amdgcn_rcpf(3.0f * (float)y2);
  const double t = yd - d;
  const int32_t ext = (int32_t)((uint32_t)__double2loint(t) & 0x1fffffffu) - (1 << 28);
  float root;
  if (fabs(d) <= 0x1p-19 * t && (ext >= 2048 || ext <= -2048)) root = (float)t;
  else root = cbrt_metric_halley(a, y);
  return value < 0.0f ? -root : root;
}

// ---------------------------------------------------------------------------
// DCT butterflies (jpxl-core/src/dct.rs:199-426), registers only.
// ---------------------------------------------------------------------------

template <int M>
__device__ __forceinline__ void dct_iv(float* v, const DctTables& t) {
  if constexpr (M == 2) {
    float x0 = v[0], x1 = v[1];
    v[0] = t.cos_pi_8 * x0 + t.sin_pi_8 * x1;
    v[1] = t.sin_pi_8 * x0 - t.cos_pi_8 * x1;
  } else {
    const float* tab = M == 4 ? t.iv4 : (M == 8 ? t.iv8 : t.iv16);
    float x[M];
#pragma unroll
    for (int n = 0; n < M; n++) x[n] = v[n];
#pragma unroll
    for (int u = 0; u < M; u++) {
      float acc = 0.0f;
#pragma unroll
      for (int n = 0; n < M; n++) acc = acc + tab[u * M + n] * x[n];
      v[u] = acc;
    }
  }
}

__device__ __forceinline__ void dct_ii_2(float* v) {
  float x0 = v[0], x1 = v[1];
  v[0] = (x0 + x1) * JPXLG_R2;
  v[1] = (x0 - x1) * JPXLG_R2;
}

template <int N>
__device__ __forceinline__ void dct_ii(float* v, const DctTables& t) {
  if constexpr (N == 1) {
  } else if constexpr (N == 2) {
    dct_ii_2(v);
  } else if constexpr (N == 4) {
    float even[2] = {v[0] + v[3], v[1] + v[2]};
    float odd[2] = {v[0] - v[3], v[1] - v[2]};
    dct_ii_2(even);
    dct_iv<2>(odd, t);
    v[0] = even[0] * JPXLG_R2;
    v[1] = odd[0] * JPXLG_R2;
    v[2] = even[1] * JPXLG_R2;
    v[3] = odd[1] * JPXLG_R2;
  } else {
    constexpr int M = N / 2;
    float even[M], odd[M];
#pragma unroll
    for (int k = 0; k < M; k++) {
      even[k] = v[k] + v[N - 1 - k];
      odd[k] = v[k] - v[N - 1 - k];
    }
    dct_ii<M>(even, t);
    dct_iv<M>(odd, t);
#pragma unroll
    for (int k = 0; k < M; k++) {
      v[2 * k] = even[k] * JPXLG_R2;
      v[2 * k + 1] = odd[k] * JPXLG_R2;
    }
  }
}

template <int N>
__device__ __forceinline__ void dct_iii(float* v, const DctTables& t) {
  if constexpr (N == 1) {
  } else if constexpr (N == 2) {
    dct_ii_2(v);
  } else if constexpr (N == 4) {
    float even[2] = {v[0] * JPXLG_R2, v[2] * JPXLG_R2};
    float odd[2] = {v[1] * JPXLG_R2, v[3] * JPXLG_R2};
    dct_ii_2(even);
    dct_iv<2>(odd, t);
    v[0] = even[0] + odd[0];
    v[3] = even[0] - odd[0];
    v[1] = even[1] + odd[1];
    v[2] = even[1] - odd[1];
  } else {
    constexpr int M = N / 2;
    float even[M], odd[M];
#pragma unroll
    for (int k = 0; k < M; k++) {
      even[k] = v[2 * k] * JPXLG_R2;
      odd[k] = v[2 * k + 1] * JPXLG_R2;
    }
    dct_iii<M>(even, t);
    dct_iv<M>(odd, t);
#pragma unroll
    for (int k = 0; k < M; k++) {
      v[k] = even[k] + odd[k];
      v[N - 1 - k] = even[k] - odd[k];
    }
  }
}

// 1/sqrt(N) and sqrt(N) in f32, as dct.rs:789-795 forms them.
template <int N>
__device__ __forceinline__ float fwd_factor() {
  return 1.0f / __fsqrt_rn((float)N);
}
template <int N>
__device__ __forceinline__ float inv_factor() {
  return __fsqrt_rn((float)N);
}

// quantize.rs:1720-1746 clamp_round: ties away from zero through a +-0.5
// shift and truncation; sets *bad on a value outside +-2^20.
__device__ __forceinline__ int32_t clamp_round(float v, int* bad) {
  if (!isfinite(v)) { *bad = 1; return 0; }
  float adj = signbit(v) ? v - 0.5f : v + 0.5f;
  if (!(adj > -2147483648.0f && adj < 2147483648.0f)) { *bad = 1; return 0; }
  int32_t value = (int32_t)adj;
  if ((uint32_t)abs(value) > (uint32_t)kMaxQuant) { *bad = 1; return 0; }
  return value;
}

// reconstruct.rs:45-59 bias_adjust.
__device__ __forceinline__ float bias_adjust(int32_t q, float bias) {
  float f = (float)q;
  if (abs(q) <= 1) return f * bias;
  return f - 0.145f / f;
}

__device__ __forceinline__ int mirror1d(int c, int size) {
  while (true) {
    if (c < 0) c = -c - 1;
    else if (c >= size) c = 2 * size - 1 - c;
    else return c;
  }
}

// acc + p[0] + p[1] + ... + p[m-1], each addition rounded in that order (a
// sequential f64 sum), by one lane.  Loads are issued 32 at a time ahead of
// the additions that consume them, and the wave raises its issue priority
// for the chain so that the other waves sharing its SIMD do not stretch the
// dependent additions.
__device__ __forceinline__ double sequential_sum(const double* p, int m, double acc) {
  constexpr int kU = 32;
  __builtin_amdgcn_s_setprio(3);
  int i = 0;
  for (; i + kU <= m; i += kU) {
    double t[kU];
#pragma unroll
    for (int k = 0; k < kU; k++) t[k] = p[i + k];
#pragma unroll
    for (int k = 0; k < kU; k++) acc += t[k];
  }
  for (; i < m; i++) acc += p[i];
  __builtin_amdgcn_s_setprio(0);
  return acc;
}

}  // namespace jpxlg
