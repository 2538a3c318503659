// Matrix forms of the square DCT_2D / IDCT_2D on the gfx950 matrix cores.
//
// The reference transform (jpxl-core/src/dct.rs:956-997) is, in exact
// arithmetic, Out = L . In^T . L^T with
//   forward  L = D / sqrt(N),   In = samples  X[y][x], Out = cells C[k][r],
//   inverse  L = sqrt(N) D^T,   In = cells    C[k][r], Out = samples X[y][x],
// where D is the orthonormal DCT-II basis (D[k][x] = c_k cos(pi (2x+1) k / 2N)).
// The butterfly kernels reproduce the reference's rounding bit for bit; these
// contractions do not, and exist for the precision study: the plain form
// multiplies in FP32 on the matrix cores (v_mfma_f32_*f32), the chosen form
// splits every operand into bf16 hi + lo parts and accumulates
// hi*hi + hi*lo + lo*hi in FP32 (v_mfma_f32_*bf16).
//
// One wave transforms one 32x32 or 16x16 matrix, or four 8x8 matrices packed
// into one 16x16 product with L8 on the diagonal of a 16x16 L.  The input is
// read from LDS tiles (row stride P) and the intermediate Z = In^T . L^T is
// written back over the same tiles, so a wave owns its tiles for the call.
#pragma once
#include <hip/hip_runtime.h>

namespace jpxlg {

enum DctForm : int { kDctExact = 0, kDctFp32 = 1, kDctBf16x3 = 2 };

// L matrices per size (row-major N x N, f32), forward and inverse.
struct DctMats {
  float fwd8[64], fwd16[256], fwd32[1024];
  float inv8[64], inv16[256], inv32[1024];
};

typedef __bf16 jbf16x8 __attribute__((ext_vector_type(8)));
typedef short js16x4 __attribute__((ext_vector_type(4)));
typedef float jf32x4 __attribute__((ext_vector_type(4)));
typedef float jf32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void bf16_split(float v, __bf16* hi, __bf16* lo) {
  const __bf16 h = (__bf16)v;
  *hi = h;
  *lo = (__bf16)(v - (float)h);
}

__device__ __forceinline__ short bf16_bits(__bf16 v) { return __builtin_bit_cast(short, v); }

// Element (r, c) of the 16x16 packed matrix built from four 8x8 tiles
// (tile 2*(r/8) + c/8 holds rows r%8, columns c%8).
__device__ __forceinline__ float* packed16(float* const t[4], int P, int r, int c) {
  return t[2 * (r >> 3) + (c >> 3)] + (r & 7) * P + (c & 7);
}

// ---------------------------------------------------------------------------
// 32 x 32 (one matrix per wave)
// ---------------------------------------------------------------------------
template <int FORM, class Put>
__device__ __forceinline__ void wave_transform32(float* tile, int P, const float* __restrict__ L, Put put) {
  const int l = threadIdx.x & 63;
  const int m = l & 31, hb = l >> 5;
  jf32x16 acc;
  if constexpr (FORM == kDctBf16x3) {
    // Operand fragments of L (identical for the B operand of Z = In^T L^T
    // and the A operand of Out = L Z): L[m][16 s + 8 hb + i].
    jbf16x8 lh[2], ll[2];
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
      for (int i = 0; i < 8; i++) {
        __bf16 h, lo;
        bf16_split(L[m * 32 + 16 * s + 8 * hb + i], &h, &lo);
        lh[s][i] = h;
        ll[s][i] = lo;
      }
    // G1: Z = In^T . L^T, A[m][k] = In[k][m].
    jbf16x8 ah[2], al[2];
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
      for (int i = 0; i < 8; i++) {
        __bf16 h, lo;
        bf16_split(tile[(16 * s + 8 * hb + i) * P + m], &h, &lo);
        ah[s][i] = h;
        al[s][i] = lo;
      }
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.0f;
#pragma unroll
    for (int s = 0; s < 2; s++) {
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[s], lh[s], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[s], ll[s], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[s], lh[s], acc, 0, 0, 0);
    }
    // Z over the input tile: lane holds Z[8(i/4) + 4 hb + i%4][m].
#pragma unroll
    for (int i = 0; i < 16; i++) tile[(8 * (i >> 2) + 4 * hb + (i & 3)) * P + m] = acc[i];
    // G2: Out = L . Z, B[k][n] = Z[k][n].
    jbf16x8 bh[2], bl[2];
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
      for (int i = 0; i < 8; i++) {
        __bf16 h, lo;
        bf16_split(tile[(16 * s + 8 * hb + i) * P + m], &h, &lo);
        bh[s][i] = h;
        bl[s][i] = lo;
      }
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.0f;
#pragma unroll
    for (int s = 0; s < 2; s++) {
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ll[s], bh[s], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lh[s], bl[s], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lh[s], bh[s], acc, 0, 0, 0);
    }
  } else {
    // FP32: v_mfma_f32_32x32x2f32, K = 2 per step, lane k index = hb.
    float lf[16];
#pragma unroll
    for (int s = 0; s < 16; s++) lf[s] = L[m * 32 + 2 * s + hb];
    float a[16];
#pragma unroll
    for (int s = 0; s < 16; s++) a[s] = tile[(2 * s + hb) * P + m];
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.0f;
#pragma unroll
    for (int s = 0; s < 16; s++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[s], lf[s], acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; i++) tile[(8 * (i >> 2) + 4 * hb + (i & 3)) * P + m] = acc[i];
#pragma unroll
    for (int s = 0; s < 16; s++) a[s] = tile[(2 * s + hb) * P + m];
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.0f;
#pragma unroll
    for (int s = 0; s < 16; s++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(lf[s], a[s], acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 16; i++) put(8 * (i >> 2) + 4 * hb + (i & 3), m, acc[i]);
}

// ---------------------------------------------------------------------------
// 16 x 16: one DCT16 matrix, or four DCT8 matrices on the diagonal of L
// ---------------------------------------------------------------------------
// Access to the 16x16 operand: either one tile or four packed 8x8 tiles.
template <bool kPacked>
__device__ __forceinline__ float* elem16(float* const t[4], int P, int r, int c) {
  if constexpr (kPacked) return packed16(t, P, r, c);
  else return t[0] + r * P + c;
}

template <int N, int FORM, class Put>
__device__ __forceinline__ void wave_transform16(float* const tiles[4], int P, const float* __restrict__ L, Put put) {
  constexpr bool kPacked = N == 8;
  const int l = threadIdx.x & 63;
  const int m = l & 15, q = l >> 4;
  // L16 element (r, c): the size's matrix, or L8 on the diagonal.
  auto lval = [&](int r, int c) -> float {
    if constexpr (kPacked) return (r >> 3) == (c >> 3) ? L[(r & 7) * 8 + (c & 7)] : 0.0f;
    else return L[r * 16 + c];
  };
  jf32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  if constexpr (FORM == kDctBf16x3) {
    // v_mfma_f32_16x16x16bf16_1k: A[m][4q + i], B[4q + i][m].
    js16x4 lh, ll, ah, al;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      __bf16 h, lo;
      bf16_split(lval(m, 4 * q + i), &h, &lo);
      lh[i] = bf16_bits(h);
      ll[i] = bf16_bits(lo);
      bf16_split(*elem16<kPacked>(tiles, P, 4 * q + i, m), &h, &lo);
      ah[i] = bf16_bits(h);
      al[i] = bf16_bits(lo);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, lh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, ll, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, lh, acc, 0, 0, 0);
    // Z over the input tiles: lane holds Z[4q + i][m].
#pragma unroll
    for (int i = 0; i < 4; i++) *elem16<kPacked>(tiles, P, 4 * q + i, m) = acc[i];
    js16x4 bh, bl;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      __bf16 h, lo;
      bf16_split(*elem16<kPacked>(tiles, P, 4 * q + i, m), &h, &lo);
      bh[i] = bf16_bits(h);
      bl[i] = bf16_bits(lo);
    }
    acc = jf32x4{0.0f, 0.0f, 0.0f, 0.0f};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ll, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lh, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lh, bh, acc, 0, 0, 0);
  } else {
    // v_mfma_f32_16x16x4f32: A[m][4 s + q], B[4 s + q][m], s = 0..3.
    float lf[4], a[4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      lf[s] = lval(m, 4 * s + q);
      a[s] = *elem16<kPacked>(tiles, P, 4 * s + q, m);
    }
#pragma unroll
    for (int s = 0; s < 4; s++) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], lf[s], acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++) *elem16<kPacked>(tiles, P, 4 * q + i, m) = acc[i];
#pragma unroll
    for (int s = 0; s < 4; s++) a[s] = *elem16<kPacked>(tiles, P, 4 * s + q, m);
    acc = jf32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 4; s++) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(lf[s], a[s], acc, 0, 0, 0);
  }
  // Out16[4q + i][m]; packed: block 2p + b16 with p = m/8 (column block)
  // and b16 = row/8, i.e. matrix (row % 8, m % 8) of tile 2 (m/8) + row/8.
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = 4 * q + i;
    if constexpr (kPacked) put(2 * (m >> 3) + (r >> 3), r & 7, m & 7, acc[i]);
    else put(0, r, m, acc[i]);
  }
}

// Out = L In^T L^T over `ntiles` N x N LDS tiles (tile t at base + t*N*P),
// shared by the workgroup's waves; put(t, row, col, v) receives every output.
template <int N, int FORM, class Put>
__device__ __forceinline__ void matrix_transform_tiles(float* base, int ntiles, const float* __restrict__ L,
                                                       Put put) {
  constexpr int P = N + 1;
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if constexpr (N == 32) {
    for (int t = wave; t < ntiles; t += nw)
      wave_transform32<FORM>(base + t * N * P, P, L, [&](int r, int c, float v) { put(t, r, c, v); });
  } else if constexpr (N == 16) {
    for (int t = wave; t < ntiles; t += nw) {
      float* tl[4] = {base + t * N * P, nullptr, nullptr, nullptr};
      wave_transform16<16, FORM>(tl, P, L, [&](int, int r, int c, float v) { put(t, r, c, v); });
    }
  } else {
    // Groups of four distinct 8x8 tiles (Z is stored over them in place, so
    // the four slots must not alias): ntiles is a multiple of 4.
    for (int g0 = 4 * wave; g0 + 4 <= ntiles; g0 += 4 * nw) {
      float* tl[4];
#pragma unroll
      for (int b = 0; b < 4; b++) tl[b] = base + (g0 + b) * N * P;
      wave_transform16<8, FORM>(tl, P, L, [&](int b, int r, int c, float v) { put(g0 + b, r, c, v); });
    }
  }
}

}  // namespace jpxlg
