/*
 * jpxl_oracle.c -- CPU restatement of JPXL's VarDCT pixel path and its
 * SSIMULACRA2 metric, used ONLY as a test oracle.
 *
 * This file is test infrastructure.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as the
 * timed CPU baseline; the product path (jpxl_amd/, libjpxlg) never calls it.
 *
 * Every function restates one reference function and cites it (paths are
 * relative to JPXL/crates/).  Floating-point expressions are written in the
 * reference's evaluation order; the file must be compiled with
 * -ffp-contract=off so that no multiply-add is fused unless the reference
 * itself calls mul_add (written here as fmaf).
 *
 * Parity: tests/test_oracle_cpu.py pins each function against the
 * reference's own known-answer tests, closed forms or brute force.  The
 * functions no such check reaches say "parity unpinned" beside them.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint64_t d2u(double f) { uint64_t u; memcpy(&u, &f, 8); return u; }
static inline double u2d(uint64_t u) { double f; memcpy(&f, &u, 8); return f; }

/* ------------------------------------------------------------------------ */
/* Cube roots                                                                */
/* ------------------------------------------------------------------------ */

/* jpxl-core/src/simd.rs:36-67 reproducible_cbrt: musl-style 5-bit estimate,
 * two f64 Halley steps t = t*(2x + t^3)/(x + 2 t^3), one rounding to f32. */
EXPORT float orc_cbrt_core(float value) {
    const uint32_t B1 = 709958130u, B2 = 642849266u;
    uint32_t bits = f2u(value);
    uint32_t mag = bits & 0x7fffffffu;
    if (mag == 0 || mag >= 0x7f800000u) return value;
    uint32_t est;
    if (mag < 0x00800000u) {
        float scaled = value * 16777216.0f;
        est = (f2u(scaled) & 0x7fffffffu) / 3u + B2;
    } else {
        est = mag / 3u + B1;
    }
    uint32_t eb = (bits & 0x80000000u) | est;
    double x = (double)value;
    double t = (double)u2f(eb);
    for (int i = 0; i < 2; i++) {
        double r = (t * t) * t;
        t = (t * ((x + x) + r)) / ((x + r) + r);
    }
    return (float)t;
}

/* jpxl-perceptual/src/color.rs:46-61 reproducible_cbrt (metric copy): double
 * bit-hack estimate plus four f64 Newton steps. */
EXPORT float orc_cbrt_metric(float value) {
    double x = (double)value;
    double a = fabs(x);
    double y = u2d(d2u(a) / 3u + (0x2A9F7893ull << 32));
    for (int i = 0; i < 4; i++) y -= (y * y * y - a) / (3.0 * y * y);
    float root = (float)y;
    float sgn = x < 0.0 ? -root : root;
    if (value == 0.0f || !isfinite(value)) return value;
    return sgn;
}

/* ------------------------------------------------------------------------ */
/* Transfer functions and XYB                                                */
/* ------------------------------------------------------------------------ */

/* jpxl-core/src/color.rs:517-528 srgb_to_linear (negatives odd-symmetric). */
EXPORT float orc_srgb_to_linear(float v) {
    float a = fabsf(v);
    float d;
    if (a <= 12.92f * 0.0031308f) d = a / 12.92f;
    else d = powf((a + 0.055f) / 1.055f, 2.4f);
    return v < 0.0f ? -d : d;
}

/* jpxl-core/src/color.rs:490-501 linear_to_srgb. */
EXPORT float orc_linear_to_srgb(float v) {
    float a = fabsf(v);
    float e;
    if (a <= 0.0031308f) e = 12.92f * a;
    else e = 1.055f * powf(a, 1.0f / 2.4f) - 0.055f;
    return v < 0.0f ? -e : e;
}

/* jpxl-core/src/color.rs:53-76 opsin constants. */
static const float OPSIN_M[3][3] = {
    {0.3f, 0.622f, 0.078f},
    {0.23f, 0.692f, 0.078f},
    {0.2434227f, 0.2047674f, 0.5518099f},
};
static const float OPSIN_BIAS = 0.0037930734f;
static const float OPSIN_BIAS_CBRT = 0.1559542f;

/* jpxl-core/src/color.rs:84-97 linear_srgb_to_xyb. */
static void xyb_of(float r, float g, float b, float *X, float *Y, float *B) {
    float lm = OPSIN_M[0][0] * r + OPSIN_M[0][1] * g + OPSIN_M[0][2] * b + OPSIN_BIAS;
    float mm = OPSIN_M[1][0] * r + OPSIN_M[1][1] * g + OPSIN_M[1][2] * b + OPSIN_BIAS;
    float sm = OPSIN_M[2][0] * r + OPSIN_M[2][1] * g + OPSIN_M[2][2] * b + OPSIN_BIAS;
    float lg = orc_cbrt_core(lm) - OPSIN_BIAS_CBRT;
    float mg = orc_cbrt_core(mm) - OPSIN_BIAS_CBRT;
    float sg = orc_cbrt_core(sm) - OPSIN_BIAS_CBRT;
    *X = 0.5f * (lg - mg);
    *Y = 0.5f * (lg + mg);
    *B = sg;
}

EXPORT void orc_xyb_pixel(float r, float g, float b, float *out3) {
    xyb_of(r, g, b, &out3[0], &out3[1], &out3[2]);
}

/* The sRGB LUT used by every 8-bit entry point:
 * jpxl-encode-policy/src/source.rs:253-254 and
 * jpxl-perceptual/src/evaluator.rs:93-94. */
EXPORT void orc_srgb8_lut(float *lut256) {
    for (int v = 0; v < 256; v++) lut256[v] = orc_srgb_to_linear((float)v / 255.0f);
}

/* jpxl-encode-policy/src/source.rs:389-398: 65536-entry LUT dividing by
 * min(65535, 2^bits - 1). */
EXPORT void orc_srgb16_lut(int bits, float *lut65536) {
    float m = (float)((1u << bits) - 1u);
    float max = 65535.0f < m ? 65535.0f : m;
    float divisor = max > 0.0f ? max : 1.0f;
    for (uint32_t v = 0; v <= 65535u; v++) lut65536[v] = orc_srgb_to_linear((float)v / divisor);
}

/* jpxl-encode-policy/src/source.rs:226-331 PreparedFrame::from_srgb8_with and
 * :355 from_srgb16_with: LUT to linear, then XYB per pixel.  Also writes the
 * linear planes (the metric reference, evaluator.rs:87-99) when non-NULL.
 * Returns the grayscale flag. */
EXPORT int orc_prepare_frame(int w, int h, int bits, const void *rgb, float *X, float *Y,
                             float *B, float *lin_r, float *lin_g, float *lin_b) {
    float *lut = (float *)malloc(sizeof(float) * 65536);
    if (bits == 8) orc_srgb8_lut(lut);
    else orc_srgb16_lut(bits, lut);
    int gray = 1;
    size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; i++) {
        uint32_t cr, cg, cb;
        if (bits == 8) {
            const uint8_t *p = (const uint8_t *)rgb + 3 * i;
            cr = p[0]; cg = p[1]; cb = p[2];
        } else {
            const uint16_t *p = (const uint16_t *)rgb + 3 * i;
            cr = p[0]; cg = p[1]; cb = p[2];
        }
        gray &= (cr == cg && cg == cb);
        float r = lut[cr], g = lut[cg], b = lut[cb];
        if (lin_r) { lin_r[i] = r; lin_g[i] = g; lin_b[i] = b; }
        xyb_of(r, g, b, &X[i], &Y[i], &B[i]);
    }
    free(lut);
    return gray;
}

/* jpxl-perceptual/src/color.rs:14-25, :66-81 linear_srgb_to_positive_xyb. */
static void positive_xyb(float r, float g, float b, float *ox, float *oy, float *ob) {
    const float BIAS = 0.0037930734f, BIAS_CBRT = 0.1559542f;
    float lm = OPSIN_M[0][0] * r + OPSIN_M[0][1] * g + OPSIN_M[0][2] * b + BIAS;
    float mm = OPSIN_M[1][0] * r + OPSIN_M[1][1] * g + OPSIN_M[1][2] * b + BIAS;
    float sm = OPSIN_M[2][0] * r + OPSIN_M[2][1] * g + OPSIN_M[2][2] * b + BIAS;
    float lg = orc_cbrt_metric(lm) - BIAS_CBRT;
    float mg = orc_cbrt_metric(mm) - BIAS_CBRT;
    float sg = orc_cbrt_metric(sm) - BIAS_CBRT;
    float x = 0.5f * (lg - mg);
    float y = 0.5f * (lg + mg);
    float bb = sg;
    *ox = x * 14.0f + 0.42f;
    *oy = y + 0.01f;
    *ob = (bb - y) + 0.55f;
}

EXPORT void orc_positive_xyb_pixel(float r, float g, float b, float *out3) {
    positive_xyb(r, g, b, &out3[0], &out3[1], &out3[2]);
}

/* jpxl-plan-render/src/lib.rs:55-68 default inverse matrix and bias, and
 * jpxl-core/src/color.rs:626-742 OpsinInverse::convert (FMA row products). */
static const float INV_M[9] = {11.031567f, -9.866944f, -0.16462299f, -3.2541474f, 4.4187705f,
                               -0.16462299f, -3.6588513f, 2.712923f, 1.9459282f};
static const float INV_BIAS = -0.0037930733f;

static void opsin_inverse(float x, float y, float b, float *o) {
    float bias_cbrt = orc_cbrt_core(INV_BIAS);
    float itscale = 255.0f / 255.0f;
    float gam[3] = {y + x, y - x, b};
    float mix[3];
    for (int c = 0; c < 3; c++) {
        float v = gam[c] - bias_cbrt;
        mix[c] = (v * v * v + INV_BIAS) * itscale;
    }
    for (int r = 0; r < 3; r++) {
        float acc = 0.0f;
        for (int c = 0; c < 3; c++) acc = fmaf(INV_M[r * 3 + c], mix[c], acc);
        o[r] = acc;
    }
}

EXPORT void orc_opsin_inverse_pixel(float x, float y, float b, float *out3) {
    opsin_inverse(x, y, b, out3);
}

/* ------------------------------------------------------------------------ */
/* Depth quantizer (jpxl-plan-render/src/lib.rs:147-468)                     */
/* ------------------------------------------------------------------------ */

static float full_scale(int bits) { return bits >= 32 ? 65535.0f : (float)((1u << bits) - 1u); }

/* lib.rs:201-216 composed_level. */
static uint32_t composed_level(float x, float max) {
    float scaled = roundf(orc_linear_to_srgb(x) * max);
    if (!isfinite(scaled)) return 0;
    if (scaled < 0.0f) scaled = 0.0f;
    if (scaled > max) scaled = max;
    return (uint32_t)scaled;
}

/* lib.rs:313-334 quantize_thresholds (bisection over the f32 bit lattice) and
 * lib.rs:162-173 linear_lut.  thresholds has levels-1 entries. */
EXPORT int orc_depth_tables(int bits, float *thresholds, float *lut) {
    float max = full_scale(bits);
    int levels = bits >= 32 ? 65536 : (1 << (bits < 1 ? 1 : (bits > 16 ? 16 : bits)));
    for (int q = 0; q < levels; q++) lut[q] = orc_srgb_to_linear((float)q / max);
    for (int q = 1; q < levels; q++) {
        uint32_t lo = 0, hi = f2u(2.0f);
        while (lo < hi) {
            uint32_t mid = lo + (hi - lo) / 2;
            if (composed_level(u2f(mid), max) >= (uint32_t)q) hi = mid;
            else lo = mid + 1;
        }
        thresholds[q - 1] = u2f(lo);
    }
    return levels;
}

/* lib.rs:341-347 level_of + lib.rs:421-454 linear_at_depth. */
static float linear_at_depth(float x, const float *thr, const float *lut, int levels) {
    if (!isfinite(x)) return lut[0];
    int lo = 0, hi = levels - 1; /* count of thresholds <= x */
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (thr[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    return lut[lo];
}

/* ------------------------------------------------------------------------ */
/* DCT (jpxl-core/src/dct.rs)                                                */
/* ------------------------------------------------------------------------ */

#define R2 0.70710677f /* FRAC_1_SQRT_2 as f32 */
static const float COS_PI_8 = 0.9238795f, SIN_PI_8 = 0.38268343f;

/* dct.rs:133-185: orthonormal DCT-IV leaves, entries
 * sqrt(2/M) cos(pi (2n+1)(2u+1) / 4M).  The entries are +/- one of M
 * magnitudes; the magnitudes are the f32 values the reference tables hold
 * (for M = 4 and 8 these are the nearest floats; four of the sixteen M = 16
 * magnitudes sit one ulp away from nearest and are written out). */
static float DCT4_4[16], DCT4_8[64], DCT4_16[256];
static int dct_tables_ready = 0;

static const uint32_t MAG16_BITS[16] = {
    0x3eb4cd21u, 0x3eb30f61u, 0x3eaf982cu, 0x3eaa700cu, 0x3ea3a3b9u, 0x3e9b43f5u,
    0x3e916567u, 0x3e862061u, 0x3e73215cu, 0x3e57aa8bu, 0x3e3a2005u, 0x3e1aca9eu,
    0x3df3ef2fu, 0x3dafefbbu, 0x3d547d0au, 0x3c8e1d7au};

static void fold_table(int M, const float *mags, float *out) {
    for (int u = 0; u < M; u++)
        for (int i = 0; i < M; i++) {
            int a = ((2 * i + 1) * (2 * u + 1)) % (8 * M);
            float s = 1.0f;
            if (a > 4 * M) a = 8 * M - a;
            if (a > 2 * M) { a = 4 * M - a; s = -1.0f; }
            out[u * M + i] = s * mags[(a - 1) / 2];
        }
}

static void init_dct_tables(void) {
    if (dct_tables_ready) return;
    float m4[4], m8[8], m16[16];
    for (int m = 0; m < 4; m++) m4[m] = (float)(sqrt(2.0 / 4) * cos(M_PI * (2 * m + 1) / 16.0));
    for (int m = 0; m < 8; m++) m8[m] = (float)(sqrt(2.0 / 8) * cos(M_PI * (2 * m + 1) / 32.0));
    for (int m = 0; m < 16; m++) m16[m] = u2f(MAG16_BITS[m]);
    fold_table(4, m4, DCT4_4);
    fold_table(8, m8, DCT4_8);
    fold_table(16, m16, DCT4_16);
    dct_tables_ready = 1;
}

EXPORT void orc_dct4_table(int M, float *out) {
    init_dct_tables();
    const float *t = M == 4 ? DCT4_4 : (M == 8 ? DCT4_8 : DCT4_16);
    memcpy(out, t, sizeof(float) * M * M);
}

/* dct.rs:199-230 dct_iv_2 / dct_iv_{4,8,16}: accumulate acc = acc + c*x. */
static void dct_iv(float *v, int M) {
    if (M == 2) {
        float x0 = v[0], x1 = v[1];
        v[0] = COS_PI_8 * x0 + SIN_PI_8 * x1;
        v[1] = SIN_PI_8 * x0 - COS_PI_8 * x1;
        return;
    }
    const float *t = M == 4 ? DCT4_4 : (M == 8 ? DCT4_8 : DCT4_16);
    float x[16];
    memcpy(x, v, sizeof(float) * M);
    for (int u = 0; u < M; u++) {
        float acc = 0.0f;
        for (int n = 0; n < M; n++) acc = acc + t[u * M + n] * x[n];
        v[u] = acc;
    }
}

static void dct_ii(float *v, int N);
static void dct_iii(float *v, int N);

/* dct.rs:236-241 dct_ii_2 (its own inverse). */
static void dct_ii_2(float *v) {
    float x0 = v[0], x1 = v[1];
    v[0] = (x0 + x1) * R2;
    v[1] = (x0 - x1) * R2;
}

/* dct.rs:245-255 dct_ii_4. */
static void dct_ii_4(float *v) {
    float even[2] = {v[0] + v[3], v[1] + v[2]};
    float odd[2] = {v[0] - v[3], v[1] - v[2]};
    dct_ii_2(even);
    dct_iv(odd, 2);
    v[0] = even[0] * R2;
    v[1] = odd[0] * R2;
    v[2] = even[1] * R2;
    v[3] = odd[1] * R2;
}

/* dct.rs:259-276 dct_iii_4. */
static void dct_iii_4(float *v) {
    float even[2] = {v[0] * R2, v[2] * R2};
    float odd[2] = {v[1] * R2, v[3] * R2};
    dct_ii_2(even);
    dct_iv(odd, 2);
    v[0] = even[0] + odd[0];
    v[3] = even[0] - odd[0];
    v[1] = even[1] + odd[1];
    v[2] = even[1] - odd[1];
}

/* dct.rs:278-362 split/merge + dct.rs:364-426 the 8/16/32 recursion. */
static void dct_ii(float *v, int N) {
    if (N == 1) return;
    if (N == 2) { dct_ii_2(v); return; }
    if (N == 4) { dct_ii_4(v); return; }
    int M = N / 2;
    float even[16], odd[16];
    for (int k = 0; k < M; k++) {
        even[k] = v[k] + v[N - 1 - k];
        odd[k] = v[k] - v[N - 1 - k];
    }
    dct_ii(even, M);
    dct_iv(odd, M);
    for (int k = 0; k < M; k++) {
        v[2 * k] = even[k] * R2;
        v[2 * k + 1] = odd[k] * R2;
    }
}

static void dct_iii(float *v, int N) {
    if (N == 1) return;
    if (N == 2) { dct_ii_2(v); return; }
    if (N == 4) { dct_iii_4(v); return; }
    int M = N / 2;
    float even[16], odd[16];
    for (int k = 0; k < M; k++) {
        even[k] = v[2 * k] * R2;
        odd[k] = v[2 * k + 1] * R2;
    }
    dct_iii(even, M);
    dct_iv(odd, M);
    for (int k = 0; k < M; k++) {
        v[k] = even[k] + odd[k];
        v[N - 1 - k] = even[k] - odd[k];
    }
}

/* dct.rs:789-812 column_pass_lanes: the 1-D kernel down every column, then
 * x 1/sqrt(rows) (forward) or x sqrt(rows) (inverse). */
static void column_pass(float *m, int rows, int cols, int forward) {
    float factor = forward ? 1.0f / sqrtf((float)rows) : sqrtf((float)rows);
    float col[32];
    for (int c = 0; c < cols; c++) {
        for (int r = 0; r < rows; r++) col[r] = m[r * cols + c];
        if (forward) dct_ii(col, rows);
        else dct_iii(col, rows);
        for (int r = 0; r < rows; r++) m[r * cols + c] = col[r] * factor;
    }
}

static void transpose_into(const float *src, float *dst, int rows, int cols) {
    for (int r = 0; r < rows; r++)
        for (int c = 0; c < cols; c++) dst[c * rows + r] = src[r * cols + c];
}

/* dct.rs:956-975 dct_2d_in_place. */
EXPORT void orc_dct2d_fwd(int rows, int cols, const float *in, float *out) {
    init_dct_tables();
    int n = rows * cols;
    float work[1024], scratch[1024];
    memcpy(work, in, sizeof(float) * n);
    column_pass(work, rows, cols, 1);
    transpose_into(work, scratch, rows, cols);
    column_pass(scratch, cols, rows, 1);
    if (cols <= rows) memcpy(out, scratch, sizeof(float) * n);
    else transpose_into(scratch, out, cols, rows);
}

/* dct.rs:982-997 idct_2d_in_place. */
EXPORT void orc_dct2d_inv(int rows, int cols, const float *in, float *out) {
    init_dct_tables();
    int n = rows * cols;
    float work[1024], scratch[1024];
    memcpy(work, in, sizeof(float) * n);
    if (cols > rows) transpose_into(work, scratch, rows, cols);
    else memcpy(scratch, work, sizeof(float) * n);
    column_pass(scratch, cols, rows, 0);
    transpose_into(scratch, work, cols, rows);
    column_pass(work, rows, cols, 0);
    memcpy(out, work, sizeof(float) * n);
}

/* dct.rs:1375-1404 dct2_naive / dct3_naive: the f64 orthonormal definition. */
EXPORT void orc_dct1d_naive(int n, const double *in, double *out, int inverse) {
    for (int k = 0; k < n; k++) {
        double acc = 0.0;
        for (int j = 0; j < n; j++) {
            int u = inverse ? j : k, x = inverse ? k : j;
            double s = u == 0 ? sqrt(1.0 / n) : sqrt(2.0 / n);
            acc += s * in[j] * cos(M_PI * (2 * x + 1) * u / (2.0 * n));
        }
        out[k] = acc;
    }
}

/* jpxl-core/src/varblock.rs:789-811 scale_f_closed_form with the argument
 * taken as the varblock dimension (lf_count * 8). */
static float scale_f(int c, int lf_count) {
    int b = lf_count * 8;
    double cf = c, bf = b;
    double inv = cos(cf * M_PI / (2.0 * bf)) * cos(cf * M_PI / bf) * cos(2.0 * cf * M_PI / bf);
    return (float)(1.0 / inv);
}

EXPORT float orc_scale_f(int c, int lf_count) { return scale_f(c, lf_count); }

/* jpxl-core/src/forward.rs:779-830 lf_from_llf_into for a square transform
 * whose LF rectangle is n x n (coefficient stride `side`). */
static void lf_from_llf_square(const float *coeff, int side, int n, float *lf) {
    if (n == 1) { lf[0] = coeff[0]; return; }
    float work[16];
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++)
            work[y * n + x] = coeff[y * side + x] / (scale_f(y, n) * scale_f(x, n));
    orc_dct2d_inv(n, n, work, lf);
}

/* jpxl-core/src/varblock.rs:817-855 llf_from_lf for a square transform. */
static void llf_from_lf_square(const float *lf, int n, float *llf) {
    if (n == 1) { llf[0] = lf[0]; return; }
    float dc[16];
    orc_dct2d_fwd(n, n, lf, dc);
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) llf[y * n + x] = dc[y * n + x] * scale_f(y, n) * scale_f(x, n);
}

/* ------------------------------------------------------------------------ */
/* Dequantization matrices (jpxl-core/src/dequant.rs)                        */
/* ------------------------------------------------------------------------ */

/* dequant.rs:824-897 Table I.6 dct_params for the three square DCTs. */
static const double P8[3][6] = {{3150.0, 0.0, -0.4, -0.4, -0.4, -2.0},
                                {560.0, 0.0, -0.3, -0.3, -0.3, -0.3},
                                {512.0, -2.0, -1.0, 0.0, -1.0, -2.0}};
static const double P16[3][7] = {
    {8996.8725711814115328, -1.3000777393353804, -0.49424529824571225, -0.439093774457103443,
     -0.6350101832695744, -0.90177264050827612, -1.6162099239887414},
    {3191.48366296844234752, -0.67424582104194355, -0.80745813428471001,
     -0.44925837484843441, -0.35865440981033403, -0.31322389111877305, -0.37615025315725483},
    {1157.50408145487200256, -2.0531423165804414, -1.4, -0.50687130033378396,
     -0.42708730624733904, -1.4856834539296244, -4.9209142884401604}};
static const double P32[3][8] = {
    {15718.40830982518931456, -1.025, -0.98, -0.9012, -0.4, -0.48819395464, -0.421064, -0.27},
    {7305.7636810695983104, -0.8041958212306401, -0.7633036457487539, -0.55660379990111464,
     -0.49785304658857626, -0.43699592683512467, -0.40180866526242109, -0.27321683125358037},
    {3803.53173721215041536, -3.060733579805728, -2.0413270132490346, -2.0235650159727417,
     -0.5495389509954993, -0.4, -0.4, -0.3}};

/* dequant.rs:400-402 Mult. */
static double mult_(double v) { return v > 0.0 ? 1.0 + v : 1.0 / (1.0 - v); }
EXPORT double orc_dequant_mult(double v) { return mult_(v); }

/* dequant.rs:418-440 interpolate; NaN where the reference returns an error
 * (no bands, or a position scaling outside them). */
static double interpolate(double pos, double max, const double *bands, int nb) {
    int last = nb - 1;
    if (last < 0) return NAN;
    if (last == 0) return bands[0];
    double scaled_pos = pos * (double)last / max;
    double si = floor(scaled_pos);
    if (!(si >= 0.0 && si < (double)last)) return NAN;
    int idx = (int)si;
    double frac = scaled_pos - si;
    double a = bands[idx], b = bands[idx + 1];
    return a * pow(b / a, frac);
}
EXPORT double orc_dequant_interpolate(double pos, double max, const double *bands, int nb) {
    return interpolate(pos, max, bands, nb);
}

/* dequant.rs:449-468 bands_from: bands[0] = params[0], bands[i] =
 * bands[i-1] * Mult(params[i]).  Returns -1 where the reference errors. */
EXPORT int orc_dequant_bands(const double *params, int np, double *bands) {
    if (np < 1 || !isfinite(params[0]) || params[0] <= 0.0) return -1;
    bands[0] = params[0];
    for (int i = 1; i < np; i++) {
        bands[i] = bands[i - 1] * mult_(params[i]);
        if (!isfinite(bands[i]) || bands[i] <= 0.0) return -1;
    }
    return 0;
}

/* dequant.rs:466-510 dct_quant_weights + dequant.rs:321-338 reciprocal. */
EXPORT void orc_dequant_matrix(int side, int channel, float *out) {
    const double *params;
    int np;
    if (side == 8) { params = P8[channel]; np = 6; }
    else if (side == 16) { params = P16[channel]; np = 7; }
    else { params = P32[channel]; np = 8; }
    double bands[8];
    orc_dequant_bands(params, np, bands);
    double max = M_SQRT2 + 1e-6;
    double span = (double)(side - 1);
    for (int y = 0; y < side; y++)
        for (int x = 0; x < side; x++) {
            double dx = (double)x / span, dy = (double)y / span;
            double dist = sqrt(dx * dx + dy * dy);
            double w = interpolate(dist, max, bands, np);
            out[y * side + x] = (float)(1.0 / w);
        }
}

/* jpxl-core/src/reconstruct.rs:27-36 hf_multiplier. */
static float hf_multiplier(uint32_t gs, uint32_t hf_mul) {
    double denom = (double)gs * (double)hf_mul;
    if (denom <= 0.0) return 0.0f;
    return (float)(65536.0 / denom);
}

/* reconstruct.rs:38-43 qm_multiplier: 0.8f.powi(qm - 2), as compiler-rt's
 * __powisf2 evaluates a run-time exponent. */
static float powi_f(float a, int b) {
    int recip = b < 0;
    float r = 1.0f;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0f / r : r;
}
static float qm_multiplier(int qm_scale) { return powi_f(0.8f, qm_scale - 2); }

/* reconstruct.rs:45-59 bias_adjust with Table L.1 defaults. */
static const float QUANT_BIAS[3] = {1.0f - 0.054650073f, 1.0f - 0.0700545f, 1.0f - 0.049935103f};
static const float QUANT_BIAS_NUM = 0.145f;

static float bias_adjust(int32_t q, float bias) {
    float f = (float)q;
    if (abs(q) <= 1) return f * bias;
    return f - QUANT_BIAS_NUM / f;
}

/* reconstruct.rs:61-76 lf_multipliers with the G.1.2 default weights. */
static void lf_multipliers(uint32_t gs, uint32_t quant_lf, float *m) {
    const float w[3] = {1.0f / 32.0f, 1.0f / 4.0f, 1.0f / 2.0f};
    double denom = (double)gs * (double)quant_lf;
    for (int c = 0; c < 3; c++) m[c] = (float)(65536.0 * (double)(w[c] / 128.0f) / denom);
}

/* reconstruct.rs:96-116 cfl_factors. */
static float cfl_k(float base, int factor) { return base + (float)factor / 84.0f; }

/* ------------------------------------------------------------------------ */
/* Natural coefficient order and truncation weights                          */
/* ------------------------------------------------------------------------ */

typedef struct { int64_t k1, k2; int lin; } hf_key;
static int cmp_hf(const void *a, const void *b) {
    const hf_key *x = a, *y = b;
    if (x->k1 != y->k1) return x->k1 < y->k1 ? -1 : 1;
    if (x->k2 != y->k2) return x->k2 < y->k2 ? -1 : 1;
    return x->lin < y->lin ? -1 : (x->lin > y->lin);
}

/* jpxl-core/src/varblock.rs:710-745 natural_coeff_order for a square side. */
EXPORT void orc_natural_order(int side, uint32_t *order) {
    int cx = side / 8, cy = side / 8;
    int64_t scale = cx;
    int nl = 0, nh = 0;
    hf_key *hf = malloc(sizeof(hf_key) * side * side);
    for (int y = 0; y < side; y++)
        for (int x = 0; x < side; x++) {
            int lin = y * side + x;
            if (x < cx && y < cy) { order[y * cx + x] = lin; nl++; continue; }
            int64_t sx = (int64_t)x * scale / cx, sy = (int64_t)y * scale / cy;
            int64_t k1 = sx + sy, k2 = sx - sy;
            if (k1 % 2 == 1) k2 = -k2;
            hf[nh].k1 = k1; hf[nh].k2 = k2; hf[nh].lin = lin; nh++;
        }
    qsort(hf, nh, sizeof(hf_key), cmp_hf); /* stable: ties never occur on (k1,k2) */
    for (int i = 0; i < nh; i++) order[nl + i] = hf[i].lin;
    free(hf);
}

/* jpxl-encode-policy/src/csf.rs:95-105 radial_frequency. */
static double radial_frequency(int u, int v, int side) {
    double fu = u, fv = v, n = side;
    return sqrt((fu / n) * (fu / n) + (fv / n) * (fv / n)) / M_SQRT2;
}

typedef struct { double f, w; int seq; } curve_pt;
static int cmp_curve(const void *a, const void *b) {
    const curve_pt *x = a, *y = b;
    if (x->f != y->f) return x->f < y->f ? -1 : 1;
    return x->seq - y->seq; /* stable sort */
}

/* csf.rs:221-316 quant_donor_weights (donor: DCT8x8 Y matrix). */
EXPORT void orc_donor_weights(int side, float *weights) {
    int llf = side / 8;
    float m8[64];
    orc_dequant_matrix(8, 1, m8);
    curve_pt pts[64];
    int np = 0;
    for (int cell = 0; cell < 64; cell++) {
        int u = cell / 8, v = cell % 8;
        if (u < 1 && v < 1) continue;
        double step = (double)m8[u * 8 + v]; /* matrix.at(v, u) = data[u*8+v] */
        pts[np].f = radial_frequency(u, v, 8);
        pts[np].w = 1.0 / (step * step);
        pts[np].seq = np;
        np++;
    }
    qsort(pts, np, sizeof(curve_pt), cmp_curve);
    int cells = side * side;
    double *raw = calloc(cells, sizeof(double));
    double sum = 0.0;
    int count = 0;
    for (int cell = 0; cell < cells; cell++) {
        int u = cell / side, v = cell % side;
        weights[cell] = 1.0f;
        if (u < llf && v < llf) continue;
        double f = radial_frequency(u, v, side), w;
        if (f <= pts[0].f) w = pts[0].w;
        else if (f >= pts[np - 1].f) w = pts[np - 1].w;
        else {
            w = pts[np - 1].w;
            for (int i = 0; i + 1 < np; i++) {
                double f0 = pts[i].f, w0 = pts[i].w, f1 = pts[i + 1].f, w1 = pts[i + 1].w;
                if (f1 >= f) {
                    w = (f1 == f0) ? w1 : w0 + (w1 - w0) * ((f - f0) / (f1 - f0));
                    break;
                }
            }
        }
        raw[cell] = w;
        sum += w;
        count++;
    }
    double mean = sum / (double)count;
    for (int cell = 0; cell < cells; cell++) {
        int u = cell / side, v = cell % side;
        if (u < llf && v < llf) continue;
        weights[cell] = (float)(raw[cell] / mean);
    }
    free(raw);
}

/* ------------------------------------------------------------------------ */
/* Quantizers (jpxl-encode-policy/src/quantize.rs)                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint32_t global_scale, hf_mul, quant_lf;  /* the wire values (after the AQ factorization) */
    int x_qm, b_qm;
    int truncate;       /* QuantizerChoiceMode::TrailingTruncation */
    float lambda_scale; /* request.lambda_scale */
    float zero_token_bits;
    const float *aq_adj; /* per-atom AQ adjustment (bw x bh), NULL = no field */
    int aq_refine_only, aq_fine;
    int cover_legacy;    /* the cover prices with CoverRateModel::Legacy and flat cell weights */
} orc_quant_params;

typedef struct {
    int side;
    float step[3][1024];
    float thr[3][1024];
    float weight[3][1024];
    uint32_t order[1024];
} hf_quantizer;

/* quantize.rs:1720-1746 clamp_round (ties away, |v| <= 2^20). Returns 0 and
 * sets *ok = 0 on a value the reference rejects. */
static int32_t clamp_round(float v, int *ok) {
    if (!isfinite(v)) { *ok = 0; return 0; }
    float adj = signbit(v) ? v - 0.5f : v + 0.5f;
    if (!(adj > -2147483648.0f && adj < 2147483648.0f)) { *ok = 0; return 0; }
    int32_t value = (int32_t)adj;
    if ((uint32_t)abs(value) > (1u << 20)) { *ok = 0; return 0; }
    return value;
}

/* lib.rs:2033-2172 HfQuantizers::new_with_scales_for_transforms lambda for
 * one channel, from the DCT8 baseline steps. */
static double lambda_for(const orc_quant_params *p, int channel) {
    float m8[64];
    orc_dequant_matrix(8, channel, m8);
    float mul = hf_multiplier(p->global_scale, p->hf_mul);
    float sc = channel == 0 ? mul * qm_multiplier(p->x_qm)
                            : (channel == 1 ? mul : mul * qm_multiplier(p->b_qm));
    double sum = 0.0;
    for (int cell = 1; cell < 64; cell++) {
        int x = cell % 8, y = cell / 8;
        double step = 8.0 * (double)(sc * m8[y * 8 + x]);
        sum += step * step;
    }
    double mean = sum / 63.0;
    double scale = (isfinite(p->lambda_scale) && p->lambda_scale > 0.0f) ? (double)p->lambda_scale : 1.0;
    return mean > 0.0 ? scale * 16.0 / mean : 0.0;
}

/* quantize.rs:301-366 HfQuantizer::new_with_matrices + lib.rs:2149-2170, for
 * the (transform, HfMul) key `hf_mul`; the Lagrange weights come from the
 * frame's baseline DCT8 quantizer (p->hf_mul) whatever the key. */
static void build_hf_quantizer_mul(hf_quantizer *q, int side, const orc_quant_params *p, uint32_t hf_mul) {
    q->side = side;
    float mul = hf_multiplier(p->global_scale, hf_mul);
    float scale[3] = {mul * qm_multiplier(p->x_qm), mul, mul * qm_multiplier(p->b_qm)};
    float freq[1024];
    orc_donor_weights(side, freq);
    int cells = side * side;
    for (int c = 0; c < 3; c++) {
        float m[1024];
        orc_dequant_matrix(side, c, m);
        double lam = lambda_for(p, c);
        double to_sample = (double)(side * side) * 1.0;
        for (int cell = 0; cell < cells; cell++) {
            float step = scale[c] * m[cell];
            q->step[c][cell] = step;
            q->thr[c][cell] = 0.5f * QUANT_BIAS[c] * step;
            q->weight[c][cell] = (float)(lam * to_sample * (double)freq[cell]);
        }
    }
    orc_natural_order(side, q->order);
}

static void build_hf_quantizer(hf_quantizer *q, int side, const orc_quant_params *p) {
    build_hf_quantizer_mul(q, side, p, p->hf_mul);
}

/* quantize.rs:1400-1499 choose (nearest reconstruction, ties to smaller |q|). */
static int32_t hf_choose(const hf_quantizer *q, float target, int c, int cell, int *ok) {
    float step = q->step[c][cell];
    if (fabsf(target) <= q->thr[c][cell]) return 0;
    int32_t e = clamp_round(target / step, ok);
    if (!*ok) return 0;
    int32_t cand[4] = {0, e - 1, e, e + 1};
    int32_t best = 0;
    float best_err = INFINITY;
    for (int i = 0; i < 4; i++) {
        int32_t v = cand[i];
        if (abs(v) > (1 << 20)) continue;
        float err = fabsf(bias_adjust(v, QUANT_BIAS[c]) * step - target);
        if (err < best_err || (err == best_err && abs(v) < abs(best))) { best = v; best_err = err; }
    }
    return best;
}

static float hf_recon(const hf_quantizer *q, int32_t v, int c, int cell) {
    return bias_adjust(v, QUANT_BIAS[c]) * q->step[c][cell];
}

static int prev_nonzero(const int32_t *quant, const uint32_t *order, int first, int before) {
    for (int pos = before - 1; pos >= first; pos--)
        if (quant[order[pos]] != 0) return pos;
    return -1;
}

/* quantize.rs:468-545 truncate_trailing. `targets` is indexed by cell;
 * `recon_y` (may be NULL) is refreshed to reconstruct(0) for dropped cells. */
static void truncate_trailing(const hf_quantizer *q, int c, int32_t *quant, int num_blocks,
                              float zero_token_bits, const float *targets, float *recon_y) {
    int size = q->side * q->side;
    if (num_blocks >= size) return;
    int last = prev_nonzero(quant, q->order, num_blocks, size);
    while (last >= 0) {
        int cell = (int)q->order[last];
        int32_t v = quant[cell];
        if (v == 0) { last = prev_nonzero(quant, q->order, num_blocks, last); continue; }
        float weight = q->weight[c][cell];
        int previous = prev_nonzero(quant, q->order, num_blocks, last);
        int exposed = previous >= 0 ? last - previous - 1 : last - num_blocks;
        if (exposed < 0) exposed = 0;
        float zeros = (float)exposed;
        uint32_t mag = (uint32_t)abs(v);
        float own_bits = (float)(33 - __builtin_clz(mag));
        float bits_saved = own_bits + zeros * zero_token_bits;
        float target = targets[cell];
        float recon = hf_recon(q, v, c, cell);
        float kept = recon - target;
        float added = fmaf(target, target, -(kept * kept));
        float cost = weight * added;
        if (bits_saved <= cost) break;
        quant[cell] = 0;
        if (recon_y) recon_y[cell] = hf_recon(q, 0, c, cell);
        last = previous;
    }
}

/* One varblock's trailing truncation (quantize.rs:468-545) under the
 * quantizer of side `side` and params p: quant (cell order, side*side) is
 * updated in place; targets are the pre-quantization coefficients. */
EXPORT void orc_truncate_block(int side, int c, const orc_quant_params *p, int32_t *quant, const float *targets) {
    hf_quantizer *q = calloc(1, sizeof(hf_quantizer));
    build_hf_quantizer(q, side, p);
    int n = side / 8;
    truncate_trailing(q, c, quant, n * n, p->zero_token_bits, targets, NULL);
    free(q);
}

/* quantize.rs:150-213 LfQuantizer. */
static float lf_recon(const float *mdc, int32_t v, int c) { return mdc[c] * (float)v / 1.0f; }
static int32_t lf_quantize(const float *mdc, float target, int c, int *ok) {
    float scaled = target * 1.0f / mdc[c];
    return clamp_round(scaled, ok);
}

/* ------------------------------------------------------------------------ */
/* Adaptive quantization (jpxl-encode-policy/src/analysis.rs:111-158,        */
/* field.rs:224-420, lib.rs:1847-1944)                                       */
/* ------------------------------------------------------------------------ */

/* field.rs:52-91 AqMode, in wire order of the C ABI. */
enum { AQ_OFF = 0, AQ_MASKING = 1, AQ_UNIFORM = 2, AQ_FINE_MASKING = 3, AQ_FINE_UNIFORM = 4, AQ_EDGE_REFINE = 5 };
#define AQ_FINE_BASELINE 16u                 /* field.rs:93 */
#define AQ_STRENGTH 0.25f                    /* field.rs:99 */
#define AQ_CLAMP 1.0f                        /* field.rs:104 */
#define AQ_EDGE_CONTRAST 6.0f                /* field.rs:108 */
#define AQ_CHROMA_WEIGHT 0.35f               /* field.rs:112 */
#define AQ_VARIANCE_TO_8BIT (255.0f * 255.0f) /* field.rs:117 */
#define MAX_GLOBAL_SCALE (8193u + 65535u)     /* jpxl-encode/src/vardct/ids.rs:20 */
#define MAX_QUANT_LF 65536u                   /* ids.rs:24 */
#define MAX_HF_MUL 2147483647u                /* ids.rs:28 */

/* The activity's f32 log2.  The reference calls the platform's log2f; this
 * is the correctly rounded value (log2 in f64, one rounding), which a
 * faithful libm log2f matches except within an ulp of a rounding boundary.
 * The field's decisions (a half-octave snap of footprint means) move only
 * at exact ties, so that ulp is parity unpinned. */
static float aq_log2f(float x) { return (float)log2((double)x); }

/* analysis.rs:111-158 AnalysisAtlas::analyze: per 8x8 atom and channel the
 * sample variance over the samples that exist, sums in f64 in raster
 * order.  var: 3 planes of bw x bh. */
EXPORT void orc_atlas_variance(int w, int h, const float *xyb[3], float *var) {
    int bw = (w + 7) / 8, bh = (h + 7) / 8;
    for (int ay = 0; ay < bh; ay++)
        for (int ax = 0; ax < bw; ax++) {
            int x0 = ax * 8, y0 = ay * 8, x1 = x0 + 8 < w ? x0 + 8 : w, y1 = y0 + 8 < h ? y0 + 8 : h;
            double count = (double)(x1 - x0) * (double)(y1 - y0);
            for (int c = 0; c < 3; c++) {
                double sum = 0.0, sum_sq = 0.0;
                for (int y = y0; y < y1; y++)
                    for (int x = x0; x < x1; x++) {
                        double v = (double)xyb[c][(size_t)y * w + x];
                        sum += v;
                        sum_sq += v * v;
                    }
                double mean = count > 0.0 ? sum / count : 0.0;
                double variance = count > 0.0 ? sum_sq / count - mean * mean : 0.0;
                if (!(variance > 0.0)) variance = 0.0; /* f64::max(0.0) */
                var[(size_t)c * bw * bh + (size_t)ay * bw + ax] = (float)variance;
            }
        }
}

/* field.rs:200-219 erode_min_3x3 (clamped at the frame edge). */
static void erode_min_3x3(const float *a, int w, int h, float *out) {
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float m = INFINITY;
            for (int ny = y > 0 ? y - 1 : 0; ny < (y + 2 < h ? y + 2 : h); ny++)
                for (int nx = x > 0 ? x - 1 : 0; nx < (x + 2 < w ? x + 2 : w); nx++) {
                    float v = a[ny * w + nx];
                    m = fminf(m, v);
                }
            out[y * w + x] = isfinite(m) ? m : 0.0f;
        }
}

/* field.rs:383-395 mul_of: baseline * 2^-adj rounded to the nearest integer,
 * clamped to [1, 2 * baseline]. */
EXPORT uint32_t orc_aq_mul_of(uint32_t baseline, float adj) {
    double v = round((double)baseline * exp2((double)-adj));
    uint32_t mul = !(v > 0.0) ? 0u : (v >= 4294967295.0 ? 4294967295u : (uint32_t)v);
    uint32_t hi = baseline > 0x7fffffffu ? 0xffffffffu : baseline * 2u;
    if (mul < 1u) mul = 1u;
    if (mul > hi) mul = hi;
    return (mul >= 1u && mul <= MAX_HF_MUL) ? mul : baseline;
}

/* field.rs:224-300 DesiredQuantField::from_atlas_tuned at the default tuning
 * (field.rs:129-158), and field.rs:312-320 is_neutral.  adj: bw x bh.
 * Returns 0 when there is no field (AQ_OFF) or it is neutral (AqSetup::build
 * then treats it as off, lib.rs:1868-1875), else 1; *fine = the lattice. */
EXPORT int orc_aq_field(int w, int h, const float *xyb[3], int mode, float *adj, int *fine) {
    int bw = (w + 7) / 8, bh = (h + 7) / 8, count = bw * bh;
    float sign;
    int is_fine, edge = mode == AQ_EDGE_REFINE;
    switch (mode) {
        case AQ_MASKING: sign = 1.0f; is_fine = 0; break;
        case AQ_UNIFORM: sign = -1.0f; is_fine = 0; break;
        case AQ_FINE_MASKING: sign = 1.0f; is_fine = 1; break;
        case AQ_FINE_UNIFORM: sign = -1.0f; is_fine = 1; break;
        case AQ_EDGE_REFINE: sign = -1.0f; is_fine = 1; break;
        default: *fine = 0; return 0;
    }
    *fine = is_fine;
    float *var = malloc(sizeof(float) * 3 * count), *act = malloc(sizeof(float) * count);
    float *ero = malloc(sizeof(float) * count);
    orc_atlas_variance(w, h, xyb, var);
    for (int i = 0; i < count; i++) {
        float vx = var[i], vy = var[count + i], vb = var[2 * count + i];
        float luma = aq_log2f(1.0f + vy * AQ_VARIANCE_TO_8BIT);
        float chroma = aq_log2f(1.0f + (vx + vb) * AQ_VARIANCE_TO_8BIT);
        act[i] = luma + AQ_CHROMA_WEIGHT * chroma;
    }
    float k = sign * AQ_STRENGTH;
    if (edge) {
        erode_min_3x3(act, bw, bh, ero);
        for (int i = 0; i < count; i++) {
            float d = act[i] - ero[i] - AQ_EDGE_CONTRAST;
            adj[i] = k * (d > 0.0f ? d : 0.0f);
        }
    } else {
        const float *a = act;
        if (is_fine) { erode_min_3x3(act, bw, bh, ero); a = ero; }
        double sum = 0.0;
        for (int i = 0; i < count; i++) sum += (double)a[i];
        float mean = count > 0 ? (float)(sum / (double)count) : 0.0f;
        for (int i = 0; i < count; i++) {
            float v = k * (a[i] - mean);
            adj[i] = v < -AQ_CLAMP ? -AQ_CLAMP : (v > AQ_CLAMP ? AQ_CLAMP : v);
        }
    }
    free(var); free(act); free(ero);
    int neutral = 1;
    for (int i = 0; i < count && neutral; i++) {
        if (is_fine) neutral = orc_aq_mul_of(AQ_FINE_BASELINE, adj[i]) == AQ_FINE_BASELINE;
        else neutral = roundf(adj[i] * 2.0f) == 0.0f;
    }
    return neutral ? 0 : 1;
}

/* field.rs:338-378 mul_for_footprint: the log-mean of the footprint's
 * adjustments (f32, rows then columns, zero outside the grid), snapped to
 * the half-octave lattice (legacy) or clamped to one octave (fine), capped
 * at zero under refine-only. */
EXPORT uint32_t orc_aq_footprint_mul(const float *adj, int bw, int bh, int bx, int by, int rows, int cols,
                                     uint32_t baseline, int refine_only, int fine) {
    if (!adj) return baseline;
    float sum = 0.0f;
    uint32_t count = 0;
    for (int dy = 0; dy < rows; dy++)
        for (int dx = 0; dx < cols; dx++) {
            int x = bx + dx, y = by + dy;
            sum += (x < bw && y < bh) ? adj[(size_t)y * bw + x] : 0.0f;
            count++;
        }
    float a = count > 0 ? sum / (float)count : 0.0f;
    if (fine) a = a < -1.0f ? -1.0f : (a > 1.0f ? 1.0f : a);
    else a = roundf(a * 2.0f) / 2.0f;
    if (refine_only) a = fminf(a, 0.0f);
    return orc_aq_mul_of(baseline, a);
}

/* lib.rs:1859-1908 AqSetup::build's factorization: global_scale / factor
 * (snapped down to a multiple), quant_lf and the baseline HfMul times the
 * factor (2 legacy, 16 fine); refine-only with the request's values when a
 * factored value is not representable.  Returns refine_only. */
EXPORT int orc_aq_setup(uint32_t gs, uint32_t quant_lf, uint32_t hf_mul, int fine, uint32_t *gs_out,
                        uint32_t *qlf_out, uint32_t *base_out) {
    uint32_t factor = fine ? AQ_FINE_BASELINE : 2u;
    uint32_t snapped = gs - gs % factor;
    uint64_t g2 = snapped / factor;
    if (g2 < 1) g2 = 1;
    uint64_t q2 = (uint64_t)quant_lf * factor, m2 = (uint64_t)hf_mul * factor;
    if (q2 > 0xffffffffu) q2 = 0xffffffffu;
    if (m2 > 0xffffffffu) m2 = 0xffffffffu;
    int ok = snapped >= factor && g2 >= 1 && g2 <= MAX_GLOBAL_SCALE && q2 >= 1 && q2 <= MAX_QUANT_LF && m2 >= 1 &&
             m2 <= MAX_HF_MUL;
    if (ok) {
        *gs_out = (uint32_t)g2; *qlf_out = (uint32_t)q2; *base_out = (uint32_t)m2;
        return 0;
    }
    *gs_out = gs; *qlf_out = quant_lf; *base_out = hf_mul;
    return 1;
}

/* lib.rs:4569-4577 mul_signal_bits and lib.rs:1933-1938: 8 bits for a
 * non-baseline multiplier on the legacy lattice, a flat 2 on the fine one. */
static double aq_signal_bits(const orc_quant_params *p, uint32_t mul) {
    if (!p->aq_adj) return 0.0;
    if (p->aq_fine) return 2.0;
    return mul == p->hf_mul ? 0.0 : 8.0;
}

static uint32_t aq_vb_mul(const orc_quant_params *p, int bw, int bh, int bx, int by, int n) {
    return orc_aq_footprint_mul(p->aq_adj, bw, bh, bx, by, n, n, p->hf_mul, p->aq_refine_only, p->aq_fine);
}

/* The (transform, HfMul) quantizers of one frame, built on first use. */
typedef struct { int side; uint32_t mul; hf_quantizer q; } hq_entry;
typedef struct { hq_entry *e; int n, cap; const orc_quant_params *p; } hq_cache;
static const hf_quantizer *hq_get(hq_cache *c, int side, uint32_t mul) {
    for (int i = 0; i < c->n; i++)
        if (c->e[i].side == side && c->e[i].mul == mul) return &c->e[i].q;
    if (c->n == c->cap) { c->cap = c->cap ? 2 * c->cap : 8; c->e = realloc(c->e, sizeof(hq_entry) * c->cap); }
    hq_entry *e = &c->e[c->n++];
    e->side = side; e->mul = mul;
    build_hf_quantizer_mul(&e->q, side, c->p, mul);
    return &e->q;
}
static void hq_free(hq_cache *c) { free(c->e); c->e = NULL; c->n = c->cap = 0; }

/* ------------------------------------------------------------------------ */
/* Frame planning: forward DCT + quantize (lib.rs:3009-3384, :4469)          */
/* ------------------------------------------------------------------------ */

/*
 * Block map: `bmap[by*bw + bx]` holds the varblock side (8, 16 or 32) at a
 * varblock's top-left 8x8 block and 0 on the blocks it covers.
 * Coefficients use the footprint layout: a varblock at block (bx,by) with
 * side s stores cell `k` of channel c at plane_c[(8*by + k/s)*stride + 8*bx +
 * k%s], stride = 8*bw.  LF planes are bw x bh.
 * cfl_x/cfl_b: per 64x64 tile HF factors (tiles_w = ceil(bw/8)); lf_x/lf_b:
 * frame LF factors (already unbiased, i.e. x_factor_lf - 128).
 */

/* lib.rs:3009-3057 gather_square: edge-replicated side x side window. */
static void gather_square(const float *plane, int w, int h, int x0, int y0, int side, float *out) {
    for (int dy = 0; dy < side; dy++)
        for (int dx = 0; dx < side; dx++) {
            int x = x0 + dx < w - 1 ? x0 + dx : w - 1;
            int y = y0 + dy < h - 1 ? y0 + dy : h - 1;
            out[dy * side + dx] = plane[(size_t)y * w + x];
        }
}

/* lib.rs:3059-3098 forward_square for every varblock of the map. */
EXPORT void orc_forward_frame(int w, int h, const float *xyb[3], const uint8_t *bmap,
                              float *coeff[3]) {
    int bw = (w + 7) / 8, bh = (h + 7) / 8, stride = 8 * bw;
    float samples[1024], out[1024];
    for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
            int side = bmap[by * bw + bx];
            if (!side) continue;
            for (int c = 0; c < 3; c++) {
                gather_square(xyb[c], w, h, 8 * bx, 8 * by, side, samples);
                orc_dct2d_fwd(side, side, samples, out);
                for (int k = 0; k < side * side; k++)
                    coeff[c][(size_t)(8 * by + k / side) * stride + 8 * bx + k % side] = out[k];
            }
        }
}

/* lib.rs:3237-3384 quantize_square_varblock + lib.rs:3162-3231 varblock_cfl,
 * over the whole map.  Returns 0 on success, -1 when the reference would
 * reject a coefficient (clamp_round). */
EXPORT int orc_quantize_frame(int w, int h, const uint8_t *bmap, const float *coeff[3],
                              const orc_quant_params *p, const int8_t *cfl_x,
                              const int8_t *cfl_b, int lf_x, int lf_b, int32_t *quant[3],
                              int32_t *lf[3]) {
    int bw = (w + 7) / 8, bh = (h + 7) / 8, stride = 8 * bw, tw = (bw + 7) / 8;
    hq_cache hc = {NULL, 0, 0, p};
    float mdc[3];
    lf_multipliers(p->global_scale, p->quant_lf, mdc);
    float k_x_lf = cfl_k(0.0f, lf_x), k_b_lf = cfl_k(1.0f, lf_b);
    int ok = 1;
    float vb[3][1024], dyh[1024], targets[1024], lfs[16], dylf[16];
    int32_t qv[3][1024];
    for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
            int side = bmap[by * bw + bx];
            if (!side) continue;
            int n = side / 8, cells = side * side;
            const hf_quantizer *q = hq_get(&hc, side, aq_vb_mul(p, bw, bh, bx, by, n));
            for (int c = 0; c < 3; c++)
                for (int k = 0; k < cells; k++)
                    vb[c][k] = coeff[c][(size_t)(8 * by + k / side) * stride + 8 * bx + k % side];
            int tile = (by / 8) * tw + bx / 8;
            float k_x_hf = cfl_k(0.0f, cfl_x[tile]), k_b_hf = cfl_k(1.0f, cfl_b[tile]);
            /* Y first */
            lf_from_llf_square(vb[1], side, n, lfs);
            for (int i = 0; i < n * n; i++) {
                int32_t v = lf_quantize(mdc, lfs[i], 1, &ok);
                lf[1][(by + i / n) * bw + bx + i % n] = v;
                dylf[i] = lf_recon(mdc, v, 1);
            }
            for (int k = 0; k < cells; k++) {
                int x = k % side, y = k / side;
                if (x < n && y < n) { qv[1][k] = 0; dyh[k] = hf_recon(q, 0, 1, k); continue; }
                qv[1][k] = hf_choose(q, vb[1][k], 1, k, &ok);
                dyh[k] = hf_recon(q, qv[1][k], 1, k);
            }
            if (p->truncate) truncate_trailing(q, 1, qv[1], n * n, p->zero_token_bits, vb[1], dyh);
            for (int ci = 0; ci < 2; ci++) {
                int c = ci == 0 ? 0 : 2;
                float k_lf = c == 0 ? k_x_lf : k_b_lf, k_hf = c == 0 ? k_x_hf : k_b_hf;
                lf_from_llf_square(vb[c], side, n, lfs);
                for (int i = 0; i < n * n; i++) {
                    float t = lfs[i] - k_lf * dylf[i];
                    lf[c][(by + i / n) * bw + bx + i % n] = lf_quantize(mdc, t, c, &ok);
                }
                for (int k = 0; k < cells; k++) {
                    int x = k % side, y = k / side;
                    if (x < n && y < n) { targets[k] = 0.0f; qv[c][k] = 0; continue; }
                    targets[k] = vb[c][k] - k_hf * dyh[k];
                    qv[c][k] = hf_choose(q, targets[k], c, k, &ok);
                }
                if (p->truncate) truncate_trailing(q, c, qv[c], n * n, p->zero_token_bits, targets, NULL);
            }
            for (int c = 0; c < 3; c++)
                for (int k = 0; k < cells; k++)
                    quant[c][(size_t)(8 * by + k / side) * stride + 8 * bx + k % side] = qv[c][k];
        }
    hq_free(&hc);
    return ok ? 0 : -1;
}

/* The HfMul of every varblock of `bmap` (at its top-left block, 0 on covered
 * blocks): lib.rs:736-750 (fixed DCT8 cover), :1618-1650 (a given map). */
EXPORT void orc_aq_block_muls(int w, int h, const uint8_t *bmap, const orc_quant_params *p, uint32_t *muls) {
    int bw = (w + 7) / 8, bh = (h + 7) / 8;
    for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
            int side = bmap[by * bw + bx];
            muls[by * bw + bx] = side ? aq_vb_mul(p, bw, bh, bx, by, side / 8) : 0u;
        }
}

/* ------------------------------------------------------------------------ */
/* Render (jpxl-plan-render/src/lib.rs:749-1309)                             */
/* ------------------------------------------------------------------------ */

/* jpxl-core/src/restoration.rs:89-143 GaborKernel::defaults. */
static void gabor_kernel(float *centre, float *edge, float *corner) {
    const float w1 = 0.115169525f, w2 = 0.061248592f;
    float sum = fmaf(4.0f, w2, fmaf(4.0f, w1, 1.0f));
    float scale = 1.0f / sum;
    *centre = scale;
    *edge = w1 * scale;
    *corner = w2 * scale;
}

/* restoration.rs:38-60 mirror1d. */
static int mirror1d(int64_t c, int size) {
    for (;;) {
        if (c < 0) c = -c - 1;
        else if (c >= size) c = 2 * (int64_t)size - 1 - c;
        else return (int)c;
    }
}

/* restoration.rs:148-193 gaborish_into. */
EXPORT void orc_gaborish(int w, int h, const float *in, float *out) {
    float ce, ed, co;
    gabor_kernel(&ce, &ed, &co);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float acc = 0.0f;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    float wgt = (dx == 0 && dy == 0) ? ce : ((dx == 0 || dy == 0) ? ed : co);
                    float s = in[(size_t)mirror1d(y + dy, h) * w + mirror1d(x + dx, w)];
                    acc = fmaf(wgt, s, acc);
                }
            out[(size_t)y * w + x] = acc;
        }
}

/* jpxl-encode/src/vardct/gaborish.rs:66-115 gaborish_forward, as the default
 * build ("simd" feature) evaluates it: border pixels by the scalar 3x3 loop
 * over mirrored samples (w * s added in dy, dx order); interior columns in
 * groups of four from x = 1 as mid*c + (n + s + w + e)*e + (nw + ne + sw + se)*k;
 * the remaining interior columns of a row by the scalar loop.  No fused
 * multiply-add anywhere (Rust does not contract). */
/* gaborish.rs:29-47 GaborKernel::new(w1, w2): unnorm = 1 + 4 w1 + 4 w2
 * evaluated left to right without fused multiply-add. */
static void gabor_kernel_enc(float *centre, float *edge, float *corner) {
    const float w1 = 0.115169525f, w2 = 0.061248592f;
    float a = 4.0f * w1, b = 4.0f * w2;
    float unnorm = (1.0f + a) + b;
    float scale = 1.0f / unnorm;
    *centre = scale;
    *edge = w1 * scale;
    *corner = w2 * scale;
}

static void gabor_forward_enc(int w, int h, const float *in, float *out) {
    float ce, ed, co;
    gabor_kernel_enc(&ce, &ed, &co);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int edge = x == 0 || y == 0 || x + 1 == w || y + 1 == h;
            int simd = 0;
            if (!edge) {
                /* groups x = 1 + 4j while x + 4 <= w - 1 */
                int full = (w - 2) / 4;  /* number of complete groups */
                simd = (x - 1) < 4 * full;
            }
            float acc;
            if (simd) {
                const float *r0 = in + (size_t)(y - 1) * w, *r1 = in + (size_t)y * w, *r2 = in + (size_t)(y + 1) * w;
                float mid = r1[x] * ce;
                float nsew = ((r0[x] + r2[x]) + r1[x - 1]) + r1[x + 1];
                float diag = ((r0[x - 1] + r0[x + 1]) + r2[x - 1]) + r2[x + 1];
                acc = (mid + nsew * ed) + diag * co;
            } else {
                acc = 0.0f;
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        float wgt = (dx == 0 && dy == 0) ? ce : ((dx == 0 || dy == 0) ? ed : co);
                        float sm = in[(size_t)mirror1d(y + dy, h) * w + mirror1d(x + dx, w)];
                        float prod = wgt * sm;
                        acc = acc + prod;
                    }
            }
            out[(size_t)y * w + x] = acc;
        }
}

EXPORT void orc_gaborish_forward_enc(int w, int h, const float *in, float *out) { gabor_forward_enc(w, h, in, out); }

/* gaborish.rs:218-315 precondition_xyb_planes / gaborish_inverse: 12 Jacobi
 * steps x <- x + (f - Kx) * (1/centre) over the flattened plane in groups of
 * four, the last n % 4 samples as x + (f - Kx) / centre. */
EXPORT void orc_gaborish_sharpen(int w, int h, int iters, const float *f, float *x) {
    float ce, ed, co;
    gabor_kernel_enc(&ce, &ed, &co);
    size_t n = (size_t)w * h;
    float *kx = (float *)malloc(n * sizeof(float));
    memcpy(x, f, n * sizeof(float));
    const float inv_c = 1.0f / ce;
    const size_t vec_end = n - n % 4;
    for (int it = 0; it < iters; it++) {
        gabor_forward_enc(w, h, x, kx);
        for (size_t i = 0; i < n; i++) {
            float d = f[i] - kx[i];
            x[i] = i < vec_end ? x[i] + d * inv_c : x[i] + d / ce;
        }
    }
    free(kx);
}

/* restoration.rs:201-256 EPF parameters (Table J.1 defaults). */
static const float EPF_SHARP_LUT[8] = {0.0f, 1.0f / 7.0f, 2.0f / 7.0f, 3.0f / 7.0f,
                                       4.0f / 7.0f, 5.0f / 7.0f, 6.0f / 7.0f, 1.0f};
static const float EPF_CHANNEL_SCALE[3] = {40.0f, 5.0f, 3.5f};

/* restoration.rs:583-731 epf_step_rows over the whole frame.  step: 0, 1, 2. */
EXPORT void orc_epf_step(int step, int w, int h, const float *in[3], const float *sigma,
                         float *out[3]) {
    static const int K13[13][2] = {{0, 0}, {-1, 0}, {1, 0}, {0, -1}, {0, 1}, {1, -1}, {1, 1},
                                   {-1, 1}, {-1, -1}, {-2, 0}, {2, 0}, {0, 2}, {0, -2}};
    static const int CROSS[5][2] = {{0, 0}, {-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    int ntaps = step == 0 ? 13 : 5;
    const int(*taps)[2] = step == 0 ? K13 : CROSS;
    int ncross = step == 2 ? 1 : 5;
    float scale = step == 0 ? 0.9f : (step == 1 ? 1.0f : 6.5f);
    float step_mul = 1.65f * scale;
    const float border_mul = 2.0f / 3.0f;
    int bw = (w + 7) / 8;
#define PX(c, xx, yy) in[c][(size_t)mirror1d((yy), h) * w + mirror1d((xx), w)]
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t idx = (size_t)y * w + x;
            float s = sigma[(y / 8) * bw + x / 8];
            if (s < 0.3f) {
                for (int c = 0; c < 3; c++) out[c][idx] = in[c][idx];
                continue;
            }
            float inv_sigma = step_mul * 4.0f * (1.0f - sqrtf(0.5f)) / s;
            int rx = x % 8, ry = y % 8;
            int at_border = ry == 0 || ry == 7 || rx == 0 || rx == 7;
            float sum_w = 0.0f, acc[3] = {0.0f, 0.0f, 0.0f};
            for (int t = 0; t < ntaps; t++) {
                int kx = taps[t][0], ky = taps[t][1];
                float d = 0.0f;
                for (int c = 0; c < 3; c++)
                    for (int j = 0; j < ncross; j++) {
                        int ix = CROSS[j][0], iy = CROSS[j][1];
                        float a = PX(c, x + ix, y + iy);
                        float b = PX(c, x + kx + ix, y + ky + iy);
                        d = fmaf(fabsf(a - b), EPF_CHANNEL_SCALE[c], d);
                    }
                float pm = at_border ? border_mul : 1.0f;
                float sd = pm * d;
                float wgt;
                if (sd == 0.0f) wgt = 1.0f;
                else {
                    float v = fmaf(sd, -inv_sigma, 1.0f);
                    wgt = v > 0.0f ? v : 0.0f;
                }
                sum_w += wgt;
                for (int c = 0; c < 3; c++) acc[c] = fmaf(PX(c, x + kx, y + ky), wgt, acc[c]);
            }
            for (int c = 0; c < 3; c++) out[c][idx] = acc[c] / sum_w;
        }
#undef PX
}

typedef struct {
    uint32_t global_scale, hf_mul, quant_lf;
    int x_qm, b_qm;
    int epf_iters;   /* 0..3 */
    int gaborish;    /* 0/1 */
    int sharpness;   /* uniform EPF sharpness 0..7 */
    int bits;        /* output depth for linear_at_depth */
    const float *aq_adj; /* per-atom AQ adjustment, NULL = hf_mul everywhere */
    int aq_refine_only, aq_fine;
} orc_render_params;

/* lib.rs:749-1309 render_inner with linear_at_depth: dequant, CfL, LLF, IDCT,
 * scatter, Gaborish, EPF, XYB^-1, depth.  xyb_out (optional) receives the
 * restored XYB planes before the colour stage. */
EXPORT void orc_render(int w, int h, const uint8_t *bmap, const int32_t *quant[3],
                       const int32_t *lfq[3], const int8_t *cfl_x, const int8_t *cfl_b, int lf_x,
                       int lf_b, const orc_render_params *p, float *lin_out[3],
                       float *xyb_out[3]) {
    int bw = (w + 7) / 8, bh = (h + 7) / 8, stride = 8 * bw, tw = (bw + 7) / 8;
    size_t n = (size_t)w * h;
    float *planes[3], *aux[3];
    for (int c = 0; c < 3; c++) {
        planes[c] = calloc(n, sizeof(float));
        aux[c] = calloc(n, sizeof(float));
    }
    float *sigma = calloc((size_t)bw * bh, sizeof(float));
    float qm[3] = {qm_multiplier(p->x_qm), 1.0f, qm_multiplier(p->b_qm)};
    float mdc[3];
    lf_multipliers(p->global_scale, p->quant_lf, mdc);
    float kx_lf = cfl_k(0.0f, lf_x), kb_lf = cfl_k(1.0f, lf_b);
    float *lfp[3];
    for (int c = 0; c < 3; c++) lfp[c] = malloc(sizeof(float) * bw * bh);
    for (int i = 0; i < bw * bh; i++) {
        /* lib.rs:1294-1320 lf_planes: lf_dequantize then cfl_apply. */
        float dx = mdc[0] * (float)lfq[0][i] / 1.0f;
        float dy = mdc[1] * (float)lfq[1][i] / 1.0f;
        float db = mdc[2] * (float)lfq[2][i] / 1.0f;
        lfp[0][i] = dx + kx_lf * dy;
        lfp[1][i] = dy;
        lfp[2][i] = db + kb_lf * dy;
    }
    float mats[3][3][1024];
    for (int s = 0; s < 3; s++)
        for (int c = 0; c < 3; c++) orc_dequant_matrix(8 << s, c, mats[s][c]);
    float co[3][1024], smp[1024], lfr[16], llf[16];
    for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
            int side = bmap[by * bw + bx];
            if (!side) continue;
            int si = side == 8 ? 0 : (side == 16 ? 1 : 2), nb = side / 8, cells = side * side;
            /* lib.rs:876-877: the varblock's own HfMul */
            uint32_t vmul = orc_aq_footprint_mul(p->aq_adj, bw, bh, bx, by, nb, nb, p->hf_mul, p->aq_refine_only,
                                                 p->aq_fine);
            float mul = hf_multiplier(p->global_scale, vmul);
            for (int c = 0; c < 3; c++) {
                float scale = mul * qm[c];
                for (int k = 0; k < cells; k++) {
                    int32_t v = quant[c][(size_t)(8 * by + k / side) * stride + 8 * bx + k % side];
                    co[c][k] = bias_adjust(v, QUANT_BIAS[c]) * scale * mats[si][c][k];
                }
            }
            int tile = (by / 8) * tw + bx / 8;
            float kx = cfl_k(0.0f, cfl_x[tile]), kb = cfl_k(1.0f, cfl_b[tile]);
            for (int k = 0; k < cells; k++) {
                float dyv = co[1][k];
                co[0][k] = co[0][k] + kx * dyv;
                co[2][k] = co[2][k] + kb * dyv;
            }
            for (int c = 0; c < 3; c++) {
                for (int dy = 0; dy < nb; dy++)
                    for (int dx = 0; dx < nb; dx++) {
                        int lx = bx + dx, ly = by + dy;
                        lfr[dy * nb + dx] = (lx < bw && ly < bh) ? lfp[c][ly * bw + lx] : 0.0f;
                    }
                llf_from_lf_square(lfr, nb, llf);
                for (int y = 0; y < nb; y++)
                    for (int x = 0; x < nb; x++) co[c][y * side + x] = llf[y * nb + x];
                orc_dct2d_inv(side, side, co[c], smp);
                for (int r = 0; r < side; r++) {
                    int fy = 8 * by + r;
                    if (fy >= h) continue;
                    for (int q = 0; q < side; q++) {
                        int fx = 8 * bx + q;
                        if (fx >= w) continue;
                        planes[c][(size_t)fy * w + fx] = smp[r * side + q];
                    }
                }
            }
            int sh = p->sharpness < 7 ? p->sharpness : 7;
            for (int dy = 0; dy < nb; dy++)
                for (int dx = 0; dx < nb; dx++) {
                    int sx = bx + dx, sy = by + dy;
                    if (sx >= bw || sy >= bh) continue;
                    sigma[sy * bw + sx] = mul * 0.46f * EPF_SHARP_LUT[sh];
                }
        }
    if (p->gaborish) {
        for (int c = 0; c < 3; c++) orc_gaborish(w, h, planes[c], aux[c]);
        for (int c = 0; c < 3; c++) { float *t = planes[c]; planes[c] = aux[c]; aux[c] = t; }
    }
    int steps[3], ns = 0;
    if (p->epf_iters == 1) { steps[0] = 1; ns = 1; }
    else if (p->epf_iters == 2) { steps[0] = 1; steps[1] = 2; ns = 2; }
    else if (p->epf_iters >= 3) { steps[0] = 0; steps[1] = 1; steps[2] = 2; ns = 3; }
    for (int i = 0; i < ns; i++) {
        const float *src[3] = {planes[0], planes[1], planes[2]};
        orc_epf_step(steps[i], w, h, src, sigma, aux);
        for (int c = 0; c < 3; c++) { float *t = planes[c]; planes[c] = aux[c]; aux[c] = t; }
    }
    if (xyb_out)
        for (int c = 0; c < 3; c++) memcpy(xyb_out[c], planes[c], sizeof(float) * n);
    int levels = p->bits >= 16 ? 65536 : (1 << p->bits);
    float *thr = malloc(sizeof(float) * levels), *lut = malloc(sizeof(float) * levels);
    orc_depth_tables(p->bits, thr, lut);
    for (size_t i = 0; i < n; i++) {
        float o[3];
        opsin_inverse(planes[0][i], planes[1][i], planes[2][i], o);
        for (int c = 0; c < 3; c++) lin_out[c][i] = linear_at_depth(o[c], thr, lut, levels);
    }
    free(thr); free(lut); free(sigma);
    for (int c = 0; c < 3; c++) { free(planes[c]); free(aux[c]); free(lfp[c]); }
}

/* ------------------------------------------------------------------------ */
/* SSIMULACRA2 (jpxl-perceptual)                                             */
/* ------------------------------------------------------------------------ */

/* blur.rs:37-50 recursion constants for sigma 1.5. */
static const double MUL_IN[3] = {0.05529523572608661, -0.05883668702694998,
                                 0.012955819110517063};
static const double MUL_PREV[3] = {1.902113032590307, 1.1755705045849463,
                                   1.2246467991473532e-16};

/* blur.rs:155-174 step, fed sum = l + r rounded to f32. */
static float blur_step(float sum, double *prev, double *prev2) {
    double s = (double)sum;
    double o[3];
    for (int k = 0; k < 3; k++) o[k] = s * MUL_IN[k] + MUL_PREV[k] * prev[k] - prev2[k];
    for (int k = 0; k < 3; k++) { prev2[k] = prev[k]; prev[k] = o[k]; }
    return (float)(o[0] + o[1] + o[2]);
}

/* blur.rs:404-424 horizontal_padded and blur.rs:590-620 vertical_strip: the
 * recursion over one line of samples (stride `stride`), zero padded. */
static void blur_line(const float *in, size_t stride, int len, float *out, size_t ostride) {
    double prev[3] = {0, 0, 0}, prev2[3] = {0, 0, 0};
    for (int nn = 1 - 5; nn < len; nn++) {
        int t = nn - 6, b = nn + 4;
        float top = (t >= 0 && t < len) ? in[t * stride] : 0.0f;
        float bot = (b >= 0 && b < len) ? in[b * stride] : 0.0f;
        float v = blur_step(top + bot, prev, prev2);
        if (nn >= 0) out[nn * ostride] = v;
    }
}

/* blur.rs:124-148 Blur::blur_input: horizontal then vertical pass.  When b
 * is non-NULL the input is the product a*b rounded to f32. */
EXPORT void orc_blur(int w, int h, const float *a, const float *b, float *out) {
    size_t n = (size_t)w * h;
    float *in = malloc(sizeof(float) * n), *tmp = malloc(sizeof(float) * n);
    for (size_t i = 0; i < n; i++) in[i] = b ? a[i] * b[i] : a[i];
    for (int y = 0; y < h; y++) blur_line(in + (size_t)y * w, 1, w, tmp + (size_t)y * w, 1);
    for (int x = 0; x < w; x++) blur_line(tmp + x, (size_t)w, h, out + x, (size_t)w);
    free(in);
    free(tmp);
}

/* pyramid.rs:19-44 downscale_by_2. */
static void downscale_by_2(const float *in, int w, int h, float *out) {
    int ow = (w + 1) / 2, oh = (h + 1) / 2;
    for (int oy = 0; oy < oh; oy++) {
        int y0 = 2 * oy < h - 1 ? 2 * oy : h - 1, y1 = 2 * oy + 1 < h - 1 ? 2 * oy + 1 : h - 1;
        for (int ox = 0; ox < ow; ox++) {
            int x0 = 2 * ox < w - 1 ? 2 * ox : w - 1, x1 = 2 * ox + 1 < w - 1 ? 2 * ox + 1 : w - 1;
            float a = in[(size_t)y0 * w + x0], b = in[(size_t)y0 * w + x1];
            float c = in[(size_t)y1 * w + x0], d = in[(size_t)y1 * w + x1];
            out[(size_t)oy * ow + ox] = (a + b + c + d) * 0.25f;
        }
    }
}

/* ssimulacra2.rs:391-500 WEIGHTS. */
static const double WEIGHTS[108] = {
    0.0, 0.0007376606707406586, 0.0, 0.0, 0.0007793481682867309, 0.0, 0.0, 0.0004371155730107379, 0.0,
    1.1041726426657346, 0.00066284834129271, 0.00015231632783718752, 0.0, 0.0016406437456599754, 0.0,
    1.8422455520539298, 11.441172603757666, 0.0, 0.0007989109436015163, 0.000176816438078653, 0.0,
    1.8787594979546387, 10.94906990605142, 0.0, 0.0007289346991508072, 0.9677937080626833, 0.0,
    0.00014003424285435884, 0.9981766977854967, 0.00031949755934435053, 0.0004550992113792063, 0.0,
    0.0, 0.0013648766163243398, 0.0, 0.0, 0.0, 0.0, 0.0, 7.466890328078848, 0.0, 17.445833984131262,
    0.0006235601634041466, 0.0, 0.0, 6.683678146179332, 0.00037724407979611296, 1.027889937768264,
    225.20515300849274, 0.0, 0.0, 19.213238186143016, 0.0011401524586618361, 0.001237755635509985,
    176.39317598450694, 0.0, 0.0, 24.43300999870476, 0.28520802612117757, 0.0004485436923833408,
    0.0, 0.0, 0.0, 34.77906344483772, 44.835625328877896, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
    0.0008680556573291698, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0005313191874358747, 0.0,
    0.00016533814161379112, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0004179171803251336, 0.0017290828234722833,
    0.0, 0.0020827005846636437, 0.0, 0.0, 8.826982764996862, 23.19243343998926, 0.0,
    95.1080498811086, 0.9863978034400682, 0.9834382792465353, 0.0012286405048278493,
    171.2667255897307, 0.9807858872435379, 0.0, 0.0, 0.0, 0.0005130064588990679, 0.0,
    0.00010854057858411537};

/* ssimulacra2.rs:531-540 remap. */
EXPORT double orc_remap(double error) {
    double s = error * 0.9562382616834844;
    s = 6.248496625763138e-5 * s * s * s + 2.326765642916932 * s - 0.020884521182843837 * s * s;
    if (s > 0.0) return 100.0 - 10.0 * pow(s, 0.6276336467831387);
    return 100.0;
}

/* pool.rs:243-270 accumulate_scalar over one plane set, f64 sums. */
static void pool_maps(size_t n, const float *i1, const float *m1, const float *v11,
                      const float *i2, const float *m2, const float *v22, const float *v12,
                      double *sums /* 6 */) {
    const float C2 = 0.0009f;
    for (size_t i = 0; i < n; i++) {
        float mu11 = m1[i] * m1[i], mu22 = m2[i] * m2[i], mu12 = m1[i] * m2[i];
        float md = m1[i] - m2[i];
        float num_m = 1.0f - md * md;
        float num_s = 2.0f * (v12[i] - mu12) + C2;
        float den_s = (v11[i] - mu11) + (v22[i] - mu22) + C2;
        double d = 1.0 - (double)((num_m * num_s) / den_s);
        if (d < 0.0) d = 0.0;
        sums[0] += d;
        sums[1] += d * d * d * d;
        double d1 = (1.0 + (double)fabsf(i2[i] - m2[i])) / (1.0 + (double)fabsf(i1[i] - m1[i])) - 1.0;
        double art = d1 > 0.0 ? d1 : 0.0;
        sums[2] += art;
        sums[3] += art * art * art * art;
        double lost = -d1 > 0.0 ? -d1 : 0.0;
        sums[4] += lost;
        sums[5] += lost * lost * lost * lost;
    }
}

EXPORT void orc_pool_maps(size_t n, const float *i1, const float *m1, const float *v11,
                          const float *i2, const float *m2, const float *v22, const float *v12,
                          double *sums) {
    pool_maps(n, i1, m1, v11, i2, m2, v22, v12, sums);
}

/* pool.rs:60-68 ChannelTerms::from_sums: [1-norm, 4-norm] per map, written
 * in (norm, map) order: ssim, artifact, detail lost. */
EXPORT void orc_from_sums(const double *sums, size_t pixels, double *out) {
    double inv = 1.0 / (double)(pixels ? pixels : 1);
    for (int m = 0; m < 3; m++) {
        out[m] = inv * sums[2 * m];
        out[3 + m] = sqrt(sqrt(inv * sums[2 * m + 1]));
    }
}

/* ssimulacra2.rs:509-528 weighted_error: the weights are consumed in
 * (component, scale, norm, map) order over the ns scales that exist. */
EXPORT double orc_weighted_error(const double *terms, int ns) {
    double err = 0.0;
    int wi = 0;
    for (int c = 0; c < 3; c++)
        for (int sc = 0; sc < ns; sc++)
            for (int k = 0; k < 6; k++) err += WEIGHTS[wi++] * fabs(terms[c * ns * 6 + sc * 6 + k]);
    return err;
}

/* ssimulacra2.rs:79-135 Ssimulacra2::score with reference.rs:90-200
 * PrecomputedReference::new (Moments retention).  terms[c*ns*6 + s*6 + n*3 +
 * map] holds the pooled terms; returns the score, writes scale count. */
EXPORT double orc_ssimulacra2(int w, int h, const float *ref[3], const float *cand[3],
                              double *terms, int *nscales, double *raw_error) {
    int cw = w, ch = h, s = 0;
    float *rr[3], *cr[3];
    for (int c = 0; c < 3; c++) {
        rr[c] = malloc(sizeof(float) * w * h);
        cr[c] = malloc(sizeof(float) * w * h);
        memcpy(rr[c], ref[c], sizeof(float) * w * h);
        memcpy(cr[c], cand[c], sizeof(float) * w * h);
    }
    double all[6][3][6];
    for (s = 0; s < 6; s++) {
        if (s > 0) {
            if (cw < 8 || ch < 8) break;
            int nw = (cw + 1) / 2, nh = (ch + 1) / 2;
            for (int c = 0; c < 3; c++) {
                float *t1 = malloc(sizeof(float) * nw * nh), *t2 = malloc(sizeof(float) * nw * nh);
                downscale_by_2(rr[c], cw, ch, t1);
                downscale_by_2(cr[c], cw, ch, t2);
                free(rr[c]); free(cr[c]);
                rr[c] = t1; cr[c] = t2;
            }
            cw = nw; ch = nh;
        }
        size_t n = (size_t)cw * ch;
        float *x1[3], *x2[3];
        for (int c = 0; c < 3; c++) { x1[c] = malloc(4 * n); x2[c] = malloc(4 * n); }
        for (size_t i = 0; i < n; i++) {
            positive_xyb(rr[0][i], rr[1][i], rr[2][i], &x1[0][i], &x1[1][i], &x1[2][i]);
            positive_xyb(cr[0][i], cr[1][i], cr[2][i], &x2[0][i], &x2[1][i], &x2[2][i]);
        }
        float *m1 = malloc(4 * n), *v11 = malloc(4 * n), *m2 = malloc(4 * n), *v22 = malloc(4 * n),
              *v12 = malloc(4 * n);
        for (int c = 0; c < 3; c++) {
            orc_blur(cw, ch, x1[c], NULL, m1);
            orc_blur(cw, ch, x1[c], x1[c], v11);
            orc_blur(cw, ch, x2[c], NULL, m2);
            orc_blur(cw, ch, x2[c], x2[c], v22);
            orc_blur(cw, ch, x1[c], x2[c], v12);
            double sums[6] = {0, 0, 0, 0, 0, 0};
            /* pooled in 32-row bands, bands folded in order (streamed.rs). */
            for (int y0 = 0; y0 < ch; y0 += 32) {
                int rows = ch - y0 < 32 ? ch - y0 : 32;
                double band[6] = {0, 0, 0, 0, 0, 0};
                size_t o = (size_t)y0 * cw, bn = (size_t)rows * cw;
                pool_maps(bn, x1[c] + o, m1 + o, v11 + o, x2[c] + o, m2 + o, v22 + o, v12 + o, band);
                for (int k = 0; k < 6; k++) sums[k] += band[k];
            }
            orc_from_sums(sums, n, all[s][c]);
        }
        free(m1); free(v11); free(m2); free(v22); free(v12);
        for (int c = 0; c < 3; c++) { free(x1[c]); free(x2[c]); }
    }
    int ns = s;
    for (int c = 0; c < 3; c++) { free(rr[c]); free(cr[c]); }
    double tl[108];
    for (int c = 0; c < 3; c++)
        for (int sc = 0; sc < ns; sc++)
            for (int k = 0; k < 6; k++) tl[c * ns * 6 + sc * 6 + k] = all[sc][c][k];
    if (terms) memcpy(terms, tl, sizeof(double) * 18 * ns);
    double err = orc_weighted_error(tl, ns);
    if (nscales) *nscales = ns;
    if (raw_error) *raw_error = err;
    return orc_remap(err);
}

/* ------------------------------------------------------------------------ */
/* Scalar rules exported for known-answer tests                              */
/* ------------------------------------------------------------------------ */
EXPORT float orc_hf_multiplier(uint32_t gs, uint32_t hf_mul) { return hf_multiplier(gs, hf_mul); }
EXPORT float orc_qm_multiplier(int qm) { return qm_multiplier(qm); }
EXPORT float orc_bias_adjust(int32_t q, float bias) { return bias_adjust(q, bias); }
EXPORT void orc_lf_multipliers(uint32_t gs, uint32_t quant_lf, float *m) { lf_multipliers(gs, quant_lf, m); }
EXPORT int32_t orc_hf_choose(float target, float step, int channel) {
    hf_quantizer *q = calloc(1, sizeof(hf_quantizer));
    q->side = 8;
    q->step[channel][1] = step;
    q->thr[channel][1] = 0.5f * QUANT_BIAS[channel] * step;
    int ok = 1;
    int32_t v = hf_choose(q, target, channel, 1, &ok);
    free(q);
    return ok ? v : INT32_MIN;
}
EXPORT float orc_hf_recon(int32_t v, float step, int channel) { return bias_adjust(v, QUANT_BIAS[channel]) * step; }

/* ------------------------------------------------------------------------ */
/* Hierarchical cover (jpxl-encode-policy/src/lib.rs:4685-5302)              */
/* ------------------------------------------------------------------------ */

/* lib.rs:3794 residual_bits. */
static uint64_t residual_bits(int32_t q) {
    if (q == 0) return 0;
    return (uint64_t)(32 - __builtin_clz((uint32_t)abs(q))) + 1;
}

/* request.rs:585-636 CoverRateModel::Calibrated (the quality and
 * target-rate requests) and ::Legacy (EncodeRequest::defaults). */
static double cover_rate_scale(const orc_quant_params *p, int side) {
    if (p->cover_legacy) return 1.0;
    return side == 16 ? 1.63 : (side == 32 ? 1.88 : 1.56);
}
static double cover_fixed_bits(const orc_quant_params *p, int side) {
    if (p->cover_legacy) return side == 8 ? 2.0 : 34.0;
    return side == 16 ? 8.4 + 4.0 : (side == 32 ? 6.5 + 4.0 : 5.5);
}

typedef struct {
    int w, h, bw, bh;
    const float *xyb[3];
    hq_cache hc;
    const orc_quant_params *p;
    double lambda[3];
    float donor[3][1024];
} cover_ctx;

/* lib.rs:4909-5062 block_cost_bounded without the cutoff (the cutoff only
 * prunes candidates whose full cost could not win), with lib.rs:4685-4890
 * score_channel_lanes for Y, then X and B against the reconstructed dY. */
static double block_cost(cover_ctx *cc, int side, int bx, int by, uint32_t mul) {
    int si = side == 8 ? 0 : (side == 16 ? 1 : 2), n = side / 8;
    const hf_quantizer *q = hq_get(&cc->hc, side, mul);
    float smp[1024], co[3][1024], dyh[1024];
    for (int c = 0; c < 3; c++) {
        gather_square(cc->xyb[c], cc->w, cc->h, 8 * bx, 8 * by, side, smp);
        orc_dct2d_fwd(side, side, smp, co[c]);
    }
    uint64_t bits = 0;
    double sse = 0.0;
    double tsd = (double)(side * side) * 1.0;
    int order[3] = {1, 0, 2};
    int ok = 1;
    for (int oi = 0; oi < 3; oi++) {
        int c = order[oi];
        for (int row = 0; row < side; row++) {
            int col0 = row < n ? n : 0;
            for (int col = col0; col < side; col++) {
                int cell = row * side + col;
                float target;
                if (c == 1) target = co[1][cell];
                else if (c == 0) target = co[0][cell] - 0.0f * dyh[cell];
                else target = co[2][cell] - 1.0f * dyh[cell];
                int32_t v = hf_choose(q, target, c, cell, &ok);
                float recon = hf_recon(q, v, c, cell);
                if (c == 1) dyh[cell] = recon;
                bits += residual_bits(v);
                double d = (double)(recon - target);
                sse += cc->lambda[c] * tsd * (double)cc->donor[si][cell] * (d * d);
            }
        }
    }
    return (double)bits * cover_rate_scale(cc->p, side) + sse;
}

/* lib.rs:5115-5223 tile_region_with. */
static double tile_region(cover_ctx *cc, int bx, int by, int size, uint8_t *bmap, int write) {
    if (bx >= cc->bw || by >= cc->bh) return 0.0;
    if (size == 1) {
        uint32_t mul = aq_vb_mul(cc->p, cc->bw, cc->bh, bx, by, 1);
        double cost = block_cost(cc, 8, bx, by, mul) + cover_fixed_bits(cc->p, 8) + aq_signal_bits(cc->p, mul);
        if (write) bmap[by * cc->bw + bx] = 8;
        return cost;
    }
    int half = size / 2;
    double split = 0.0;
    int qs[4][2] = {{bx, by}, {bx + half, by}, {bx, by + half}, {bx + half, by + half}};
    for (int i = 0; i < 4; i++) split += tile_region(cc, qs[i][0], qs[i][1], half, bmap, 0);
    int fits = bx + size <= cc->bw && by + size <= cc->bh;
    int side = size * 8;
    int single = 0;
    double best = split;
    if (fits) {
        uint32_t mul = aq_vb_mul(cc->p, cc->bw, cc->bh, bx, by, size);
        double fixed = cover_fixed_bits(cc->p, side) + aq_signal_bits(cc->p, mul);
        if (!(fixed >= split)) {
            double c = block_cost(cc, side, bx, by, mul) + fixed;
            if (c < split) { single = 1; best = c; }
        }
    }
    if (write) {
        if (single) {
            for (int y = 0; y < size; y++)
                for (int x = 0; x < size; x++)
                    if (by + y < cc->bh && bx + x < cc->bw) bmap[(by + y) * cc->bw + bx + x] = 0;
            bmap[by * cc->bw + bx] = (uint8_t)side;
        } else {
            for (int i = 0; i < 4; i++) tile_region(cc, qs[i][0], qs[i][1], half, bmap, 1);
        }
    }
    return best;
}

/* lib.rs:5260-5302 select_blocks over the frame (one LF group per 2048 px;
 * regions never cross a group, so the frame-level walk is the same). */
EXPORT void orc_cover(int w, int h, const float *xyb[3], const orc_quant_params *p, uint8_t *bmap) {
    cover_ctx cc;
    cc.w = w; cc.h = h; cc.bw = (w + 7) / 8; cc.bh = (h + 7) / 8;
    for (int c = 0; c < 3; c++) cc.xyb[c] = xyb[c];
    cc.hc = (hq_cache){NULL, 0, 0, p};
    cc.p = p;
    /* request.rs:650-690: quantizer-donor cell weights, or flat (every
     * cell 1.0) under the fixed-quantizer request. */
    for (int i = 0; i < 3; i++) {
        orc_donor_weights(8 << i, cc.donor[i]);
        if (p->cover_legacy)
            for (int k = 0; k < (64 << (2 * i)); k++) cc.donor[i][k] = 1.0f;
    }
    for (int c = 0; c < 3; c++) cc.lambda[c] = lambda_for(p, c);
    memset(bmap, 0, (size_t)cc.bw * cc.bh);
    for (int sby = 0; sby < cc.bh; sby += 4)
        for (int sbx = 0; sbx < cc.bw; sbx += 4) tile_region(&cc, sbx, sby, 4, bmap, 1);
    hq_free(&cc.hc);
}

/* ------------------------------------------------------------------------ */
/* Chroma-from-luma estimate (jpxl-encode-policy/src/lib.rs:3437-4014,       */
/* quantize.rs:1599-1717)                                                    */
/* ------------------------------------------------------------------------ */

typedef struct { double s_yy, s_yc, s_cc; } cfl_acc;

static void acc_add(cfl_acc *a, float y, float c) { /* quantize.rs:1628-1634 */
    double yy = (double)y, cc = (double)c;
    a->s_yy += yy * yy;
    a->s_yc += yy * cc;
    a->s_cc += cc * cc;
}

static float cfl_multiplier(float base, int factor) { return base + (float)factor / (float)84; }

/* quantize.rs:1652-1717 best_factor + consider. */
static int best_factor(const cfl_acc *a, float base, int lo, int hi) {
    if (a->s_yy <= 0.0) return 0;
    double k_ls = a->s_yc / a->s_yy;
    double star = (k_ls - (double)base) * 84.0;
    double r = round(star);
    int64_t centre = r >= 9.2e18 ? INT64_MAX : (r <= -9.2e18 ? INT64_MIN : (int64_t)r);
    int have = 0, best = 0;
    double best_e = 0.0;
    for (int i = 0; i < 6; i++) {
        int cand;
        if (i < 5) {
            int64_t d = i - 2, c;
            if (d > 0 && centre > INT64_MAX - d) c = INT64_MAX;      /* saturating_add */
            else if (d < 0 && centre < INT64_MIN - d) c = INT64_MIN;
            else c = centre + d;
            if (c < lo) c = lo;
            if (c > hi) c = hi;
            cand = (int)c;
        } else cand = 0;
        double k = (double)cfl_multiplier(base, cand);
        double e = a->s_cc - 2.0 * k * a->s_yc + k * k * a->s_yy;
        if (!have || e < best_e || (e == best_e && abs(cand) < abs(best))) { have = 1; best = cand; best_e = e; }
    }
    return best;
}

/* quantize.rs:1628-1634 + 1656-1684: the accumulator over n (y, c) pairs
 * and best_factor with colour factor 84. */
EXPORT int orc_cfl_best_factor(int n, const float *y, const float *c, float base, int lo, int hi) {
    cfl_acc a = {0.0, 0.0, 0.0};
    for (int i = 0; i < n; i++) acc_add(&a, y[i], c[i]);
    return best_factor(&a, base, lo, hi);
}

static uint64_t factor_bits(int f) { return f == 0 ? 0 : (uint64_t)(32 - __builtin_clz((uint32_t)abs(f))) + 2; }

/* lib.rs:3771-3783 factor_candidates: seed-1..seed+1 clamped, deduplicated, plus 0. */
static int factor_candidates(int seed, int lo, int hi, int *out) {
    int n = 0;
    for (int d = -1; d <= 1; d++) {
        int c = seed + d;
        if (c < lo) c = lo;
        if (c > hi) c = hi;
        int dup = 0;
        for (int i = 0; i < n; i++) dup |= out[i] == c;
        if (!dup) out[n++] = c;
    }
    int has0 = 0;
    for (int i = 0; i < n; i++) has0 |= out[i] == 0;
    if (!has0) out[n++] = 0;
    return n;
}

typedef struct { float source, ry, dy; int cell; } cfl_sample;
typedef struct { cfl_acc acc; cfl_sample *s; int n, cap; } cfl_set;

static void set_push(cfl_set *s, float source, float y, float dy, int cell) {
    if (s->n == s->cap) { s->cap = s->cap ? 2 * s->cap : 1024; s->s = realloc(s->s, sizeof(cfl_sample) * s->cap); }
    acc_add(&s->acc, y, source);
    s->s[s->n].source = source; s->s[s->n].ry = y; s->s[s->n].dy = dy; s->s[s->n].cell = cell;
    s->n++;
}

/* lib.rs:3899-3955 hf_residual_cost_bounded (no cutoff: it cannot change the winner). */
static uint64_t hf_residual_cost(const cfl_set *s, float base, int factor, int channel, const hf_quantizer *q) {
    float k = cfl_multiplier(base, factor);
    uint64_t bits = 0;
    int ok = 1;
    for (int i = 0; i < s->n; i++) {
        float t = s->s[i].source - k * s->s[i].dy;
        bits += residual_bits(hf_choose(q, t, channel, s->s[i].cell, &ok));
    }
    return bits;
}

/* lib.rs:3966-4014 refine_hf_factor. */
static int refine_hf_factor(const cfl_set *s, float base, int channel, const hf_quantizer *q) {
    if (s->n == 0) return 0;
    int seed = best_factor(&s->acc, base, -128, 127);
    if (seed == 0) return 0;
    int best = 0;
    uint64_t best_cost = hf_residual_cost(s, base, 0, channel, q);
    int cands[4];
    int nc = factor_candidates(seed, -128, 127, cands);
    for (int i = 0; i < nc; i++) {
        int f = cands[i];
        if (f == 0) continue;
        uint64_t cost = hf_residual_cost(s, base, f, channel, q) + factor_bits(f);
        if (cost < best_cost || (cost == best_cost && abs(f) < abs(best))) { best_cost = cost; best = f; }
    }
    return best;
}

static uint64_t lf_residual_cost(const cfl_set *s, float base, int factor, int channel, const float *mdc) {
    float k = cfl_multiplier(base, factor);
    uint64_t bits = 0;
    int ok = 1;
    for (int i = 0; i < s->n; i++)
        bits += residual_bits(lf_quantize(mdc, s->s[i].source - k * s->s[i].dy, channel, &ok));
    return bits;
}

/* lib.rs:3831-3894 refine_lf_factors. */
static void refine_lf_factors(const cfl_set *x, const cfl_set *b, const float *mdc, int *ox, int *ob) {
    int xs = best_factor(&x->acc, 0.0f, -128, 127), bs = best_factor(&b->acc, 1.0f, -128, 127);
    int xc[4], bc[4];
    int nx = factor_candidates(xs, -128, 127, xc), nb = factor_candidates(bs, -128, 127, bc);
    uint64_t xcost[4], bcost[4], neutral_x = 0, neutral_b = 0;
    for (int i = 0; i < nx; i++) { xcost[i] = lf_residual_cost(x, 0.0f, xc[i], 0, mdc); if (xc[i] == 0) neutral_x = xcost[i]; }
    for (int i = 0; i < nb; i++) { bcost[i] = lf_residual_cost(b, 1.0f, bc[i], 2, mdc); if (bc[i] == 0) neutral_b = bcost[i]; }
    int best_x = 0, best_b = 0;
    uint64_t best_cost = neutral_x + neutral_b;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < nb; j++) {
            uint64_t cost = xcost[i] + bcost[j];
            if (xc[i] != 0 || bc[j] != 0) cost += 50 + factor_bits(xc[i]) + factor_bits(bc[j]);
            unsigned mag = (unsigned)abs(xc[i]) + (unsigned)abs(bc[j]);
            unsigned bmag = (unsigned)abs(best_x) + (unsigned)abs(best_b);
            if (cost < best_cost || (cost == best_cost && mag < bmag)) { best_x = xc[i]; best_b = bc[j]; best_cost = cost; }
        }
    *ox = best_x; *ob = best_b;
}

/* lib.rs:3437-3763 estimate_cfl over the whole frame.  LF groups are 256
 * blocks square; within a group, tile rows (bands) in order, varblocks in
 * raster order inside each band. */
EXPORT void orc_estimate_cfl(int w, int h, const uint8_t *bmap, const float *coeff[3],
                             const orc_quant_params *p, int8_t *cfl_x, int8_t *cfl_b, int *lf_x, int *lf_b) {
    int bw = (w + 7) / 8, bh = (h + 7) / 8, stride = 8 * bw, tw = (bw + 7) / 8, th = (bh + 7) / 8;
    hq_cache hc = {NULL, 0, 0, p};
    float mdc[3];
    lf_multipliers(p->global_scale, p->quant_lf, mdc);
    cfl_set lfx = {0}, lfb = {0};
    cfl_set *tx = calloc((size_t)tw * th, sizeof(cfl_set)), *tb = calloc((size_t)tw * th, sizeof(cfl_set));
    int ok = 1;
    int gw = (bw + 255) / 256, gh = (bh + 255) / 256;
    float vb[3][1024], lfs[3][16];
    for (int gy = 0; gy < gh; gy++)
        for (int gx = 0; gx < gw; gx++) {
            int gbx0 = gx * 256, gby0 = gy * 256;
            int gbx1 = gbx0 + 256 < bw ? gbx0 + 256 : bw, gby1 = gby0 + 256 < bh ? gby0 + 256 : bh;
            for (int band = gby0; band < gby1; band += 8)
                for (int by = band; by < band + 8 && by < gby1; by++)
                    for (int bx = gbx0; bx < gbx1; bx++) {
                        int side = bmap[by * bw + bx];
                        if (!side) continue;
                        int n = side / 8, cells = side * side;
                        const hf_quantizer *q = hq_get(&hc, side, aq_vb_mul(p, bw, bh, bx, by, n));
                        for (int c = 0; c < 3; c++)
                            for (int k = 0; k < cells; k++)
                                vb[c][k] = coeff[c][(size_t)(8 * by + k / side) * stride + 8 * bx + k % side];
                        lf_from_llf_square(vb[1], side, n, lfs[1]);
                        lf_from_llf_square(vb[0], side, n, lfs[0]);
                        lf_from_llf_square(vb[2], side, n, lfs[2]);
                        for (int i = 0; i < n * n; i++) {
                            float y = lfs[1][i];
                            float dy = lf_recon(mdc, lf_quantize(mdc, y, 1, &ok), 1);
                            set_push(&lfx, lfs[0][i], y, dy, 0);
                            set_push(&lfb, lfs[2][i], y, dy, 0);
                        }
                        int tile = (by / 8) * tw + bx / 8, fold = side / 8;
                        for (int row = 0; row < side; row++) {
                            int c0 = row < n ? n : 0;
                            for (int col = c0; col < side; col++) {
                                int cell = row * side + col;
                                int cell8 = (cell / side / fold) * 8 + (cell % side / fold);
                                float y = vb[1][cell];
                                float dy = hf_recon(q, hf_choose(q, y, 1, cell, &ok), 1, cell);
                                set_push(&tx[tile], vb[0][cell], y, dy, cell8);
                                set_push(&tb[tile], vb[2][cell], y, dy, cell8);
                            }
                        }
                    }
        }
    refine_lf_factors(&lfx, &lfb, mdc, lf_x, lf_b);
    const hf_quantizer *base8 = hq_get(&hc, 8, p->hf_mul);  /* lib.rs:3737 the baseline DCT8 quantizer */
    for (int t = 0; t < tw * th; t++) {
        cfl_x[t] = (int8_t)refine_hf_factor(&tx[t], 0.0f, 0, base8);
        cfl_b[t] = (int8_t)refine_hf_factor(&tb[t], 1.0f, 2, base8);
        free(tx[t].s); free(tb[t].s);
    }
    free(tx); free(tb); free(lfx.s); free(lfb.s);
    hq_free(&hc);
}

/* ---- first-guess features of the quality search --------------------------- */

/* preanalysis_features (jpxl-encode-policy/src/preanalysis.rs:277-540) over
 * the 8-bit samples preanalysis_srgb8 / preanalysis_srgb16 hand it
 * (:542-574: 9..16-bit samples reduced to their top eight bits).  Returns the
 * number of whole 32x32 cells (0: None, nothing written); out[22] = share[7],
 * log2var mean/std/q10/q90, edge mean/q90, flat mean, noise mean/q90, chroma
 * log2 mean/q90, orientation entropy, direction dominance, axis-aligned
 * share and heterogeneous fraction, as the struct's f32 fields. */
static int cmp_f64(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static double pa_quantile(const double *s, int n, double q) { /* preanalysis.rs:243-268 */
    if (n == 0) return 0.0;
    double pos = q * (double)(n - 1);
    int lo = (int)floor(pos);
    double frac = pos - (double)lo;
    double a = lo < n ? s[lo] : s[n - 1];
    double b = lo + 1 < n ? s[lo + 1] : a;
    return a + (b - a) * frac;
}

EXPORT uint32_t orc_preanalysis(int w, int h, int bits, const void *rgb, float *out) {
    int cells_x = w / 32, cells_y = h / 32;
    if (cells_x == 0 || cells_y == 0) return 0;
    size_t px = (size_t)w * h;
    uint8_t *lum = malloc(px), *cbp = malloc(px), *crp = malloc(px);
    int shift = bits - 8;
    for (size_t i = 0; i < px; i++) {
        int s[3];
        for (int c = 0; c < 3; c++) {
            if (bits == 8) s[c] = ((const uint8_t *)rgb)[3 * i + c];
            else {
                int v = ((const uint16_t *)rgb)[3 * i + c] >> shift;
                s[c] = v < 255 ? v : 255;
            }
        }
        int r = s[0], g = s[1], b = s[2];
        lum[i] = (uint8_t)((77 * r + 150 * g + 29 * b + 128) >> 8);
        int cb = ((-43 * r - 85 * g + 128 * b + 128) >> 8) + 128;
        int cr = ((128 * r - 107 * g - 21 * b + 128) >> 8) + 128;
        cbp[i] = (uint8_t)(cb < 0 ? 0 : (cb > 255 ? 255 : cb));
        crp[i] = (uint8_t)(cr < 0 ? 0 : (cr > 255 ? 255 : cr));
    }
#define LUM(xx, yy) ((int)lum[(size_t)((yy) < 0 ? 0 : ((yy) >= h ? h - 1 : (yy))) * w + \
                               ((xx) < 0 ? 0 : ((xx) >= w ? w - 1 : (xx)))])
    int n = cells_x * cells_y;
    int *cls = malloc(sizeof(int) * n), *edge_q8 = malloc(sizeof(int) * n), *flat_q8 = malloc(sizeof(int) * n);
    int *noise = malloc(sizeof(int) * n), *has_edges = malloc(sizeof(int) * n);
    int *ent = malloc(sizeof(int) * n), *dom = malloc(sizeof(int) * n), *axis = malloc(sizeof(int) * n);
    uint32_t *var = malloc(sizeof(uint32_t) * n), *chroma = malloc(sizeof(uint32_t) * n);
    const int64_t N = 1024;
    for (int cy = 0; cy < cells_y; cy++)
        for (int cx = 0; cx < cells_x; cx++) {
            int k = cy * cells_x + cx, x0 = cx * 32, y0 = cy * 32;
            int64_t sum = 0, sum_sq = 0, edge_count = 0, flat_count = 0, noise_sum = 0, weak = 0;
            int64_t dir[4] = {0, 0, 0, 0}, cb_sum = 0, cb_sq = 0, cr_sum = 0, cr_sq = 0;
            for (int y = y0; y < y0 + 32; y++)
                for (int x = x0; x < x0 + 32; x++) {
                    int c = LUM(x, y);
                    sum += c;
                    sum_sq += c * c;
                    int tl = LUM(x - 1, y - 1), tc = LUM(x, y - 1), tr = LUM(x + 1, y - 1);
                    int ml = LUM(x - 1, y), mr = LUM(x + 1, y);
                    int bl = LUM(x - 1, y + 1), bc = LUM(x, y + 1), br = LUM(x + 1, y + 1);
                    int gx = (tr + 2 * mr + br) - (tl + 2 * ml + bl);
                    int gy = (bl + 2 * bc + br) - (tl + 2 * tc + tr);
                    int grad = abs(gx) + abs(gy);
                    if (grad > 48) {
                        edge_count++;
                        int ax = abs(gx), ay = abs(gy);
                        int bin = ax >= 2 * ay ? 0 : (ay >= 2 * ax ? 1 : ((gx > 0) == (gy > 0) ? 2 : 3));
                        dir[bin]++;
                    }
                    if (grad < 12) flat_count++;
                    if (grad < 24) {
                        int box_mean = (tl + tc + tr + ml + c + mr + bl + bc + br) / 9;
                        noise_sum += abs(c - box_mean);
                        weak++;
                    }
                    size_t i = (size_t)y * w + x;
                    int64_t vb = cbp[i], vr = crp[i];
                    cb_sum += vb; cb_sq += vb * vb; cr_sum += vr; cr_sq += vr * vr;
                }
            int64_t mean = sum / N;
            int64_t v = sum_sq / N - mean * mean;
            var[k] = (uint32_t)(v > 0 ? v : 0);
            int64_t cbm = cb_sum / N, crm = cr_sum / N;
            int64_t vcb = cb_sq / N - cbm * cbm, vcr = cr_sq / N - crm * crm;
            int64_t ch = (vcb > 0 ? vcb : 0) + (vcr > 0 ? vcr : 0);
            chroma[k] = (uint32_t)(ch < 65535 ? ch : 65535);
            int64_t edge_total = dir[0] + dir[1] + dir[2] + dir[3];
            int64_t den = edge_total > 1 ? edge_total : 1;
            int64_t dmax = dir[0];
            for (int b = 1; b < 4; b++) dmax = dir[b] > dmax ? dir[b] : dmax;
            edge_q8[k] = (int)((edge_count * 256) / N);
            flat_q8[k] = (int)((flat_count * 256) / N);
            noise[k] = weak > 0 ? (int)(noise_sum / weak) : 0;
            has_edges[k] = edge_total > 0;
            /* entropy4_q8 (preanalysis.rs:192-212) */
            int e = 0;
            if (edge_total > 0) {
                double t = (double)edge_total, hh = 0.0;
                for (int b = 0; b < 4; b++)
                    if (dir[b] > 0) {
                        double p = (double)dir[b] / t;
                        hh -= p * log2(p);
                    }
                double q = round(hh / 2.0 * 256.0);
                e = (int)(q < 0 ? 0 : (q > 256 ? 256 : q));
            }
            ent[k] = e;
            dom[k] = (int)((dmax * 256) / den);
            axis[k] = (int)(((dir[0] + dir[1]) * 256) / den);
            /* classify (preanalysis.rs:214-241) */
            int c;
            if (edge_q8[k] >= 64 && ent[k] < 110 && axis[k] >= 160 && noise[k] < 7) c = 6;
            else if (edge_q8[k] >= 38 && dom[k] >= 128) c = 5;
            else if (noise[k] >= 7 && var[k] < 64) c = 3;
            else if (var[k] >= 64 && ent[k] >= 150) c = 4;
            else if (var[k] < 12 && edge_q8[k] < 13) c = 0;
            else if (edge_q8[k] < 26) c = 1;
            else if (chroma[k] >= 96) c = 2;
            else c = 4;
            cls[k] = c;
        }
#undef LUM
    double count = (double)n;
    float share[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < n; k++) share[cls[k]] += 1.0f;
    for (int i = 0; i < 7; i++) share[i] /= (float)count;
    double *lv = malloc(sizeof(double) * n), *ed = malloc(sizeof(double) * n);
    double *no = malloc(sizeof(double) * n), *chv = malloc(sizeof(double) * n);
    double lv_sum = 0.0, ed_sum = 0.0, fl_sum = 0.0, no_sum = 0.0, ch_sum = 0.0;
    for (int k = 0; k < n; k++) {
        lv[k] = log2(1.0 + (double)var[k]);
        lv_sum += lv[k];
    }
    double lv_mean = lv_sum / count, ss = 0.0, het = 0.0;
    for (int k = 0; k < n; k++) ss += (lv[k] - lv_mean) * (lv[k] - lv_mean);
    for (int k = 0; k < n; k++) het += fabs(lv[k] - lv_mean) > 2.0 ? 1.0 : 0.0;
    for (int k = 0; k < n; k++) { ed[k] = (double)edge_q8[k] / 256.0; ed_sum += ed[k]; }
    for (int k = 0; k < n; k++) fl_sum += (double)flat_q8[k] / 256.0;
    for (int k = 0; k < n; k++) { no[k] = (double)noise[k]; no_sum += no[k]; }
    for (int k = 0; k < n; k++) { chv[k] = log2(1.0 + (double)chroma[k]); ch_sum += chv[k]; }
    qsort(lv, n, sizeof(double), cmp_f64);
    qsort(ed, n, sizeof(double), cmp_f64);
    qsort(no, n, sizeof(double), cmp_f64);
    qsort(chv, n, sizeof(double), cmp_f64);
    double e_sum = 0.0, d_sum = 0.0, a_sum = 0.0;
    int edged = 0;
    for (int k = 0; k < n; k++) if (has_edges[k]) { e_sum += (double)ent[k] / 256.0; edged++; }
    for (int k = 0; k < n; k++) if (has_edges[k]) d_sum += (double)dom[k] / 256.0;
    for (int k = 0; k < n; k++) if (has_edges[k]) a_sum += (double)axis[k] / 256.0;
    for (int i = 0; i < 7; i++) out[i] = share[i];
    out[7] = (float)lv_mean;
    out[8] = (float)sqrt(ss / count);
    out[9] = (float)pa_quantile(lv, n, 0.1);
    out[10] = (float)pa_quantile(lv, n, 0.9);
    out[11] = (float)(ed_sum / count);
    out[12] = (float)pa_quantile(ed, n, 0.9);
    out[13] = (float)(fl_sum / count);
    out[14] = (float)(no_sum / count);
    out[15] = (float)pa_quantile(no, n, 0.9);
    out[16] = (float)(ch_sum / count);
    out[17] = (float)pa_quantile(chv, n, 0.9);
    out[18] = edged ? (float)(e_sum / (double)edged) : 0.0f;
    out[19] = edged ? (float)(d_sum / (double)edged) : 0.0f;
    out[20] = edged ? (float)(a_sum / (double)edged) : 0.0f;
    out[21] = (float)(het / count);
    free(lum); free(cbp); free(crp); free(cls); free(edge_q8); free(flat_q8); free(noise); free(has_edges);
    free(ent); free(dom); free(axis); free(var); free(chroma); free(lv); free(ed); free(no); free(chv);
    return (uint32_t)n;
}

/* quality_transform_summary (jpxl-encode-policy/src/lib.rs:2769-2955): the
 * aligned DCT8x8 of every block, LF groups (2048-sample squares) in raster
 * order and blocks in raster order within each, with the reference's
 * running sums.  out[11] = blocks, ln_ac_y_mean, ln_ac_y_q50, _q90, _q99,
 * high_low_ratio, directional_asymmetry, chroma_ac_ratio, near_zero_frac_1e3,
 * near_zero_frac_1e2, dc_variance_y; hist[512] = the ln-energy histogram. */
EXPORT void orc_transform_summary(int w, int h, const float *xyb[3], double *out, uint32_t *hist) {
    const double EPS = 1e-30, HIST_LO = -46.0, HIST_WIDTH = 0.125;
    const int BINS = 512, LFG = 2048;
    for (int b = 0; b < BINS; b++) hist[b] = 0;
    uint64_t blocks = 0, ac_cells = 0, near3 = 0, near2 = 0;
    double total_y = 0, total_low = 0, total_high = 0, total_row = 0, total_col = 0, total_xb = 0;
    double dc_sum = 0, dc_sumsq = 0;
    float samples[64], co[3][64];
    int gx_n = (w + LFG - 1) / LFG, gy_n = (h + LFG - 1) / LFG;
    for (int gy = 0; gy < gy_n; gy++)
        for (int gx = 0; gx < gx_n; gx++) {
            /* one group's partial, combined in group order */
            uint64_t p_blocks = 0, p_ac = 0, p3 = 0, p2 = 0;
            double py_ = 0, plo = 0, phi = 0, prow = 0, pcol = 0, pxb = 0, pdc = 0, pdq = 0;
            int rx0 = gx * LFG, ry0 = gy * LFG;
            int rw = (w - rx0 < LFG ? w - rx0 : LFG), rh = (h - ry0 < LFG ? h - ry0 : LFG);
            int gbw = (rw + 7) / 8, gbh = (rh + 7) / 8;
            for (int by = 0; by < gbh; by++)
                for (int bx = 0; bx < gbw; bx++) {
                    int px = rx0 + 8 * bx, py = ry0 + 8 * by;
                    for (int c = 0; c < 3; c++) {
                        gather_square(xyb[c], w, h, px, py, 8, samples);
                        orc_dct2d_fwd(8, 8, samples, co[c]);
                    }
                    double block_y = 0.0;
                    for (int cell = 0; cell < 64; cell++) {
                        double v = (double)co[1][cell];
                        if (cell == 0) { pdc += v; pdq += v * v; continue; }
                        int row = cell / 8, col = cell % 8;
                        double energy = v * v;
                        block_y += energy;
                        if (row + col <= 2) plo += energy;
                        if ((row > col ? row : col) >= 4) phi += energy;
                        if (row > col) prow += energy;
                        else if (col > row) pcol += energy;
                        double m = fabs(v);
                        if (m < 1e-3) p3++;
                        if (m < 1e-2) p2++;
                        p_ac++;
                    }
                    for (int c = 0; c < 3; c += 2)
                        for (int cell = 1; cell < 64; cell++) pxb += (double)co[c][cell] * (double)co[c][cell];
                    py_ += block_y;
                    p_blocks++;
                    double bin = (log(block_y + EPS) - HIST_LO) / HIST_WIDTH;
                    bin = bin < 0.0 ? 0.0 : (bin > (double)(BINS - 1) ? (double)(BINS - 1) : bin);
                    hist[(int)bin]++;
                }
            blocks += p_blocks; ac_cells += p_ac; total_y += py_; total_low += plo; total_high += phi;
            total_row += prow; total_col += pcol; total_xb += pxb; near3 += p3; near2 += p2;
            dc_sum += pdc; dc_sumsq += pdq;
        }
    double q_out[3];
    const double qs[3] = {0.5, 0.9, 0.99};
    for (int qi = 0; qi < 3; qi++) {
        double wf = qs[qi] * (double)blocks, cap = (double)(blocks ? blocks - 1 : 0);
        uint64_t want = (uint64_t)(wf < cap ? wf : cap), seen = 0;
        q_out[qi] = HIST_LO + ((double)BINS - 0.5) * HIST_WIDTH;
        for (int b = 0; b < BINS; b++) {
            seen += hist[b];
            if (seen > want) { q_out[qi] = HIST_LO + ((double)b + 0.5) * HIST_WIDTH; break; }
        }
    }
    double nb = blocks > 1 ? (double)blocks : 1.0;
    double acn = ac_cells > 1 ? (double)ac_cells : 1.0;
    out[0] = (double)blocks;
    out[1] = log(total_y / nb + EPS);
    out[2] = q_out[0];
    out[3] = q_out[1];
    out[4] = q_out[2];
    out[5] = total_high / (total_low + EPS);
    out[6] = fabs(total_row - total_col) / (total_row + total_col + EPS);
    out[7] = total_xb / (total_y + EPS);
    out[8] = (double)near3 / acn;
    out[9] = (double)near2 / acn;
    double mean = dc_sum / nb, dv = dc_sumsq / nb - mean * mean;
    out[10] = dv > 0.0 ? dv : 0.0;
}
