// Host-side construction of the constant tables the kernels read, and of the
// per-probe quantizer scalars.  Everything here runs once per context or once
// per probe on the CPU with the platform libm (glibc), exactly as JPXL builds
// the same tables, and is uploaded to HBM.
#include "tables.h"

#include <math.h>
#include <string.h>

#include <algorithm>

namespace jpxlg {

namespace {

float srgb_to_linear(float v) {  // jpxl-core/src/color.rs:517-528
  float a = fabsf(v);
  float d = a <= 12.92f * 0.0031308f ? a / 12.92f : powf((a + 0.055f) / 1.055f, 2.4f);
  return v < 0.0f ? -d : d;
}

float linear_to_srgb(float v) {  // jpxl-core/src/color.rs:490-501
  float a = fabsf(v);
  float e = a <= 0.0031308f ? 12.92f * a : 1.055f * powf(a, 1.0f / 2.4f) - 0.055f;
  return v < 0.0f ? -e : e;
}

// Table I.6 dct_params of the three square transforms
// (jpxl-core/src/dequant.rs:824-897).
const double kParams8[3][6] = {{3150.0, 0.0, -0.4, -0.4, -0.4, -2.0},
                               {560.0, 0.0, -0.3, -0.3, -0.3, -0.3},
                               {512.0, -2.0, -1.0, 0.0, -1.0, -2.0}};
const double kParams16[3][7] = {
    {8996.8725711814115328, -1.3000777393353804, -0.49424529824571225, -0.439093774457103443,
     -0.6350101832695744, -0.90177264050827612, -1.6162099239887414},
    {3191.48366296844234752, -0.67424582104194355, -0.80745813428471001, -0.44925837484843441,
     -0.35865440981033403, -0.31322389111877305, -0.37615025315725483},
    {1157.50408145487200256, -2.0531423165804414, -1.4, -0.50687130033378396,
     -0.42708730624733904, -1.4856834539296244, -4.9209142884401604}};
const double kParams32[3][8] = {
    {15718.40830982518931456, -1.025, -0.98, -0.9012, -0.4, -0.48819395464, -0.421064, -0.27},
    {7305.7636810695983104, -0.8041958212306401, -0.7633036457487539, -0.55660379990111464,
     -0.49785304658857626, -0.43699592683512467, -0.40180866526242109, -0.27321683125358037},
    {3803.53173721215041536, -3.060733579805728, -2.0413270132490346, -2.0235650159727417,
     -0.5495389509954993, -0.4, -0.4, -0.3}};

// I.2.4 GetDctQuantWeights and its reciprocal (dequant.rs:400-510, :321-338).
void dequant_matrix(int side, int channel, float* out) {
  const double* p = side == 8 ? kParams8[channel] : (side == 16 ? kParams16[channel] : kParams32[channel]);
  const int np = side == 8 ? 6 : (side == 16 ? 7 : 8);
  double bands[8];
  bands[0] = p[0];
  for (int i = 1; i < np; i++) bands[i] = bands[i - 1] * (p[i] > 0.0 ? 1.0 + p[i] : 1.0 / (1.0 - p[i]));
  const double max = M_SQRT2 + 1e-6;
  const double span = (double)(side - 1);
  const int last = np - 1;
  for (int y = 0; y < side; y++)
    for (int x = 0; x < side; x++) {
      double dx = (double)x / span, dy = (double)y / span;
      double dist = sqrt(dx * dx + dy * dy);
      double sp = dist * (double)last / max;
      double si = floor(sp);
      int idx = (int)si;
      double a = bands[idx], b = bands[idx + 1];
      double w = a * pow(b / a, sp - si);
      out[y * side + x] = (float)(1.0 / w);
    }
}

// I.3.2 natural coefficient order for a square (varblock.rs:710-745).
void natural_order(int side, uint16_t* order) {
  const int cx = side / 8;
  struct Key { long k1, k2; int lin; };
  std::vector<Key> hf;
  int nl = 0;
  for (int y = 0; y < side; y++)
    for (int x = 0; x < side; x++) {
      int lin = y * side + x;
      if (x < cx && y < cx) { order[y * cx + x] = (uint16_t)lin; nl++; continue; }
      long sx = x, sy = y;  // scale / cx == 1 for squares
      long k1 = sx + sy, k2 = sx - sy;
      if (k1 % 2 == 1) k2 = -k2;
      hf.push_back({k1, k2, lin});
    }
  std::stable_sort(hf.begin(), hf.end(), [](const Key& a, const Key& b) {
    return a.k1 != b.k1 ? a.k1 < b.k1 : a.k2 < b.k2;
  });
  for (size_t i = 0; i < hf.size(); i++) order[nl + i] = (uint16_t)hf[i].lin;
}

// csf.rs:95-105 radial_frequency.
double radial(int u, int v, int side) {
  double fu = u, fv = v, n = side;
  return sqrt((fu / n) * (fu / n) + (fv / n) * (fv / n)) / M_SQRT2;
}

// csf.rs:221-316 quant_donor_weights (donor: the DCT8x8 Y matrix).
void donor_weights(int side, float* weights) {
  float m8[64];
  dequant_matrix(8, 1, m8);
  std::vector<std::pair<double, double>> pts;
  for (int cell = 1; cell < 64; cell++) {
    int u = cell / 8, v = cell % 8;
    double step = (double)m8[u * 8 + v];
    pts.push_back({radial(u, v, 8), 1.0 / (step * step)});
  }
  std::stable_sort(pts.begin(), pts.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  const int llf = side / 8, cells = side * side;
  std::vector<double> raw(cells, 0.0);
  double sum = 0.0;
  int count = 0;
  for (int cell = 0; cell < cells; cell++) {
    int u = cell / side, v = cell % side;
    weights[cell] = 1.0f;
    if (u < llf && v < llf) continue;
    double f = radial(u, v, side), w = pts.back().second;
    if (f <= pts.front().first) {
      w = pts.front().second;
    } else if (f < pts.back().first) {
      for (size_t i = 0; i + 1 < pts.size(); i++) {
        double f0 = pts[i].first, w0 = pts[i].second, f1 = pts[i + 1].first, w1 = pts[i + 1].second;
        if (f1 >= f) {
          w = f1 == f0 ? w1 : w0 + (w1 - w0) * ((f - f0) / (f1 - f0));
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
}

// The sixteen DCT-IV_16 magnitudes as the reference's table holds them (four
// of them one ulp away from the nearest float, so they are written as bits).
const uint32_t kMag16Bits[16] = {0x3eb4cd21u, 0x3eb30f61u, 0x3eaf982cu, 0x3eaa700cu,
                                 0x3ea3a3b9u, 0x3e9b43f5u, 0x3e916567u, 0x3e862061u,
                                 0x3e73215cu, 0x3e57aa8bu, 0x3e3a2005u, 0x3e1aca9eu,
                                 0x3df3ef2fu, 0x3dafefbbu, 0x3d547d0au, 0x3c8e1d7au};

// sqrt(2/M) cos(pi (2n+1)(2u+1) / 4M) as +/- one of M magnitudes.
void fold_iv(int M, const float* mags, float* out) {
  for (int u = 0; u < M; u++)
    for (int i = 0; i < M; i++) {
      int a = ((2 * i + 1) * (2 * u + 1)) % (8 * M);
      float s = 1.0f;
      if (a > 4 * M) a = 8 * M - a;
      if (a > 2 * M) { a = 4 * M - a; s = -1.0f; }
      out[u * M + i] = s * mags[(a - 1) / 2];
    }
}

// varblock.rs:789-811 scale_f_closed_form, argument = varblock dimension.
float scale_f(int c, int lf_count) {
  double cf = c, bf = lf_count * 8;
  double inv = cos(cf * M_PI / (2.0 * bf)) * cos(cf * M_PI / bf) * cos(2.0 * cf * M_PI / bf);
  return (float)(1.0 / inv);
}

// reconstruct.rs:38-43: 0.8f.powi(qm - 2) as compiler-rt's __powisf2.
float powi_f(float a, int b) {
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

}  // namespace

void natural_order_table(int side, uint16_t* order) { natural_order(side, order); }

float qm_multiplier(uint32_t qm_scale) { return powi_f(0.8f, (int)qm_scale - 2); }

float hf_multiplier(uint32_t gs, uint32_t hf_mul) {  // reconstruct.rs:27-36
  double denom = (double)gs * (double)hf_mul;
  return denom <= 0.0 ? 0.0f : (float)(65536.0 / denom);
}

void build_dct_mats(DctMats* m) {
  float* fwd[3] = {m->fwd8, m->fwd16, m->fwd32};
  float* inv[3] = {m->inv8, m->inv16, m->inv32};
  for (int si = 0; si < 3; si++) {
    const int n = kSides[si];
    for (int k = 0; k < n; k++)
      for (int x = 0; x < n; x++) {
        // orthonormal DCT-II basis D[k][x]
        const double d = (k == 0 ? sqrt(1.0 / n) : sqrt(2.0 / n)) * cos(M_PI * (2 * x + 1) * k / (2.0 * n));
        fwd[si][k * n + x] = (float)(d / sqrt((double)n));  // L = D / sqrt(N)
        inv[si][x * n + k] = (float)(d * sqrt((double)n));  // L = sqrt(N) D^T
      }
  }
}

void build_const_tables(ConstTables* t) {
  memset(t, 0, sizeof(*t));
  t->dct.cos_pi_8 = 0.9238795f;
  t->dct.sin_pi_8 = 0.38268343f;
  float m4[4], m8[8], m16[16];
  for (int m = 0; m < 4; m++) m4[m] = (float)(sqrt(2.0 / 4) * cos(M_PI * (2 * m + 1) / 16.0));
  for (int m = 0; m < 8; m++) m8[m] = (float)(sqrt(2.0 / 8) * cos(M_PI * (2 * m + 1) / 32.0));
  for (int m = 0; m < 16; m++) memcpy(&m16[m], &kMag16Bits[m], 4);
  fold_iv(4, m4, t->dct.iv4);
  fold_iv(8, m8, t->dct.iv8);
  fold_iv(16, m16, t->dct.iv16);
  for (int si = 0; si < 3; si++) {
    int side = kSides[si];
    for (int c = 0; c < 3; c++) dequant_matrix(side, c, t->matrix[c] + kCellBase[si]);
    donor_weights(side, t->donor + kCellBase[si]);
    natural_order(side, t->order + kCellBase[si]);
    const int cb = kCellBase[si];
    for (int pos = 0; pos < side * side; pos++) {
      const int cell = t->order[cb + pos];
      for (int c = 0; c < 3; c++) t->matrix_pos[c][cb + pos] = t->matrix[c][cb + cell];
      t->donor_pos[cb + pos] = t->donor[cb + cell];
    }
    int n = side / 8;
    for (int c = 0; c < n; c++) t->scale_f[si][c] = scale_f(c, n);
  }
  // The metric blur's impulse response: the recursion of blur.rs:155-174
  // fed one unit sample, in f64.
  const double mul_in[3] = {0.05529523572608661, -0.05883668702694998, 0.012955819110517063};
  const double mul_prev[3] = {1.902113032590307, 1.1755705045849463, 1.2246467991473532e-16};
  const int len = 21, centre = 10;
  double prev[3] = {0, 0, 0}, prev2[3] = {0, 0, 0}, resp[21] = {0};
  for (int nn = -4; nn < len; nn++) {
    int a = nn - 6, b = nn + 4;
    double s = (a == centre ? 1.0 : 0.0) + (b == centre ? 1.0 : 0.0);
    double o[3];
    for (int k = 0; k < 3; k++) o[k] = s * mul_in[k] + mul_prev[k] * prev[k] - prev2[k];
    for (int k = 0; k < 3; k++) { prev2[k] = prev[k]; prev[k] = o[k]; }
    if (nn >= 0) resp[nn] = o[0] + o[1] + o[2];
  }
  for (int k = 0; k <= 4; k++) t->fir[k] = resp[centre + k];
}

void srgb8_lut(float* lut) {  // source.rs:253-254, evaluator.rs:93-94
  for (int v = 0; v < 256; v++) lut[v] = srgb_to_linear((float)v / 255.0f);
}

void srgb16_lut(int bits, float* lut) {  // source.rs:389-398, evaluator.rs:109-117
  float m = (float)((1u << bits) - 1u);
  float max = 65535.0f < m ? 65535.0f : m;
  float divisor = max > 0.0f ? max : 1.0f;
  for (uint32_t v = 0; v <= 65535u; v++) lut[v] = srgb_to_linear((float)v / divisor);
}

// jpxl-plan-render/src/lib.rs:147-216 + :313-334: bisection of the composed
// depth quantizer's level boundaries over the f32 bit lattice.
int depth_tables(int bits, std::vector<float>* thr, std::vector<float>* lut) {
  const float max = bits >= 32 ? 65535.0f : (float)((1u << bits) - 1u);
  const int levels = bits >= 32 ? 65536 : (1 << std::min(std::max(bits, 1), 16));
  lut->resize(levels);
  thr->resize(levels - 1);
  for (int q = 0; q < levels; q++) (*lut)[q] = srgb_to_linear((float)q / max);
  auto level = [&](float x) -> uint32_t {
    float scaled = roundf(linear_to_srgb(x) * max);
    if (!isfinite(scaled)) return 0;
    scaled = std::min(std::max(scaled, 0.0f), max);
    return (uint32_t)scaled;
  };
  uint32_t two;
  float f2 = 2.0f;
  memcpy(&two, &f2, 4);
  for (int q = 1; q < levels; q++) {
    uint32_t lo = 0, hi = two;
    while (lo < hi) {
      uint32_t mid = lo + (hi - lo) / 2;
      float fm;
      memcpy(&fm, &mid, 4);
      if (level(fm) >= (uint32_t)q) hi = mid;
      else lo = mid + 1;
    }
    memcpy(&(*thr)[q - 1], &lo, 4);
  }
  return levels;
}

// quantize.rs:301-366, reconstruct.rs:27-76, lib.rs:2081-2109, restoration.rs:313
void build_probe_params(const jpxlg_quantizer& q, int sharpness, ProbeParams* p) {
  memset(p, 0, sizeof(*p));
  const float mul = hf_multiplier(q.global_scale, q.hf_mul);
  const float qm[3] = {qm_multiplier(q.x_qm_scale), 1.0f, qm_multiplier(q.b_qm_scale)};
  for (int c = 0; c < 3; c++) p->sc[c] = c == 1 ? mul : mul * qm[c];
  // The Lagrange multiplier from the DCT8x8 baseline steps.
  float m8[64];
  const double scale = (isfinite(q.lambda_scale) && q.lambda_scale > 0.0f) ? (double)q.lambda_scale : 1.0;
  for (int c = 0; c < 3; c++) {
    dequant_matrix(8, c, m8);
    double sum = 0.0;
    for (int cell = 1; cell < 64; cell++) {
      double step = 8.0 * (double)(p->sc[c] * m8[cell]);
      sum += step * step;
    }
    double mean = sum / 63.0;
    p->lambda[c] = mean > 0.0 ? scale * 16.0 / mean : 0.0;
  }
  const float w[3] = {1.0f / 32.0f, 1.0f / 4.0f, 1.0f / 2.0f};
  const double denom = (double)q.global_scale * (double)q.quant_lf;
  for (int c = 0; c < 3; c++) p->mdc[c] = denom > 0.0 ? (float)(65536.0 * (double)(w[c] / 128.0f) / denom) : 0.0f;
  p->mul = mul;
  const float lut[8] = {0.0f, 1.0f / 7.0f, 2.0f / 7.0f, 3.0f / 7.0f, 4.0f / 7.0f, 5.0f / 7.0f, 6.0f / 7.0f, 1.0f};
  p->sig_lut = lut[std::min(std::max(sharpness, 0), 7)];
  p->sigma = mul * 0.46f * p->sig_lut;
  p->gs = q.global_scale;
  p->aq_baseline = q.hf_mul;
  for (int c = 0; c < 3; c++) p->qm[c] = qm[c];
  p->zero_token_bits = q.zero_token_bits;
  p->truncate = q.truncate ? 1 : 0;
  set_cover_objective(p, false);
}

void set_cover_objective(ProbeParams* p, bool legacy) {
  // CoverRateModel::Calibrated (the quality and target-rate requests) or
  // ::Legacy (EncodeRequest::defaults); request.rs:585-636.
  const double cal_rate[3] = {1.56, 1.63, 1.88}, cal_fixed[3] = {5.5, 8.4 + 4.0, 6.5 + 4.0};
  const double leg_rate[3] = {1.0, 1.0, 1.0}, leg_fixed[3] = {2.0, 34.0, 34.0};
  for (int i = 0; i < 3; i++) {
    p->cover_rate[i] = legacy ? leg_rate[i] : cal_rate[i];
    p->cover_fixed[i] = legacy ? leg_fixed[i] : cal_fixed[i];
  }
  p->cover_flat = legacy ? 1 : 0;
}

}  // namespace jpxlg
