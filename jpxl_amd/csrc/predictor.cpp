// The quality search's first guess (jpxl-encode-policy/src/
// quality_features.rs, preanalysis.rs, quality_prediction.rs and the
// routing of quality.rs:2485-2530).  Host code: the device supplies the
// per-atom variances, the 32x32 cell sums and the DCT8 summary partials; the
// order statistics, the cell classes, the frame reductions and the
// predictors run here, in the reference's order and with the same libm.
#include "predictor.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <utility>

#include "controller.h"

extern "C" const unsigned char jpxlg_predictor_tables[];
extern "C" const unsigned char jpxlg_predictor_tables_end[];

namespace jpxlg {

namespace {

// ---------------------------------------------------------------------------
// Source features (quality_features.rs)
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Preanalysis (preanalysis.rs)
// ---------------------------------------------------------------------------

// bpg-rs cell thresholds (preanalysis.rs:65-83).
constexpr int kFlatVar = 12, kFlatEdgeQ8 = 13, kGradEdgeQ8 = 26, kEdgeDenseQ8 = 38, kTextEdgeQ8 = 64;
constexpr int kDirDominantQ8 = 128, kAxisAlignedQ8 = 160, kLowEntropyQ8 = 110, kHighEntropyQ8 = 150;
constexpr int kTextureVar = 64, kNoiseHi = 7, kChromaHi = 96;
constexpr double kHeteroBits = 2.0;
constexpr int64_t kCellPixels = 32 * 32;

enum CellClass { kFlat = 0, kGradient, kChromaCritical, kNoisy, kTexture, kDirectionalEdge, kTextLike };

struct Cell {
  int cls;
  uint32_t variance;
  int edge_q8, flat_q8, noise;
  uint32_t chroma;
  bool has_edges;
  int entropy_q8, dominance_q8, axis_q8;
};

// entropy4_q8 (preanalysis.rs:192-212).
int entropy4_q8(const int64_t dir[4], int64_t total) {
  if (total <= 0) return 0;
  const double t = (double)total;
  double h = 0.0;
  for (int i = 0; i < 4; i++)
    if (dir[i] > 0) {
      const double p = (double)dir[i] / t;
      h -= p * log2(p);
    }
  return (int)std::min(std::max(round(h / 2.0 * 256.0), 0.0), 256.0);
}

// classify (preanalysis.rs:214-241), bpg-rs priority order.
int classify(const Cell& c) {
  if (c.edge_q8 >= kTextEdgeQ8 && c.entropy_q8 < kLowEntropyQ8 && c.axis_q8 >= kAxisAlignedQ8 && c.noise < kNoiseHi)
    return kTextLike;
  if (c.edge_q8 >= kEdgeDenseQ8 && c.dominance_q8 >= kDirDominantQ8) return kDirectionalEdge;
  if (c.noise >= kNoiseHi && c.variance < (uint32_t)kTextureVar) return kNoisy;
  if (c.variance >= (uint32_t)kTextureVar && c.entropy_q8 >= kHighEntropyQ8) return kTexture;
  if (c.variance < (uint32_t)kFlatVar && c.edge_q8 < kFlatEdgeQ8) return kFlat;
  if (c.edge_q8 < kGradEdgeQ8) return kGradient;
  if (c.chroma >= (uint32_t)kChromaHi) return kChromaCritical;
  return kTexture;
}

// log2(1 + v) for the integer cell statistics (variance and chroma, both
// below 65536), as the reference evaluates it.
const double* log2_1p_table() {
  static const std::vector<double> t = [] {
    std::vector<double> v(65536);
    for (int i = 0; i < 65536; i++) v[i] = log2(1.0 + (double)i);
    return v;
  }();
  return t.data();
}

// quantile (preanalysis.rs:243-268), linear interpolation between the order
// statistics at floor(q (n-1)) and the next, over values f(key) of integer
// keys that f maps monotonically (so the keys' order is the values').
template <typename F>
double lerp_quantile_keys(std::vector<uint32_t> keys, double q, F f) {
  if (keys.empty()) return 0.0;
  const size_t n = keys.size();
  const double pos = q * (double)(n - 1);
  const size_t lo = (size_t)floor(pos);
  const double frac = pos - (double)lo;
  std::nth_element(keys.begin(), keys.begin() + lo, keys.end());
  const double a = f(keys[lo]);
  double b = a;
  if (lo + 1 < n) b = f(*std::min_element(keys.begin() + lo + 1, keys.end()));
  return a + (b - a) * frac;
}

// ---------------------------------------------------------------------------
// Generated tables (quality_predictor_v2.rs, quality_predictor_cases.rs)
// ---------------------------------------------------------------------------

struct Tables {
  bool ok = false;
  int d2 = 0, knots = 0, d = 0, ncases = 0, nt = 0, neighbours = 0;
  const double* v2_centers = nullptr;
  const double* v2_scales = nullptr;
  const double* v2_zrange = nullptr;   // d2 x (lo, hi)
  const double* v2_knots = nullptr;    // knots x (1 + 6 (d2 + 1))
  double ratio = 0, spread = 0, dist_eps = 0;
  const double* weights = nullptr;
  const double* centers = nullptr;
  const double* scales = nullptr;
  const double* knot_targets = nullptr;
  const double* zrange = nullptr;      // d x (lo, hi)
  const double* cases = nullptr;       // ncases x (d + 2 nt)
};

const Tables& tables() {
  static const Tables t = [] {
    Tables r;
    const unsigned char* p = jpxlg_predictor_tables;
    const size_t n = (size_t)(jpxlg_predictor_tables_end - jpxlg_predictor_tables);
    if (n < 36 || memcmp(p, "JPXLQPT1", 8) != 0) return r;
    uint32_t c[7];
    memcpy(c, p + 8, sizeof(c));
    r.d2 = (int)c[0];
    r.knots = (int)c[1];
    r.d = (int)c[2];
    r.ncases = (int)c[3];
    r.nt = (int)c[4];
    r.neighbours = (int)c[5];
    // the doubles start at byte 36: copy them to aligned storage
    static std::vector<double> vals;
    vals.resize((n - 36) / 8);
    memcpy(vals.data(), p + 36, vals.size() * 8);
    const double* q = vals.data();
    const double* end = q + vals.size();
    auto take = [&](size_t k) {
      const double* s = q;
      q += k;
      return s;
    };
    r.v2_centers = take(r.d2);
    r.v2_scales = take(r.d2);
    r.v2_zrange = take(2 * r.d2);
    r.v2_knots = take((size_t)r.knots * (1 + 6 * (r.d2 + 1)));
    const double* m = take(3);
    r.ratio = m[0];
    r.spread = m[1];
    r.dist_eps = m[2];
    r.weights = take(r.d);
    r.centers = take(r.d);
    r.scales = take(r.d);
    r.knot_targets = take(r.nt);
    r.zrange = take(2 * r.d);
    r.cases = take((size_t)r.ncases * (r.d + 2 * r.nt));
    r.ok = q == end && r.d2 == 19 && r.d == 44 && r.nt == 7;
    return r;
  }();
  return t;
}

// ---------------------------------------------------------------------------
// Predictors (quality_prediction.rs)
// ---------------------------------------------------------------------------

constexpr double kFeatureEps = 1e-9;              // :23
constexpr double kLossEpsilon = 1e-3;             // :26
constexpr double kFallbackLogWidth = 0.405465;    // :30
constexpr double kFallbackSaturationRisk = 0.5;   // :33
constexpr double kOodRangeMargin = 0.25;          // :37
constexpr uint32_t kMinDomainSide = 128;          // :43
constexpr double kBetaLo = 0.2, kBetaHi = 3.0;    // :47
constexpr int kLog2PixelsFeature = 6;

double loss(double score) { return std::max(100.0 - score, kLossEpsilon); }
double ln_eps(float v) { return log(std::max((double)v, 0.0) + kFeatureEps); }

// qpv2_features (quality_prediction.rs:85-113): the qpv2-st/1 vector.
void st_features(const jpxlg_source_features& f, const jpxlg_transform_summary& t, double* out) {
  const double w = (double)std::max(f.width, 1u), h = (double)std::max(f.height, 1u);
  const double v[19] = {ln_eps(f.luma_variance_q10), ln_eps(f.luma_variance_q50), ln_eps(f.luma_variance_q90),
                        ln_eps(f.chroma_variance_q50), (double)f.flat_fraction, ln_eps(f.edge_proxy),
                        log2(w * h), log(w / h), f.grayscale ? 1.0 : 0.0, t.chroma_ac_ratio, t.dc_variance_y,
                        t.directional_asymmetry, t.high_low_ratio, t.ln_ac_y_mean, t.ln_ac_y_q50, t.ln_ac_y_q90,
                        t.ln_ac_y_q99, t.near_zero_frac_1e2, t.near_zero_frac_1e3};
  memcpy(out, v, sizeof(v));
}

// PreanalysisFeatures::vector (preanalysis.rs:141-172).
void pa_vector(const jpxlg_source_features& f, double* out) {
  const float s[22] = {f.share[0], f.share[1], f.share[2], f.share[3], f.share[4], f.share[5], f.share[6],
                       f.log2var_mean, f.log2var_std, f.log2var_q10, f.log2var_q90, f.edge_mean, f.edge_q90,
                       f.flat_mean, f.noise_mean, f.noise_q90, f.chroma_log2_mean, f.chroma_log2_q90,
                       f.orient_entropy_mean, f.dir_dominance_mean, f.axis_aligned_mean, f.hetero_frac};
  for (int i = 0; i < 22; i++) out[i] = (double)s[i];
  out[22] = log((double)f.edge_mean + 1e-4);
  out[23] = log(1.0 - (double)f.share[0] + 1e-3);
  out[24] = log((double)f.noise_mean + 0.1);
}

uint32_t rung_top() { return Rung::top(); }

// rung_ceil / rung_floor (quality_prediction.rs:196-231).
uint32_t rung_floor(double scale) {
  if (!std::isfinite(scale)) return rung_top();
  const double c = std::min(std::max(scale, 1.0), 9.0e18);
  return Rung::for_effective_scale((uint64_t)c);
}
uint32_t rung_ceil(double scale) {
  if (!std::isfinite(scale)) return rung_top();
  const uint32_t fl = rung_floor(scale);
  if ((double)Rung::effective_scale(fl) < scale) return std::min(fl + 1, rung_top());
  return fl;
}

// The out-of-distribution rule shared by both models: a standardized
// feature beyond its training range by more than a quarter of the span.
void ood_check(const double* z, const double* zrange, const double* weights, int n, uint64_t* flags,
               uint64_t* above) {
  for (int i = 0; i < n; i++) {
    if (weights && !(weights[i] > 0.0)) continue;
    const double lo = zrange[2 * i], hi = zrange[2 * i + 1];
    const double span = std::max(hi - lo, 1e-6);
    if (z[i] < lo - kOodRangeMargin * span || z[i] > hi + kOodRangeMargin * span) {
      *flags |= 1ull << i;
      if (z[i] > hi + kOodRangeMargin * span) *above |= 1ull << i;
    }
  }
}

bool predict_v2(const jpxlg_source_features& f, const jpxlg_transform_summary& t, double target,
                jpxlg_crossing_prediction* out) {
  const Tables& T = tables();
  if (!T.ok || T.knots == 0) return false;
  double raw[19], z[19];
  st_features(f, t, raw);
  for (int i = 0; i < 19; i++) z[i] = (raw[i] - T.v2_centers[i]) / T.v2_scales[i];
  uint64_t flags = 0, above = 0;
  if (std::min(f.width, f.height) < kMinDomainSide) flags |= JPXLG_OOD_TINY_FRAME;
  ood_check(z, T.v2_zrange, nullptr, 19, &flags, &above);
  const uint64_t size_bit = 1ull << kLog2PixelsFeature;
  const bool large_only = flags == size_bit && above == size_bit;
  // evaluate_knot + the monotone projection (max-accumulate over ascending
  // targets) of the four crossing quantiles.
  const int K = T.knots, W = T.d2 + 1, stride = 1 + 6 * W;
  std::vector<std::array<double, 7>> ev(K);
  double run[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int k = 0; k < K; k++) {
    const double* kn = T.v2_knots + (size_t)k * stride;
    auto affine = [&](int field) {
      const double* w = kn + 1 + field * W;
      double acc = w[0];
      for (int i = 0; i < 19; i++) acc += w[1 + i] * z[i];
      return acc;
    };
    std::array<double, 7>& e = ev[k];
    e[0] = log(loss(kn[0]));
    for (int fld = 0; fld < 6; fld++) e[1 + fld] = affine(fld);
    for (int q = 0; q < 4; q++) {
      run[q] = std::max(run[q], e[1 + q]);
      e[1 + q] = run[q];
    }
  }
  const double x = log(loss(target));
  auto interp = [&](int fld) {
    if (x >= ev[0][0]) return ev[0][fld];
    if (x <= ev[K - 1][0]) return ev[K - 1][fld];
    for (int k = 0; k + 1 < K; k++) {
      const std::array<double, 7>& a = ev[k];
      const std::array<double, 7>& b = ev[k + 1];
      if (x <= a[0] && x >= b[0]) {
        const double span = a[0] - b[0];
        if (span <= 2.220446049250313e-16) return a[fld];
        const double tt = (a[0] - x) / span;
        return a[fld] + tt * (b[fld] - a[fld]);
      }
    }
    return ev[K - 1][fld];
  };
  const double ln_lower = interp(1), ln_median = interp(2), ln_candidate = interp(3), ln_upper = interp(4);
  const double ln_beta = interp(5), sat_logit = interp(6);
  const double sat = 1.0 / (1.0 + exp(-std::min(std::max(sat_logit, -30.0), 30.0)));
  const double width = std::max(ln_upper - ln_lower, 0.0);
  int fallback = JPXLG_FALLBACK_NONE;
  if (flags) fallback = JPXLG_FALLBACK_OOD_FEATURE;
  else if (width > kFallbackLogWidth) fallback = JPXLG_FALLBACK_WIDE_INTERVAL;
  else if (sat > kFallbackSaturationRisk) fallback = JPXLG_FALLBACK_SATURATION_RISK;
  out->median_rung = rung_floor(exp(ln_median));
  out->candidate_rung = rung_ceil(exp(ln_candidate));
  out->interval_low = rung_floor(exp(ln_lower));
  out->interval_high = rung_ceil(exp(ln_upper));
  out->local_loss_exponent = std::min(std::max(exp(ln_beta), kBetaLo), kBetaHi);
  out->saturation_risk = sat;
  out->ood_flags = flags;
  out->fallback = fallback;
  out->large_frame_only_ood = large_only ? 1 : 0;
  out->model = JPXLG_PREDICTOR_KNOTS;
  return true;
}

// interpolate_knots (quality_prediction.rs:405-435).
double interpolate_knots(const Tables& T, const double* values, double x) {
  double xs[16];
  for (int i = 0; i < T.nt; i++) xs[i] = log(loss(T.knot_targets[i]));
  if (x >= xs[0]) return values[0];
  if (x <= xs[T.nt - 1]) return values[T.nt - 1];
  for (int i = 0; i + 1 < T.nt; i++) {
    const double xa = xs[i], xb = xs[i + 1];
    if (xb <= x && x <= xa) {
      const double span = xa - xb;
      if (span <= 1e-12) return values[i];
      return values[i] + (xa - x) / span * (values[i + 1] - values[i]);
    }
  }
  return values[T.nt - 1];
}

bool predict_cases(const jpxlg_source_features& f, const jpxlg_transform_summary& t, double target, double reserve,
                   jpxlg_crossing_prediction* out) {
  const Tables& T = tables();
  if (!T.ok || T.ncases == 0 || !f.has_preanalysis) return false;
  double raw[44], z[44];
  st_features(f, t, raw);
  pa_vector(f, raw + 19);
  for (int i = 0; i < 44; i++) z[i] = (raw[i] - T.centers[i]) / T.scales[i];
  // nearest cases in weighted standardized space, ties by table order
  const int stride = T.d + 2 * T.nt;
  std::vector<std::pair<double, int>> ranked(T.ncases);
  for (int c = 0; c < T.ncases; c++) {
    const double* cz = T.cases + (size_t)c * stride;
    double s = 0.0;
    for (int i = 0; i < 44; i++) {
      const double delta = T.weights[i] * (cz[i] - z[i]);
      s += delta * delta;
    }
    ranked[c] = {sqrt(s), c};
  }
  const size_t keep = (size_t)std::max(T.neighbours, 1);
  std::partial_sort(ranked.begin(), ranked.begin() + std::min(keep, ranked.size()), ranked.end(),
                    [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
                      return a.first < b.first || (a.first == b.first && a.second < b.second);
                    });
  ranked.resize(std::min(keep, ranked.size()));
  if (T.ratio > 0.0) {
    const double nearest = ranked.empty() ? 0.0 : ranked[0].first;
    std::vector<std::pair<double, int>> kept;
    for (auto& r : ranked)
      if (r.first <= nearest * T.ratio + 1e-9) kept.push_back(r);
    ranked.swap(kept);
  }
  const double x_target = log(loss(target));
  const double aim = 100.0 - loss(target) * (1.0 - reserve);
  const double x_aim = log(loss(aim));
  double total = 0.0, ln_median = 0.0, ln_candidate = 0.0, ln_beta = 0.0;
  double ln_low = INFINITY, ln_high = -INFINITY;
  for (auto& r : ranked) {
    const double* cz = T.cases + (size_t)r.second * stride;
    const double* crossing = cz + T.d;
    const double* beta = crossing + T.nt;
    const double w = 1.0 / (r.first + T.dist_eps);
    const double at_target = interpolate_knots(T, crossing, x_target);
    const double at_aim = interpolate_knots(T, crossing, x_aim);
    ln_median += w * at_target;
    ln_candidate += w * at_aim;
    ln_beta += w * interpolate_knots(T, beta, x_target);
    ln_low = std::min(ln_low, at_target);
    ln_high = std::max(ln_high, at_target);
    total += w;
  }
  if (!(total > 0.0)) return false;
  ln_median /= total;
  ln_candidate = std::max(ln_candidate / total, ln_median);
  ln_beta /= total;
  uint64_t flags = 0, above = 0;
  if (std::min(f.width, f.height) < kMinDomainSide) flags |= JPXLG_OOD_TINY_FRAME;
  ood_check(z, T.zrange, T.weights, 44, &flags, &above);
  const uint64_t size_bit = 1ull << kLog2PixelsFeature;
  const bool large_only = flags == size_bit && above == size_bit;
  int fallback = flags ? JPXLG_FALLBACK_OOD_FEATURE : JPXLG_FALLBACK_NONE;
  if (!large_only && ln_high - ln_low > T.spread) {
    flags |= JPXLG_OOD_CASE_SPREAD;
    fallback = JPXLG_FALLBACK_WIDE_INTERVAL;
  }
  out->median_rung = rung_floor(exp(ln_median));
  out->candidate_rung = rung_ceil(exp(ln_candidate));
  out->interval_low = rung_floor(exp(ln_low));
  out->interval_high = rung_ceil(exp(ln_high));
  out->local_loss_exponent = std::min(std::max(exp(ln_beta), kBetaLo), kBetaHi);
  out->saturation_risk = 0.0;
  out->ood_flags = flags;
  out->fallback = fallback;
  out->large_frame_only_ood = large_only ? 1 : 0;
  out->model = JPXLG_PREDICTOR_CASES;
  return true;
}

}  // namespace

void source_features_set(float q10, float q50, float q90, float chroma_q50, uint32_t flat, uint32_t n,
                         jpxlg_source_features* f) {
  f->flat_fraction = n ? (float)flat / (float)n : 0.0f;
  f->luma_variance_q10 = q10;
  f->luma_variance_q50 = q50;
  f->luma_variance_q90 = q90;
  f->chroma_variance_q50 = chroma_q50;
  f->edge_proxy = q90 - q50;
}

void preanalysis_from_cells(const CellStats* cs, int n, jpxlg_source_features* f) {
  f->has_preanalysis = n > 0 ? 1 : 0;
  f->cells = (uint32_t)n;
  if (n <= 0) return;
  std::vector<Cell> cells(n);
  for (int k = 0; k < n; k++) {
    const CellStats& s = cs[k];
    Cell& c = cells[k];
    const int64_t mean = (int64_t)s.sum / kCellPixels;
    c.variance = (uint32_t)std::max<int64_t>((int64_t)s.sum_sq / kCellPixels - mean * mean, 0);
    const int64_t cbm = (int64_t)s.cb_sum / kCellPixels, crm = (int64_t)s.cr_sum / kCellPixels;
    c.chroma = (uint32_t)std::min<int64_t>(std::max<int64_t>((int64_t)s.cb_sq / kCellPixels - cbm * cbm, 0) +
                                               std::max<int64_t>((int64_t)s.cr_sq / kCellPixels - crm * crm, 0),
                                           65535);
    const int64_t dir[4] = {s.dir[0], s.dir[1], s.dir[2], s.dir[3]};
    const int64_t edge_total = dir[0] + dir[1] + dir[2] + dir[3];
    const int64_t denom = std::max<int64_t>(edge_total, 1);
    const int64_t dir_max = std::max(std::max(dir[0], dir[1]), std::max(dir[2], dir[3]));
    c.edge_q8 = (int)(((int64_t)s.edge * 256) / kCellPixels);
    c.flat_q8 = (int)(((int64_t)s.flat * 256) / kCellPixels);
    c.noise = s.weak > 0 ? (int)((int64_t)s.noise_sum / s.weak) : 0;
    c.has_edges = edge_total > 0;
    c.entropy_q8 = entropy4_q8(dir, edge_total);
    c.dominance_q8 = (int)((dir_max * 256) / denom);
    c.axis_q8 = (int)(((dir[0] + dir[1]) * 256) / denom);
    c.cls = classify(c);
  }
  const double count = (double)n;
  float share[7] = {0, 0, 0, 0, 0, 0, 0};
  for (const Cell& c : cells) share[c.cls] += 1.0f;
  for (float& s : share) s /= (float)count;
  const double* L = log2_1p_table();
  std::vector<uint32_t> var_k(n), edge_k(n), noise_k(n), chroma_k(n);
  double lv_sum = 0.0, edge_sum = 0.0, noise_sum = 0.0, chroma_sum = 0.0, flat_sum = 0.0;
  for (int k = 0; k < n; k++) {
    const Cell& c = cells[k];
    var_k[k] = c.variance;
    edge_k[k] = (uint32_t)c.edge_q8;
    noise_k[k] = (uint32_t)c.noise;
    chroma_k[k] = c.chroma;
    lv_sum += L[c.variance];
  }
  const double lv_mean = lv_sum / count;
  double ss = 0.0;
  int64_t hetero = 0;
  for (int k = 0; k < n; k++) {
    const double v = L[cells[k].variance];
    ss += (v - lv_mean) * (v - lv_mean);
  }
  for (int k = 0; k < n; k++) hetero += fabs(L[cells[k].variance] - lv_mean) > kHeteroBits ? 1 : 0;
  const double lv_std = sqrt(ss / count);
  for (int k = 0; k < n; k++) edge_sum += (double)cells[k].edge_q8 / 256.0;
  for (const Cell& c : cells) flat_sum += (double)c.flat_q8 / 256.0;
  for (int k = 0; k < n; k++) noise_sum += (double)cells[k].noise;
  for (int k = 0; k < n; k++) chroma_sum += L[cells[k].chroma];
  const double edge_mean = edge_sum / count, noise_mean = noise_sum / count, chroma_mean = chroma_sum / count;
  double ent = 0.0, dom = 0.0, axis = 0.0;
  int64_t edged = 0;
  for (const Cell& c : cells)
    if (c.has_edges) {
      ent += (double)c.entropy_q8 / 256.0;
      edged++;
    }
  for (const Cell& c : cells)
    if (c.has_edges) dom += (double)c.dominance_q8 / 256.0;
  for (const Cell& c : cells)
    if (c.has_edges) axis += (double)c.axis_q8 / 256.0;
  auto lv_of = [L](uint32_t v) { return L[v]; };
  auto edge_of = [](uint32_t v) { return (double)v / 256.0; };
  auto noise_of = [](uint32_t v) { return (double)v; };
  for (int i = 0; i < 7; i++) f->share[i] = share[i];
  f->log2var_mean = (float)lv_mean;
  f->log2var_std = (float)lv_std;
  f->log2var_q10 = (float)lerp_quantile_keys(var_k, 0.1, lv_of);
  f->log2var_q90 = (float)lerp_quantile_keys(var_k, 0.9, lv_of);
  f->edge_mean = (float)edge_mean;
  f->edge_q90 = (float)lerp_quantile_keys(edge_k, 0.9, edge_of);
  f->flat_mean = (float)(flat_sum / count);
  f->noise_mean = (float)noise_mean;
  f->noise_q90 = (float)lerp_quantile_keys(noise_k, 0.9, noise_of);
  f->chroma_log2_mean = (float)chroma_mean;
  f->chroma_log2_q90 = (float)lerp_quantile_keys(chroma_k, 0.9, lv_of);
  f->orient_entropy_mean = edged ? (float)(ent / (double)edged) : 0.0f;
  f->dir_dominance_mean = edged ? (float)(dom / (double)edged) : 0.0f;
  f->axis_aligned_mean = edged ? (float)(axis / (double)edged) : 0.0f;
  f->hetero_frac = (float)((double)hetero / count);
}

void transform_summary_from(const SummaryPartial* parts, int n, const uint32_t* hist, jpxlg_transform_summary* t) {
  constexpr double kEps = 1e-30, kHistLo = -46.0, kHistWidth = 0.125;
  uint64_t blocks = 0, ac_cells = 0, near3 = 0, near2 = 0;
  double y = 0, low = 0, high = 0, row = 0, col = 0, xb = 0, dcs = 0, dcq = 0;
  for (int i = 0; i < n; i++) {
    const SummaryPartial& p = parts[i];
    blocks += p.blocks;
    ac_cells += p.ac_cells;
    near3 += p.near_1e3;
    near2 += p.near_1e2;
    y += p.total_y;
    low += p.total_low;
    high += p.total_high;
    row += p.total_row;
    col += p.total_col;
    xb += p.total_xb;
    dcs += p.dc_sum;
    dcq += p.dc_sumsq;
  }
  auto quantile = [&](double q) {
    const double want_f = std::min(q * (double)blocks, (double)(blocks == 0 ? 0 : blocks - 1));
    const uint64_t want = (uint64_t)want_f;
    uint64_t seen = 0;
    for (int b = 0; b < kSummaryBins; b++) {
      seen += hist[b];
      if (seen > want) return kHistLo + ((double)b + 0.5) * kHistWidth;
    }
    return kHistLo + ((double)kSummaryBins - 0.5) * kHistWidth;
  };
  const double nb = std::max((double)blocks, 1.0);
  t->blocks = blocks;
  t->ln_ac_y_mean = log(y / nb + kEps);
  t->ln_ac_y_q50 = quantile(0.5);
  t->ln_ac_y_q90 = quantile(0.9);
  t->ln_ac_y_q99 = quantile(0.99);
  t->high_low_ratio = high / (low + kEps);
  t->directional_asymmetry = fabs(row - col) / (row + col + kEps);
  t->chroma_ac_ratio = xb / (y + kEps);
  t->near_zero_frac_1e3 = (double)near3 / std::max((double)ac_cells, 1.0);
  t->near_zero_frac_1e2 = (double)near2 / std::max((double)ac_cells, 1.0);
  const double mean = dcs / nb;
  t->dc_variance_y = std::max(dcq / nb - mean * mean, 0.0);
}

bool predict_crossing(const jpxlg_source_features& f, const jpxlg_transform_summary& t, double target,
                      double reserve, int which, jpxlg_crossing_prediction* out) {
  memset(out, 0, sizeof(*out));
  if (which == JPXLG_PREDICTOR_KNOTS) return predict_v2(f, t, target, out);
  if (which == JPXLG_PREDICTOR_CASES) return predict_cases(f, t, target, reserve, out);
  // predict (quality_prediction.rs:351-367): the case table when it knows
  // the frame (or the frame is only larger than its domain), else the knot
  // model.
  jpxlg_crossing_prediction p;
  if (predict_cases(f, t, target, reserve, &p) && (p.fallback == JPXLG_FALLBACK_NONE || p.large_frame_only_ood)) {
    *out = p;
    return true;
  }
  memset(out, 0, sizeof(*out));
  return predict_v2(f, t, target, out);
}

QualityStart quality_start(const jpxlg_source_features& f, const jpxlg_transform_summary& t, double target,
                           double reserve, uint32_t legacy_rung) {
  jpxlg_crossing_prediction p;
  if (predict_crossing(f, t, target, reserve, JPXLG_PREDICTOR_DISPATCH, &p)) {
    if (p.fallback == JPXLG_FALLBACK_NONE) return {p.candidate_rung, p.local_loss_exponent};
    if (p.ood_flags == 0) return {p.median_rung, p.local_loss_exponent};
    if (p.large_frame_only_ood) return {p.median_rung, NAN};  // flagged-median-start
  }
  return {legacy_rung, NAN};
}

}  // namespace jpxlg

extern "C" int32_t jpxlg_predict_crossing(const jpxlg_source_features* source, const jpxlg_transform_summary* transform,
                                          double target, double reserve, int32_t which,
                                          jpxlg_crossing_prediction* out) {
  if (!source || !transform || !out || which < 0 || which > 2) return JPXLG_E_INVALID;
  return jpxlg::predict_crossing(*source, *transform, target, reserve, which, out) ? 1 : 0;
}
