// Quality controller: the dense quantizer ladder, the first-guess predictor
// and the score navigator.
#include "controller.h"

#include <math.h>

#include <algorithm>
#include <cmath>

namespace jpxlg {

namespace {

// quantizer_ladder.rs:16-17 HF_MUL_RUNGS = 64, MAX_LADDER_MUL = 65.
constexpr uint32_t kMaxLadderMul = 65;

uint32_t segment_first_scale(uint32_t k) {  // quantizer_ladder.rs segment_first_scale
  uint64_t prev_top = (uint64_t)(k - 1) * kMaxGlobalScale;
  uint32_t floor_v = (uint32_t)(prev_top / k);
  return floor_v >= kMaxGlobalScale ? kMaxGlobalScale : floor_v + 1;
}

uint32_t segment_len(uint32_t k) { return kMaxGlobalScale - segment_first_scale(k) + 1; }

uint32_t upper_rungs() {
  uint32_t total = 0;
  for (uint32_t k = 2; k <= kMaxLadderMul; k++) total += segment_len(k);
  return total;
}

// quality.rs:160-242 navigator constants.
constexpr double kBracketRatio = 1.8;
constexpr double kPriorLossExponent = 0.9;
constexpr double kExpansionMargin = 1.25;
constexpr double kMaxExpansionJump = 16.0;
constexpr double kStructureRebuildRatio = 1.8;
constexpr double kLossEpsilon = 1e-3;
constexpr double kMetOvershootBand = 1.0;
constexpr double kMetBandReserveMultiple = 2.0;
constexpr double kMinAimMargin = 0.25;
constexpr double kCorrectedStopMoveGate = 0.5;
constexpr double kF64Epsilon = 2.220446049250313e-16;

double loss(double score) { return std::max(100.0 - score, kLossEpsilon); }
double neg_ln_loss(double score) { return -log(loss(score)); }
double ln_scale(uint32_t rung) { return log((double)Rung::effective_scale(rung)); }

uint32_t rung_clamped(uint64_t r) { return r > Rung::top() ? Rung::top() : (uint32_t)r; }  // Rung::new

uint32_t clamp_rung(uint32_t r, uint32_t lo, uint32_t hi) { return std::min(std::max(r, lo), hi); }

// navigation.rs:162-182 force_progress under Progress::NudgeWhenStuck.
uint32_t nudge_when_stuck(uint32_t aim, uint32_t from, bool finer) {
  if (aim != from) return aim;
  return finer ? rung_clamped((uint64_t)from + 1) : (from == 0 ? 0 : from - 1);
}

struct Point {
  uint32_t rung;
  double score;
};

// quality.rs:933-988 extrapolated_step_from: the crossing predictor's loss
// exponent (prior_beta, NaN when none) or else PRIOR_LOSS_EXPONENT stands in
// while no measured slope fits.
bool extrapolated_step_from(std::vector<Point> pts, double threshold, double prior_beta, bool finer,
                            uint32_t attempt, uint32_t* out) {
  if (pts.empty()) return false;
  std::stable_sort(pts.begin(), pts.end(), [](const Point& a, const Point& b) { return a.rung < b.rung; });
  Point from;
  bool has_other = false;
  Point other{};
  if (finer) {
    from = pts.back();
    if (pts.size() >= 2) { other = pts[pts.size() - 2]; has_other = true; }
  } else {
    from = pts.front();
    if (pts.size() >= 2) { other = pts[1]; has_other = true; }
  }
  if ((finer && from.rung == Rung::top()) || (!finer && from.rung == 0)) return false;
  double alpha = std::isnan(prior_beta) ? kPriorLossExponent : prior_beta;
  if (has_other) {
    double x0 = ln_scale(other.rung), y0 = neg_ln_loss(other.score);
    double x1 = ln_scale(from.rung), y1 = neg_ln_loss(from.score);
    double slope = (y1 - y0) / (x1 - x0);
    if (fabs(x1 - x0) > kF64Epsilon && slope > 0.0) alpha = std::min(std::max(slope, 0.2), 3.0);
  }
  double margin = pow(kExpansionMargin, (double)(attempt + 1));
  double ratio = pow(loss(from.score) / loss(threshold), 1.0 / alpha);
  ratio = finer ? ratio * margin : ratio / margin;
  ratio = std::min(std::max(ratio, 1.0 / kMaxExpansionJump), kMaxExpansionJump);
  if ((finer && ratio <= 1.0) || (!finer && ratio >= 1.0)) {
    *out = geometric_step(from.rung, finer);
    return true;
  }
  uint32_t rung = rung_for_scale(exp(ln_scale(from.rung)) * ratio);
  *out = rung == from.rung ? geometric_step(from.rung, finer) : rung;
  return true;
}

// quality.rs:1016-1019 structure_is_far.
bool structure_is_far(uint32_t anchor, uint32_t rung) {
  return fabs(ln_scale(rung) - ln_scale(anchor)) > log(kStructureRebuildRatio);
}

// quality_predictor.rs INITIAL_RUNG_TABLE: (target, luma bucket, flat
// bucket, global_scale); LUMA_BUCKET_EDGES, FLAT_BUCKET_EDGES,
// FALLBACK_LOG_FIT.  Calibration data of the reference, kept as printed.
struct RungEntry {
  float target;
  uint8_t luma_bucket, flat_bucket;
  uint32_t global_scale;
};
const RungEntry kInitialRungTable[] = {
    {30.0f, 0, 1, 6896},
    {30.0f, 0, 2, 471},
    {30.0f, 1, 1, 4783},
    {30.0f, 1, 2, 572},
    {30.0f, 2, 0, 1690},
    {30.0f, 3, 0, 1785},
    {30.0f, 3, 1, 3854},
    {30.0f, 4, 0, 3821},
    {50.0f, 0, 1, 15154},
    {50.0f, 0, 2, 801},
    {50.0f, 1, 1, 10711},
    {50.0f, 1, 2, 787},
    {50.0f, 2, 0, 4136},
    {50.0f, 3, 0, 4695},
    {50.0f, 3, 1, 8391},
    {50.0f, 4, 0, 10966},
    {70.0f, 0, 1, 35247},
    {70.0f, 0, 2, 1650},
    {70.0f, 1, 1, 24427},
    {70.0f, 1, 2, 1299},
    {70.0f, 2, 0, 14834},
    {70.0f, 3, 0, 14981},
    {70.0f, 3, 1, 22939},
    {70.0f, 4, 0, 30507},
    {80.0f, 0, 1, 65292},
    {80.0f, 0, 2, 3048},
    {80.0f, 1, 1, 55021},
    {80.0f, 1, 2, 2111},
    {80.0f, 2, 0, 32236},
    {80.0f, 3, 0, 30721},
    {80.0f, 3, 1, 44354},
    {80.0f, 4, 0, 49429},
    {85.0f, 0, 1, 73728},
    {85.0f, 0, 2, 5043},
    {85.0f, 1, 1, 73728},
    {85.0f, 1, 2, 3202},
    {85.0f, 2, 0, 53561},
    {85.0f, 3, 0, 50434},
    {85.0f, 3, 1, 68744},
    {85.0f, 4, 0, 66284},
    {90.0f, 0, 1, 73728},
    {90.0f, 0, 2, 9688},
    {90.0f, 1, 1, 73728},
    {90.0f, 1, 2, 9139},
    {90.0f, 2, 0, 73728},
    {90.0f, 3, 0, 73728},
    {90.0f, 3, 1, 73728},
    {90.0f, 4, 0, 73728},
    {95.0f, 0, 1, 73728},
    {95.0f, 0, 2, 43968},
    {95.0f, 1, 1, 73728},
    {95.0f, 1, 2, 73728},
    {95.0f, 2, 0, 73728},
    {95.0f, 3, 0, 73728},
    {95.0f, 3, 1, 73728},
    {95.0f, 4, 0, 73728},
};
const float kLumaBucketEdges[4] = {1.2599952e-07f, 5.6737383e-05f, 9.813392e-05f, 0.00017977731f};
const float kFlatBucketEdges[2] = {0.2f, 0.6f};
const double kFallbackLogFit[4] = {15.615602561070592, -1.6033739861412692, 0.009891719405523072,
                                   -2.5749219514805652};
// quality_features.rs:19 FLAT_VARIANCE.
constexpr float kFlatVariance = 2e-5f;

uint8_t bucket(float value, const float* edges, int n) {  // quality.rs:1021-1027
  int i = 0;
  while (i < n && value >= edges[i]) i++;
  return (uint8_t)i;
}

}  // namespace

namespace Rung {

uint32_t ladder_len() { return kMaxGlobalScale + upper_rungs(); }
uint32_t top() { return ladder_len() - 1; }
uint32_t for_global_scale(uint32_t scale) { return scale == 0 ? 0 : std::min(scale - 1, top()); }

void fields(uint32_t rung, uint32_t* gs, uint32_t* mul) {
  rung = std::min(rung, top());
  if (rung < kMaxGlobalScale) { *gs = rung + 1; *mul = 1; return; }
  uint32_t offset = rung - kMaxGlobalScale, k = 2;
  while (k < kMaxLadderMul && offset >= segment_len(k)) { offset -= segment_len(k); k++; }
  *gs = std::min(segment_first_scale(k) + offset, kMaxGlobalScale);
  *mul = k;
}

uint64_t effective_scale(uint32_t rung) {
  uint32_t gs, mul;
  fields(rung, &gs, &mul);
  return (uint64_t)gs * mul;
}

uint32_t for_effective_scale(uint64_t scale) {  // quantizer_ladder.rs rung_for_effective_scale
  const uint64_t max = kMaxGlobalScale;
  if (scale <= max) return (uint32_t)std::min<uint64_t>(scale == 0 ? 0 : scale - 1, top());
  if (scale >= max * kMaxLadderMul) return top();
  uint32_t k = (uint32_t)((scale + max - 1) / max);
  k = std::min(std::max(k, 2u), kMaxLadderMul);
  uint32_t gs = (uint32_t)(scale / k);
  uint32_t index = kMaxGlobalScale;
  for (uint32_t j = 2; j < k; j++) index += segment_len(j);
  if (gs < segment_first_scale(k)) return std::min(index - 1, top());
  return std::min(index + (gs - segment_first_scale(k)), top());
}

uint32_t coupled_quant_lf(uint32_t base, uint32_t k) {  // quantizer_ladder.rs coupled_quant_lf
  uint32_t best = base;
  for (uint32_t j = 2; j <= k; j++) {
    uint64_t cand = (uint64_t)base * j;
    if (cand >= 1 && cand <= 65536 && cand <= 16) best = (uint32_t)cand;
    else break;
  }
  return best;
}

}  // namespace Rung

uint32_t rung_for_scale(double scale) {  // navigation.rs:65-71
  if (!isfinite(scale)) return 0;
  double r = std::min(std::max(round(scale), 1.0), 18446744073709551615.0);
  return Rung::for_effective_scale(r >= 18446744073709551615.0 ? UINT64_MAX : (uint64_t)r);
}

jpxlg_quantizer QualityPolicy::quantizer_at(uint32_t rung) const {
  jpxlg_quantizer q{};
  uint32_t gs, mul;
  Rung::fields(rung, &gs, &mul);
  q.global_scale = gs;
  q.hf_mul = mul;
  q.quant_lf = Rung::coupled_quant_lf(quant_lf, mul);
  q.x_qm_scale = x_qm;
  q.b_qm_scale = b_qm;
  q.truncate = truncate;
  q.lambda_scale = lambda_scale;
  q.zero_token_bits = zero_token_bits;
  return q;
}

QualityPolicy quality_policy(int32_t effort) {
  QualityPolicy p{};
  p.quant_lf = 4;             // request.rs:900 (Phase Q1)
  p.lambda_scale = 4.0f;      // request.rs:915
  p.zero_token_bits = 1.0f;   // lib.rs:3200 ZERO_TOKEN_BITS
  p.epf_iters = 1;            // request.rs:905
  p.sharpness = 7;            // request.rs:906 Uniform7
  p.structural_builds = 2;    // quality.rs:397-414
  if (effort == JPXLG_EFFORT_FAST) {
    // Fast navigation: neutral chroma, nearest quantizer, fixed DCT8x8 cover
    // (request.rs:965-967, lib.rs:628-642) and no CfL estimate
    // (planning_tiers, quality.rs:1739).
    p.x_qm = 2; p.b_qm = 2;
    p.truncate = 0;
    p.cover_mode = JPXLG_COVER_FIXED_DCT8;
    p.enable_cfl = 0;
    p.pixel_probes = 3;
    p.exact_prices = 2;
    p.reserve = 0.06;
  } else {
    p.x_qm = 3; p.b_qm = 3;   // request.rs:968-971
    p.truncate = 1;           // TrailingTruncation
    p.cover_mode = JPXLG_COVER_HIERARCHICAL;
    p.enable_cfl = 1;
    p.pixel_probes = 5;
    p.exact_prices = 3;
    p.reserve = 0.03;
  }
  return p;
}

uint32_t geometric_step(uint32_t from, int finer) {  // quality.rs:914-926
  double scale = exp(ln_scale(from));
  double next = finer ? scale * kBracketRatio : scale / kBracketRatio;
  return nudge_when_stuck(rung_for_scale(next), from, finer != 0);
}

double aim_score_for(double threshold, double reserve) {  // quality.rs:991-994
  double relative = 100.0 - loss(threshold) * (1.0 - reserve);
  return std::max(relative, threshold + kMinAimMargin);
}

double met_band(double threshold, double reserve) {  // quality.rs:219-228 (reserve-coupled-band)
  return std::max(kMetBandReserveMultiple * reserve * loss(threshold), kMetOvershootBand);
}

uint32_t log_loss_crossing(uint32_t lo_rung, double lo_score, uint32_t hi_rung, double hi_score,
                           double aim_score, int* ok) {  // quality.rs:888-911
  *ok = 0;
  if (hi_rung <= lo_rung + 1) return 0;
  double x_lo = ln_scale(lo_rung), x_hi = ln_scale(hi_rung);
  double y_lo = neg_ln_loss(lo_score), y_hi = neg_ln_loss(hi_score);
  if (y_hi <= y_lo || x_hi <= x_lo) return 0;
  double slope = (y_hi - y_lo) / (x_hi - x_lo);
  double x = x_lo + (neg_ln_loss(aim_score) - y_lo) / slope;
  uint32_t rung = rung_for_scale(exp(x));
  *ok = 1;
  return clamp_rung(rung, lo_rung + 1, hi_rung - 1);
}

double predicted_effective_scale(float luma_q50, float flat_fraction, double target) {  // quality.rs:1030-1079
  const uint8_t lb = bucket(luma_q50, kLumaBucketEdges, 4);
  const uint8_t fb = bucket(flat_fraction, kFlatBucketEdges, 2);
  const double target_loss = log(loss(target));
  std::vector<std::pair<double, double>> entries;
  for (const RungEntry& e : kInitialRungTable)
    if (e.luma_bucket == lb && e.flat_bucket == fb)
      entries.push_back({log(loss((double)e.target)), log((double)std::max(e.global_scale, 1u))});
  std::stable_sort(entries.begin(), entries.end(),
                   [](const std::pair<double, double>& a, const std::pair<double, double>& b) { return a.first < b.first; });
  if (entries.size() >= 2) {
    std::pair<double, double> a, b;
    bool found = false;
    for (size_t i = 0; i + 1 < entries.size(); i++)
      if (entries[i].first <= target_loss && target_loss <= entries[i + 1].first) {
        a = entries[i];
        b = entries[i + 1];
        found = true;
        break;
      }
    if (!found) {
      if (target_loss < entries.front().first) { a = entries[0]; b = entries[1]; }
      else { a = entries[entries.size() - 2]; b = entries.back(); }
    }
    if (fabs(b.first - a.first) > kF64Epsilon) {
      const double t = (target_loss - a.first) / (b.first - a.first);
      return exp(a.second + t * (b.second - a.second));
    }
  }
  if (!entries.empty()) return exp(entries[0].second);
  return exp(kFallbackLogFit[0] + kFallbackLogFit[1] * target_loss +
             kFallbackLogFit[2] * log((double)luma_q50 + 1e-6) + kFallbackLogFit[3] * (double)flat_fraction);
}

void source_features(std::vector<float> luma, float* luma_q50, float* flat_fraction) {  // quality_features.rs:154-213
  size_t flat = 0;
  for (float v : luma) flat += v < kFlatVariance ? 1 : 0;
  *flat_fraction = luma.empty() ? 0.0f : (float)flat / (float)luma.size();
  if (luma.empty()) { *luma_q50 = 0.0f; return; }
  const size_t last = luma.size() - 1;
  size_t index = (size_t)floor((double)last * (double)0.5f);
  index = std::min(index, last);
  std::nth_element(luma.begin(), luma.begin() + index, luma.end(),
                   [](float a, float b) { return a < b; });
  *luma_q50 = luma[index];
}

int navigate_with_budget(double target, const NavBudget& budget, const ProbeFn& probe, ControllerOutcome* out,
                         uint32_t predicted_rung, double prior_beta, const PriceFn* price) {
  const double threshold = target;  // DEFAULT_SCORE_GUARD = 0
  const double band = met_band(threshold, budget.reserve);
  std::vector<ProbeResult>& probes = out->probes;
  bool anchored = false;
  auto already = [&](uint32_t r) {
    for (auto& p : probes) if (p.rung == r) return true;
    return false;
  };
  // Navigator::probe (quality.rs:1153-1243); returns the record's index.
  auto run = [&](uint32_t r, bool fresh, int* index) -> int {
    const bool build = fresh || !anchored;
    double s = 0, ms = 0;
    int rc = probe(r, build, &s, &ms);
    if (rc) return rc;
    if (build) {
      out->structural_builds++;
      if (!anchored) { anchored = true; out->anchor_rung = r; }
    }
    out->pixel_probes++;
    probes.push_back({r, s, ms, s >= threshold, build});
    if (index) *index = (int)probes.size() - 1;
    return 0;
  };
  auto points = [&]() {
    std::vector<Point> v;
    for (auto& p : probes) v.push_back({p.rung, p.score});
    return v;
  };
  // quality.rs:1343-1357: the finest infeasible and the coarsest feasible.
  auto bracket = [&](Point* lo, Point* hi) {
    bool hl = false, hh = false;
    for (auto& p : probes) {
      if (!p.feasible && (!hl || p.rung >= lo->rung)) { *lo = {p.rung, p.score}; hl = true; }
      if (p.feasible && (!hh || p.rung < hi->rung)) { *hi = {p.rung, p.score}; hh = true; }
    }
    return hl && hh && lo->rung < hi->rung;
  };
  auto budget_left = [&]() { return out->pixel_probes < budget.pixel_probes; };
  auto in_band = [&](int i) { return probes[i].feasible && probes[i].score - threshold <= band; };
  int rc = 0;

  // The predicted first probe builds the anchor structure.
  if ((rc = run(predicted_rung, true, nullptr))) return rc;
  // trusted-first-probe (quality.rs:1832-1845).
  const bool one_shot_hit = in_band(0);
  // corrected-stop (quality.rs:1536-1600).
  bool corrected_hit = false;
  if (!one_shot_hit) {
    const Point first{probes[0].rung, probes[0].score};
    const double aim = aim_score_for(threshold, budget.reserve);
    auto corrected_aim = [&](Point from, double beta) {
      double ratio = pow(loss(from.score) / loss(aim), 1.0 / beta);
      ratio = std::min(std::max(ratio, 1.0 / kMaxExpansionJump), kMaxExpansionJump);
      const uint32_t r = rung_for_scale(exp(ln_scale(from.rung)) * ratio);
      return nudge_when_stuck(r, from.rung, from.score < aim);
    };
    const double implied = fabs(log(loss(first.score) / loss(aim))) / kPriorLossExponent;
    if (!(implied > kCorrectedStopMoveGate) && budget_left()) {
      const uint32_t second_rung = corrected_aim(first, kPriorLossExponent);
      if (!already(second_rung)) {
        int si = 0;
        if ((rc = run(second_rung, false, &si))) return rc;
        if (in_band(si)) {
          corrected_hit = true;
        } else if (budget_left()) {
          const Point second{probes[si].rung, probes[si].score};
          const double x1 = ln_scale(first.rung), x2 = ln_scale(second.rung);
          if (fabs(x2 - x1) > kF64Epsilon) {
            const double slope = (neg_ln_loss(second.score) - neg_ln_loss(first.score)) / (x2 - x1);
            if (!isnan(slope) && slope > 0.0) {
              const Point near = fabs(second.score - aim) < fabs(first.score - aim) ? second : first;
              const uint32_t third_rung = corrected_aim(near, std::min(std::max(slope, 0.2), 3.0));
              if (!already(third_rung)) {
                int ti = 0;
                if ((rc = run(third_rung, false, &ti))) return rc;
                corrected_hit = in_band(ti);
              }
            }
          }
        }
      }
    }
  }
  // expand_until_bracketed (quality.rs:1381-1398).
  Point lo{}, hi{};
  if (!one_shot_hit && !corrected_hit) {
    uint32_t attempt = 0;
    while (!bracket(&lo, &hi) && budget_left()) {
      const bool finer = !probes.back().feasible;
      uint32_t next;
      if (!extrapolated_step_from(points(), threshold, prior_beta, finer, attempt, &next)) break;
      if (already(next)) break;
      if ((rc = run(next, false, nullptr))) return rc;
      attempt++;
    }
  }
  // tighten (quality.rs:1603-1611): crossing_is_tight / tighten_aim.
  while (budget_left()) {
    if (!bracket(&lo, &hi)) break;
    if (hi.rung <= lo.rung + 1 || hi.score - threshold <= band) break;
    const double aim = aim_score_for(threshold, budget.reserve);
    if (aim >= hi.score) break;
    int ok = 0;
    const uint32_t r = log_loss_crossing(lo.rung, lo.score, hi.rung, hi.score, aim, &ok);
    if (!ok || already(r)) break;
    if ((rc = run(r, false, nullptr))) return rc;
  }
  // rescue_probe (quality.rs:1404-1417): one probe beyond the budget when
  // nothing met.
  {
    bool any_feasible = false;
    for (auto& p : probes) any_feasible |= p.feasible;
    uint32_t next;
    if (!any_feasible && extrapolated_step_from(points(), threshold, prior_beta, true, 3, &next) && !already(next))
      if ((rc = run(next, false, nullptr))) return rc;
  }

  // Finalists (quality.rs:1870-1990): the coarsest feasible probes.
  struct Finalist {
    uint32_t rung;
    double score;
    bool feasible;
    bool fresh;  // rebuilt at its own rung
  };
  std::vector<Finalist> finalists;
  const int max_finalists = std::min(std::max(budget.exact_prices, 1), 2);
  std::vector<int> ordered;
  for (int i = 0; i < (int)probes.size(); i++)
    if (probes[i].feasible) ordered.push_back(i);
  std::stable_sort(ordered.begin(), ordered.end(), [&](int a, int b) { return probes[a].rung < probes[b].rung; });
  if ((int)ordered.size() > max_finalists) ordered.resize(max_finalists);
  bool under_target = false, saturated = false, rescued = false;
  if (ordered.empty()) {
    under_target = true;
    int idx = 0;
    for (int i = 0; i < (int)probes.size(); i++)
      if (probes[i].rung >= probes[idx].rung) idx = i;  // max_by_key keeps the last maximum
    saturated = probes[idx].rung == Rung::top();
    ordered.push_back(idx);
  }
  for (int index : ordered) {
    if (out->priced >= budget.exact_prices) break;
    const ProbeResult p = probes[index];
    const bool reused = !(p.fresh && p.rung == out->anchor_rung);
    const bool rebuild = reused && structure_is_far(out->anchor_rung, p.rung) &&
                         out->structural_builds < budget.structural_builds && budget_left() && !under_target;
    if (rebuild) {
      int fi = 0;
      if ((rc = run(p.rung, true, &fi))) return rc;
      if (probes[fi].feasible) {
        out->priced++;
        finalists.push_back({p.rung, probes[fi].score, true, true});
        rescued = true;
        continue;
      }
    }
    out->priced++;
    finalists.push_back({p.rung, p.score, p.feasible, false});
  }
  // Selection (quality.rs:1996-2010): the smallest exact feasible stream
  // (the first of equal sizes) when there is a choice to price; else the
  // feasible finalist with the coarsest rung; with none feasible, the finest.
  const Finalist* pick = nullptr;
  int nfeasible = 0;
  for (auto& f : finalists) nfeasible += f.feasible;
  if (price && nfeasible >= 2) {
    uint64_t best = UINT64_MAX;
    for (auto& f : finalists) {
      if (!f.feasible) continue;
      uint64_t bytes = 0;
      if ((rc = (*price)(f.rung, f.fresh, &bytes))) return rc;
      out->exact_prices++;
      if (bytes < best) {
        best = bytes;
        pick = &f;
      }
    }
    out->bytes = best;
  }
  if (!pick)
    for (auto& f : finalists)
      if (f.feasible && (!pick || f.rung < pick->rung)) pick = &f;
  if (!pick)
    for (auto& f : finalists)
      if (!pick || f.rung >= pick->rung) pick = &f;
  if (!pick) return JPXLG_E_INVALID;
  out->rung = pick->rung;
  out->score = pick->score;
  out->met = pick->feasible ? 1 : 0;
  out->chosen_fresh = pick->fresh;
  // QualityStatus (quality.rs:2710-2724).
  bool adjacent_infeasible = false;
  {
    const ProbeResult* cf = nullptr;
    for (auto& q : probes)
      if (q.feasible && (!cf || q.rung < cf->rung)) cf = &q;
    if (cf)
      for (auto& q : probes)
        if (!q.feasible && (uint64_t)q.rung + 1 == cf->rung) adjacent_infeasible = true;
  }
  if (saturated) out->status = kStatusSaturatedTop;
  else if (under_target) out->status = kStatusUnderTargetWorkCap;
  else if (rescued && pick->fresh) out->status = kStatusRescuedFreshStructure;
  else if (pick->rung == 0) out->status = kStatusSaturatedFloor;
  else if (pick->score - threshold <= band) out->status = kStatusMet;
  else if (adjacent_infeasible) out->status = kStatusMetAdjacentRungs;
  else out->status = kStatusMetWorkCap;
  return 0;
}

int navigate_controller(double target, int32_t effort, int max_probes, const ProbeFn& probe,
                        ControllerOutcome* out, uint32_t predicted_rung, double prior_beta, const PriceFn* price) {
  const QualityPolicy pol = quality_policy(effort);
  NavBudget b{max_probes > 0 ? max_probes : pol.pixel_probes, pol.exact_prices, pol.structural_builds, pol.reserve};
  return navigate_with_budget(target, b, probe, out, predicted_rung, prior_beta, price);
}

int bisect_controller(double target, int max_probes, const ProbeFn& probe, ControllerOutcome* out) {
  const int budget = max_probes > 0 ? max_probes : 8;
  // Bisection in log effective scale between the floor and the top rung.
  uint32_t lo = 0, hi = Rung::top();  // lo: coarse side, hi: fine side
  for (int i = 0; i < budget && lo + 1 < hi; i++) {
    double xm = 0.5 * (ln_scale(lo) + ln_scale(hi));
    uint32_t mid = clamp_rung(rung_for_scale(exp(xm)), lo + 1, hi - 1);
    double s = 0, ms = 0;
    const bool first = out->probes.empty();
    int rc = probe(mid, first, &s, &ms);
    if (rc) return rc;
    if (first) { out->structural_builds++; out->anchor_rung = mid; }
    out->pixel_probes++;
    bool ok = s >= target;
    out->probes.push_back({mid, s, ms, ok, first});
    if (ok) hi = mid;
    else lo = mid;
  }
  const ProbeResult* best = nullptr;
  for (auto& p : out->probes)
    if (p.feasible && (!best || p.rung < best->rung)) best = &p;
  if (!best)
    for (auto& p : out->probes)
      if (!best || p.rung > best->rung) best = &p;
  if (!best) return JPXLG_E_INVALID;
  out->rung = best->rung;
  out->score = best->score;
  out->met = best->feasible ? 1 : 0;
  out->status = best->feasible ? kStatusMet : kStatusUnderTargetWorkCap;
  return 0;
}


// ----------------------------------------------------------------------------
// The exact rate loop's ladder search (jpxl-encode-policy/src/rate.rs:387-921):
// a cold bracket doubling the effective scale, aimed bisection alternating
// with midpoints, and a fill of the rungs just above the best feasible one.
// ----------------------------------------------------------------------------
namespace {

// RateTolerance::bytes_for (request.rs:205-240): max(bytes, floor(fraction x target)).
uint64_t rate_slack(uint64_t target, uint64_t bytes, double fraction) {
  const double scaled = std::floor(std::max(fraction, 0.0) * (double)target);
  return std::max<uint64_t>(bytes, std::isfinite(scaled) ? (uint64_t)scaled : 0);
}

double ln_scale_of(uint32_t rung) { return std::log((double)Rung::effective_scale(rung)); }

uint32_t rung_for_scale_f(double scale) {  // navigation.rs:65-70
  if (!std::isfinite(scale)) return 0;
  return Rung::for_effective_scale((uint64_t)std::min(std::max(std::round(scale), 1.0), 1.8e19));
}

}  // namespace

uint32_t rate_cold_step(uint32_t current, bool up) {  // rate.rs:490-502
  const uint64_t scale = Rung::effective_scale(current);
  const uint64_t aimed = std::max<uint64_t>(up ? scale * 2 : scale / 2, 1);
  const uint32_t aim = Rung::for_effective_scale(aimed);
  // force_progress(.., AtLeastOneRung) (navigation.rs:162-170)
  if (up) return std::max(aim, std::min(current + 1, Rung::top()));
  return std::min(aim, current == 0 ? 0u : current - 1);
}

bool rate_interpolated_rung(uint32_t lo_rung, uint64_t lo_bytes, uint32_t hi_rung, uint64_t hi_bytes,
                            uint64_t target, uint32_t* out) {  // rate.rs:522-557
  if (hi_rung <= lo_rung + 1 || lo_bytes == 0 || target == 0) return false;
  if (hi_bytes <= lo_bytes) return false;
  const double lo_x = ln_scale_of(lo_rung), hi_x = ln_scale_of(hi_rung);
  const double lo_y = std::log((double)lo_bytes), hi_y = std::log((double)hi_bytes), t_y = std::log((double)target);
  const double span = hi_y - lo_y;  // bounded_fraction_crossing (navigation.rs:131-138)
  if (!std::isfinite(span) || span <= 0.0) return false;
  const double fraction = std::min(std::max((t_y - lo_y) / span, 0.0), 1.0);
  const double guess_x = lo_x + fraction * (hi_x - lo_x);
  if (!std::isfinite(guess_x)) return false;
  const double guess_scale = std::exp(guess_x);
  if (!std::isfinite(guess_scale) || guess_scale < 0.0) return false;
  const uint32_t guess = rung_for_scale_f(guess_scale);
  const uint32_t low = lo_rung + 1;
  if (hi_rung == 0) return false;
  const uint32_t high = hi_rung - 1;
  if (low > high) return false;
  *out = std::min(std::max(guess, low), high);
  return true;
}

int rate_search_ladder(uint32_t start, uint64_t target, uint64_t tol_bytes, double tol_fraction, uint32_t max_prices,
                       uint32_t fill_probes, const RatePriceFn& price, RateOutcome* out) {
  out->trace.clear();
  out->saturated = false;
  out->floor_bytes = 0;
  max_prices = std::max(max_prices, 1u);
  const uint32_t top = Rung::top();
  start = std::min(start, top);
  const uint64_t slack = rate_slack(target, tol_bytes, tol_fraction);
  bool have_best = false;
  uint32_t best_rung = 0;
  uint64_t best_bytes = 0;
  auto affordable = [&]() { return out->trace.size() < max_prices; };
  // Search::eval (rate.rs:660-680): a rung priced once; the largest feasible
  // size is the best (the finer rung among equal sizes).
  auto eval = [&](uint32_t rung, int phase, uint64_t* bytes) -> int {
    for (auto& s : out->trace)
      if (s.rung == rung) {
        *bytes = s.bytes;
        return 0;
      }
    uint64_t b = 0;
    const int rc = price(rung, &b);
    if (rc) return rc;
    const bool feasible = b <= target;
    out->trace.push_back({rung, b, feasible, phase});
    if (feasible && (!have_best || b > best_bytes || (b == best_bytes && rung > best_rung))) {
      have_best = true;
      best_rung = rung;
      best_bytes = b;
    }
    *bytes = b;
    return 0;
  };
  int rc;
  bool have_lo = false, have_hi = false, floor_hit = false;
  uint32_t lo = 0, hi = 0;
  uint64_t lo_b = 0, hi_b = 0, first = 0;
  if ((rc = eval(start, kRateBracket, &first))) return rc;
  if (first <= target) {
    have_lo = true, lo = start, lo_b = first;
    uint32_t current = start;
    while (current < top && affordable()) {
      const uint32_t next = rate_cold_step(current, true);
      if (next <= current) break;
      uint64_t b;
      if ((rc = eval(next, kRateBracket, &b))) return rc;
      if (b > target) {
        have_hi = true, hi = next, hi_b = b;
        break;
      }
      lo = next, lo_b = b;
      current = next;
    }
  } else {
    have_hi = true, hi = start, hi_b = first;
    uint32_t current = start;
    for (;;) {
      if (current == 0) {
        floor_hit = true;
        out->floor_bytes = hi_b;
        break;
      }
      if (!affordable()) break;
      const uint32_t next = rate_cold_step(current, false);
      uint64_t b;
      if ((rc = eval(next, kRateBracket, &b))) return rc;
      if (b <= target) {
        have_lo = true, lo = next, lo_b = b;
        break;
      }
      hi = next, hi_b = b;
      current = next;
    }
  }
  if (!have_lo) return floor_hit ? kRateUnreachable : kRateBudgetExhausted;
  // Aimed bisection (rate.rs:760-800): an aimed probe may raise lo but never
  // becomes hi (a spike above the crossing must not fence off finer rungs).
  bool take_aimed = true, aiming_pays = true;
  while (have_hi) {
    if (hi <= lo + 1 || !affordable()) break;
    const uint64_t bb = have_best ? best_bytes : 0;
    if (hi_b - std::min(hi_b, bb) <= slack) break;
    if (target - std::min(target, bb) <= slack) break;
    const uint32_t mid = lo + (hi - lo) / 2;
    uint32_t probe = mid;
    bool aimed = false;
    if (take_aimed && aiming_pays) {
      uint32_t guess;
      if (rate_interpolated_rung(lo, lo_b, hi, hi_b, target, &guess) && guess != mid) {
        probe = guess;
        aimed = true;
      }
    }
    take_aimed = !take_aimed;
    uint64_t b;
    if ((rc = eval(probe, kRateBisect, &b))) return rc;
    if (b <= target) {
      lo = probe, lo_b = b;
    } else if (aimed) {
      aiming_pays = false;
    } else {
      hi = probe, hi_b = b;
    }
  }
  // Fill (rate.rs:802-820): the rungs just above the best, while undershoot remains.
  if (have_best && target - best_bytes > slack) {
    const uint32_t base = best_rung;
    for (uint32_t off = 1; off <= fill_probes; off++) {
      if (!affordable()) break;
      const uint32_t cand = std::min(base + off, top);
      if (cand <= base) break;
      bool seen = false;
      for (auto& s : out->trace) seen |= s.rung == cand;
      if (seen) continue;
      uint64_t b;
      if ((rc = eval(cand, kRateFill, &b))) return rc;
    }
  }
  if (!have_best) return kRateBudgetExhausted;
  out->rung = best_rung;
  out->bytes = best_bytes;
  for (auto& s : out->trace) out->saturated |= s.rung == top && s.feasible;
  return 0;
}

}  // namespace jpxlg
