// Quality controller: the dense quantizer ladder and the score navigator.
#include "controller.h"

#include <math.h>

#include <algorithm>

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

// quality.rs:160-233 navigator constants.
constexpr double kBracketRatio = 1.8;
constexpr double kPriorLossExponent = 0.9;
constexpr double kExpansionMargin = 1.25;
constexpr double kMaxExpansionJump = 16.0;
constexpr double kLossEpsilon = 1e-3;
constexpr double kMetOvershootBand = 1.0;
constexpr double kMinAimMargin = 0.25;

double loss(double score) { return std::max(100.0 - score, kLossEpsilon); }
double neg_ln_loss(double score) { return -log(loss(score)); }
double ln_scale(uint32_t rung) { return log((double)Rung::effective_scale(rung)); }

uint32_t rung_for_scale(double scale) {  // navigation.rs:65-71
  if (!isfinite(scale)) return 0;
  double r = std::min(std::max(round(scale), 1.0), 18446744073709551615.0);
  return Rung::for_effective_scale((uint64_t)r);
}

uint32_t clamp_rung(uint32_t r, uint32_t lo, uint32_t hi) { return std::min(std::max(r, lo), hi); }

struct Point {
  uint32_t rung;
  double score;
};

// quality.rs:933-988 extrapolated_step_from.
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
  double alpha = prior_beta;
  if (has_other) {
    double x0 = ln_scale(other.rung), y0 = neg_ln_loss(other.score);
    double x1 = ln_scale(from.rung), y1 = neg_ln_loss(from.score);
    double slope = (y1 - y0) / (x1 - x0);
    if (fabs(x1 - x0) > 2.220446049250313e-16 && slope > 0.0) alpha = std::min(std::max(slope, 0.2), 3.0);
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
  if (effort == JPXLG_EFFORT_FAST) {
    // Fast navigation: neutral chroma, nearest quantizer, fixed DCT8x8 cover
    // (request.rs:965-967, lib.rs:640-650).
    p.x_qm = 2; p.b_qm = 2;
    p.truncate = 0;
    p.cover_mode = JPXLG_COVER_FIXED_DCT8;
    p.pixel_probes = 3;
    p.reserve = 0.06;
  } else {
    p.x_qm = 3; p.b_qm = 3;   // request.rs:968-971
    p.truncate = 1;           // TrailingTruncation
    p.cover_mode = JPXLG_COVER_HIERARCHICAL;
    p.pixel_probes = 5;
    p.reserve = 0.03;
  }
  return p;
}

uint32_t geometric_step(uint32_t from, int finer) {  // quality.rs:914-926
  double scale = exp(ln_scale(from));
  double next = finer ? scale * kBracketRatio : scale / kBracketRatio;
  uint32_t aim = rung_for_scale(next);
  if (aim != from) return aim;
  return finer ? std::min(from + 1, Rung::top()) : (from == 0 ? 0 : from - 1);
}

double aim_score_for(double threshold, double reserve) {  // quality.rs:991-994
  double relative = 100.0 - loss(threshold) * (1.0 - reserve);
  return std::max(relative, threshold + kMinAimMargin);
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

int navigate_controller(double target, int32_t effort, int max_probes, const ProbeFn& probe,
                        ControllerOutcome* out, uint32_t predicted_rung) {
  QualityPolicy pol = quality_policy(effort);
  const int budget = max_probes > 0 ? max_probes : pol.pixel_probes;
  const double threshold = target;  // DEFAULT_SCORE_GUARD = 0
  std::vector<ProbeResult>& probes = out->probes;
  auto already = [&](uint32_t r) {
    for (auto& p : probes) if (p.rung == r) return true;
    return false;
  };
  auto run = [&](uint32_t r) -> int {
    double s = 0, ms = 0;
    int rc = probe(r, &s, &ms);
    if (rc) return rc;
    probes.push_back({r, s, ms, s >= threshold});
    return 0;
  };
  auto points = [&]() {
    std::vector<Point> v;
    for (auto& p : probes) v.push_back({p.rung, p.score});
    return v;
  };
  auto bracket = [&](Point* lo, Point* hi) {  // quality.rs:1343-1357
    bool hl = false, hh = false;
    for (auto& p : probes) {
      if (!p.feasible && (!hl || p.rung > lo->rung)) { *lo = {p.rung, p.score}; hl = true; }
      if (p.feasible && (!hh || p.rung < hi->rung)) { *hi = {p.rung, p.score}; hh = true; }
    }
    return hl && hh && lo->rung < hi->rung;
  };
  auto budget_left = [&]() { return (int)probes.size() < budget; };
  // The predicted start (the source-feature predictor is not restated; the
  // caller's prediction or the ladder rung of the default global scale).
  uint32_t start = predicted_rung ? predicted_rung : Rung::for_global_scale(2048);
  int rc = run(start);
  if (rc) return rc;
  // expand_until_bracketed (quality.rs:1381-1398)
  uint32_t attempt = 0;
  Point lo{}, hi{};
  while (!bracket(&lo, &hi) && budget_left()) {
    bool finer = !probes.back().feasible;
    uint32_t next;
    if (!extrapolated_step_from(points(), threshold, kPriorLossExponent, finer, attempt, &next)) break;
    if (already(next)) break;
    if ((rc = run(next))) return rc;
    attempt++;
  }
  // tighten (quality.rs:1603-1611)
  const double band = kMetOvershootBand;
  while (budget_left()) {
    if (!bracket(&lo, &hi)) break;
    if (hi.rung <= lo.rung + 1 || hi.score - threshold <= band) break;
    double aim = aim_score_for(threshold, pol.reserve);
    if (aim >= hi.score) break;
    int ok = 0;
    uint32_t r = log_loss_crossing(lo.rung, lo.score, hi.rung, hi.score, aim, &ok);
    if (!ok || already(r)) break;
    if ((rc = run(r))) return rc;
  }
  // rescue_probe (quality.rs:1404-1417): one extra finer probe when nothing met.
  bool any_feasible = false;
  for (auto& p : probes) any_feasible |= p.feasible;
  if (!any_feasible) {
    uint32_t next;
    if (extrapolated_step_from(points(), threshold, kPriorLossExponent, true, 3, &next) && !already(next))
      if ((rc = run(next))) return rc;
  }
  // Finalist: the coarsest feasible probe (the smallest stream); with none,
  // the finest probe.
  const ProbeResult* best = nullptr;
  for (auto& p : probes)
    if (p.feasible && (!best || p.rung < best->rung)) best = &p;
  if (!best)
    for (auto& p : probes)
      if (!best || p.rung > best->rung) best = &p;
  out->rung = best->rung;
  out->score = best->score;
  out->met = best->feasible ? 1 : 0;
  return 0;
}

int bisect_controller(double target, int max_probes, const ProbeFn& probe, ControllerOutcome* out) {
  const int budget = max_probes > 0 ? max_probes : 8;
  // Bisection in log effective scale between the floor and the top rung.
  uint32_t lo = 0, hi = Rung::top();  // lo: coarse side, hi: fine side
  bool have_feasible = false;
  for (int i = 0; i < budget && lo + 1 < hi; i++) {
    double xm = 0.5 * (ln_scale(lo) + ln_scale(hi));
    uint32_t mid = clamp_rung(rung_for_scale(exp(xm)), lo + 1, hi - 1);
    double s = 0, ms = 0;
    int rc = probe(mid, &s, &ms);
    if (rc) return rc;
    bool ok = s >= target;
    out->probes.push_back({mid, s, ms, ok});
    if (ok) { hi = mid; have_feasible = true; }
    else lo = mid;
  }
  const ProbeResult* best = nullptr;
  for (auto& p : out->probes)
    if (p.feasible && (!best || p.rung < best->rung)) best = &p;
  if (!best)
    for (auto& p : out->probes)
      if (!best || p.rung > best->rung) best = &p;
  if (!best) return JPXLG_E_INVALID;
  (void)have_feasible;
  out->rung = best->rung;
  out->score = best->score;
  out->met = best->feasible ? 1 : 0;
  return 0;
}

}  // namespace jpxlg
