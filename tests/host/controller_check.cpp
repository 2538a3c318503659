// Host-only test shim over the library's controller and metric pooling
// (jpxl_amd/csrc/controller.cpp, metric_host.cpp), compiled with g++ by the
// CPU tests: the navigator runs against closed-form score curves, the way
// the reference's controller tests drive it with their CurveEvaluator
// (jpxl-encode-policy/src/quality.rs:2752-2797).  Test infrastructure only.
#include <math.h>
#include <string.h>

#include "../../jpxl_amd/csrc/controller.h"
#include "../../jpxl_amd/csrc/metric_host.h"
#include "../../jpxl_amd/csrc/predictor.h"

using namespace jpxlg;

extern "C" {

struct cc_outcome {
  uint32_t rung;
  double score;
  int32_t met, status, chosen_fresh;
  uint32_t anchor_rung;
  int32_t pixel_probes, structural_builds, priced, nprobes;
  uint32_t probe_rungs[32];
  double probe_scores[32];
  int32_t probe_fresh[32];
};

double cc_met_band(double t, double r) { return met_band(t, r); }
double cc_aim(double t, double r) { return aim_score_for(t, r); }
uint32_t cc_log_loss_crossing(uint32_t lo, double ls, uint32_t hi, double hs, double aim, int* ok) {
  return log_loss_crossing(lo, ls, hi, hs, aim, ok);
}
uint32_t cc_geometric_step(uint32_t from, int finer) { return geometric_step(from, finer); }
uint32_t cc_top() { return Rung::top(); }
uint64_t cc_effective_scale(uint32_t rung) { return Rung::effective_scale(rung); }
uint32_t cc_rung_for_scale(double s) { return rung_for_scale(s); }
double cc_predicted_scale(float q50, float flat, double target) {
  return predicted_effective_scale(q50, flat, target);
}
void cc_features(const float* v, int n, float* q50, float* flat) {
  source_features(std::vector<float>(v, v + n), q50, flat);
}
int cc_policy(int effort, int32_t* out6, double* reserve) {
  const QualityPolicy p = quality_policy(effort);
  out6[0] = p.pixel_probes;
  out6[1] = p.exact_prices;
  out6[2] = p.structural_builds;
  out6[3] = p.enable_cfl;
  out6[4] = p.cover_mode;
  out6[5] = p.truncate;
  *reserve = p.reserve;
  return 0;
}

// Navigates a curve loss = a * (scale / 1000)^-b in the rung's effective
// scale; a fresh (rebuilt) structure away from the anchor scores `fresh_gain`
// points higher.  effort < 0: the explicit budget; mode 1: bisection.
int cc_navigate_curve(double target, int effort, int pixel, int exact, int builds, double reserve,
                      uint32_t predicted, double a, double b, double fresh_gain, int mode, cc_outcome* out) {
  memset(out, 0, sizeof(*out));
  uint32_t anchor = 0;
  bool anchored = false;
  ProbeFn probe = [&](uint32_t rung, bool fresh, double* score, double* ms) -> int {
    const double scale = (double)Rung::effective_scale(rung);
    double s = 100.0 - a * pow(scale / 1000.0, -b);
    if (fresh && anchored && rung != anchor) s += fresh_gain;
    if (fresh && !anchored) {
      anchored = true;
      anchor = rung;
    }
    *score = s;
    *ms = 0.0;
    return 0;
  };
  ControllerOutcome oc;
  int rc;
  if (mode == 1) rc = bisect_controller(target, pixel, probe, &oc);
  else if (effort >= 0) rc = navigate_controller(target, effort, pixel, probe, &oc, predicted);
  else rc = navigate_with_budget(target, NavBudget{pixel, exact, builds, reserve}, probe, &oc, predicted);
  if (rc) return rc;
  out->rung = oc.rung;
  out->score = oc.score;
  out->met = oc.met;
  out->status = oc.status;
  out->chosen_fresh = oc.chosen_fresh;
  out->anchor_rung = oc.anchor_rung;
  out->pixel_probes = oc.pixel_probes;
  out->structural_builds = oc.structural_builds;
  out->priced = oc.priced;
  out->nprobes = (int32_t)oc.probes.size();
  for (size_t i = 0; i < oc.probes.size() && i < 32; i++) {
    out->probe_rungs[i] = oc.probes[i].rung;
    out->probe_scores[i] = oc.probes[i].score;
    out->probe_fresh[i] = oc.probes[i].fresh;
  }
  return 0;
}

// The navigator over a caller's probe (the CPU restatement's whole encode):
// probe(rung, fresh, &score) -> status.
typedef int (*cc_probe_fn)(uint32_t rung, int fresh, double* score);
int cc_navigate_callback(double target, int effort, int max_probes, uint32_t predicted, double prior_beta, int mode,
                         cc_probe_fn fn, cc_outcome* out) {
  memset(out, 0, sizeof(*out));
  ProbeFn probe = [&](uint32_t rung, bool fresh, double* score, double* ms) -> int {
    *ms = 0.0;
    return fn(rung, fresh ? 1 : 0, score);
  };
  ControllerOutcome oc;
  const int rc = mode == 1 ? bisect_controller(target, max_probes, probe, &oc)
                           : navigate_controller(target, effort, max_probes, probe, &oc, predicted, prior_beta);
  if (rc) return rc;
  out->rung = oc.rung;
  out->score = oc.score;
  out->met = oc.met;
  out->status = oc.status;
  out->chosen_fresh = oc.chosen_fresh;
  out->anchor_rung = oc.anchor_rung;
  out->pixel_probes = oc.pixel_probes;
  out->structural_builds = oc.structural_builds;
  out->priced = oc.priced;
  out->nprobes = (int32_t)oc.probes.size();
  for (size_t i = 0; i < oc.probes.size() && i < 32; i++) {
    out->probe_rungs[i] = oc.probes[i].rung;
    out->probe_scores[i] = oc.probes[i].score;
    out->probe_fresh[i] = oc.probes[i].fresh;
  }
  return 0;
}

// The quantizer and planning tier of a quality probe at `rung`.
int cc_rung_policy(uint32_t rung, int effort, uint32_t* fields6, int32_t* tiers3, float* ls_ztb) {
  const QualityPolicy p = quality_policy(effort);
  const jpxlg_quantizer q = p.quantizer_at(rung);
  fields6[0] = q.global_scale;
  fields6[1] = q.hf_mul;
  fields6[2] = q.quant_lf;
  fields6[3] = q.x_qm_scale;
  fields6[4] = q.b_qm_scale;
  fields6[5] = (uint32_t)q.truncate;
  tiers3[0] = p.cover_mode;
  tiers3[1] = p.enable_cfl;
  tiers3[2] = p.epf_iters * 16 + p.sharpness;
  ls_ztb[0] = q.lambda_scale;
  ls_ztb[1] = q.zero_token_bits;
  return 0;
}

double cc_pool_terms(const double* sums, int ns, const int* sw, const int* sh, double* terms, double* raw) {
  MetricResult m = pool_terms(sums, ns, sw, sh);
  memcpy(terms, m.terms, sizeof(double) * 18 * ns);
  *raw = m.raw_error;
  return m.score;
}

// The first probe and prior exponent of a quality search from a frame's
// features (the library's routing of the crossing predictor).
int cc_quality_start(const jpxlg_source_features* f, const jpxlg_transform_summary* t, double target, int effort,
                     uint32_t legacy, uint32_t* rung, double* prior_beta) {
  const QualityStart q = quality_start(*f, *t, target, quality_policy(effort).reserve, legacy);
  *rung = q.rung;
  *prior_beta = q.prior_beta;
  return 0;
}


// The navigator with exact finalist pricing (quality.rs:1886-2010) over a
// caller's probe and price: price(rung, fresh, &bytes) -> status.
typedef int (*cc_price_fn)(uint32_t rung, int fresh, uint64_t* bytes);
int cc_navigate_priced(double target, int effort, int max_probes, uint32_t predicted, cc_probe_fn fn,
                       cc_price_fn pfn, cc_outcome* out, int32_t* exact_prices, uint64_t* bytes) {
  memset(out, 0, sizeof(*out));
  ProbeFn probe = [&](uint32_t rung, bool fresh, double* score, double* ms) -> int {
    *ms = 0.0;
    return fn(rung, fresh ? 1 : 0, score);
  };
  PriceFn price = [&](uint32_t rung, bool fresh, uint64_t* b) -> int { return pfn(rung, fresh ? 1 : 0, b); };
  ControllerOutcome oc;
  const int rc = navigate_controller(target, effort, max_probes, probe, &oc, predicted, NAN, pfn ? &price : nullptr);
  if (rc) return rc;
  out->rung = oc.rung;
  out->score = oc.score;
  out->met = oc.met;
  out->status = oc.status;
  out->chosen_fresh = oc.chosen_fresh;
  out->anchor_rung = oc.anchor_rung;
  out->pixel_probes = oc.pixel_probes;
  out->structural_builds = oc.structural_builds;
  out->priced = oc.priced;
  out->nprobes = (int32_t)oc.probes.size();
  for (size_t i = 0; i < oc.probes.size() && i < 32; i++) {
    out->probe_rungs[i] = oc.probes[i].rung;
    out->probe_scores[i] = oc.probes[i].score;
    out->probe_fresh[i] = oc.probes[i].fresh;
  }
  *exact_prices = oc.exact_prices;
  *bytes = oc.bytes;
  return 0;
}

// The rate loop's ladder search over a caller's price: price(rung, &bytes).
typedef int (*cc_rate_price_fn)(uint32_t rung, uint64_t* bytes);
int cc_rate_search(uint32_t start, uint64_t target, uint64_t tol_bytes, double tol_fraction, uint32_t max_prices,
                   uint32_t fill_probes, cc_rate_price_fn fn, uint32_t* rung, uint64_t* bytes, int32_t* saturated,
                   uint64_t* floor_bytes, uint32_t* trace_rungs, uint64_t* trace_bytes, int32_t* ntrace) {
  RateOutcome oc;
  RatePriceFn price = [&](uint32_t r, uint64_t* b) -> int { return fn(r, b); };
  const int rc = rate_search_ladder(start, target, tol_bytes, tol_fraction, max_prices, fill_probes, price, &oc);
  *rung = oc.rung;
  *bytes = oc.bytes;
  *saturated = oc.saturated ? 1 : 0;
  *floor_bytes = oc.floor_bytes;
  *ntrace = (int32_t)oc.trace.size();
  for (size_t i = 0; i < oc.trace.size() && i < 256; i++) {
    trace_rungs[i] = oc.trace[i].rung;
    trace_bytes[i] = oc.trace[i].bytes;
  }
  return rc;
}

uint32_t cc_rate_cold_step(uint32_t rung, int up) { return rate_cold_step(rung, up != 0); }

int cc_rate_interpolated(uint32_t lo, uint64_t lo_b, uint32_t hi, uint64_t hi_b, uint64_t target, uint32_t* out) {
  return rate_interpolated_rung(lo, lo_b, hi, hi_b, target, out) ? 1 : 0;
}

// Every rung's (global_scale, hf_mul) (quantizer_ladder.rs QuantizerChoice::at).
uint32_t cc_ladder_fields(uint32_t* gs, uint32_t* mul) {
  const uint32_t n = Rung::ladder_len();
  for (uint32_t r = 0; r < n; r++) Rung::fields(r, &gs[r], &mul[r]);
  return n;
}

}  // extern "C"
