// The quality controller: the quantizer ladder, the first-guess predictor
// and the probe navigator.
#pragma once
#include <math.h>
#include <stdint.h>

#include <functional>
#include <vector>

#include "../../include/jpxlg.h"

namespace jpxlg {

// jpxl-encode/src/vardct/ids.rs:20 MAX_GLOBAL_SCALE.
constexpr uint32_t kMaxGlobalScale = 73728;

// jpxl-encode-policy/src/quantizer_ladder.rs: the dense rung ladder.
namespace Rung {
uint32_t ladder_len();
uint32_t top();
uint32_t for_global_scale(uint32_t scale);
void fields(uint32_t rung, uint32_t* global_scale, uint32_t* hf_mul);
uint64_t effective_scale(uint32_t rung);
uint32_t for_effective_scale(uint64_t scale);
uint32_t coupled_quant_lf(uint32_t base, uint32_t k);
}  // namespace Rung

// The per-effort request of EncodeRequest::for_quality (request.rs:955-972),
// the planning tiers (quality.rs:1737-1749) and the budget of
// QualityBudget::for_preset (quality.rs:395-420).
struct QualityPolicy {
  uint32_t quant_lf;
  uint32_t x_qm, b_qm;
  int32_t truncate;
  float lambda_scale;
  float zero_token_bits;
  int32_t epf_iters;
  int32_t sharpness;
  int32_t cover_mode;
  int32_t enable_cfl;
  int32_t pixel_probes;
  int32_t exact_prices;
  int32_t structural_builds;
  double reserve;
  jpxlg_quantizer quantizer_at(uint32_t rung) const;
};
QualityPolicy quality_policy(int32_t effort);

// QualityStatus (quality.rs:440-470), in the order of jpxlg.h.
enum QualityStatusCode {
  kStatusMet = JPXLG_QUALITY_MET,
  kStatusMetAdjacentRungs = JPXLG_QUALITY_MET_ADJACENT_RUNGS,
  kStatusMetWorkCap = JPXLG_QUALITY_MET_WORK_CAP,
  kStatusSaturatedTop = JPXLG_QUALITY_SATURATED_TOP,
  kStatusSaturatedFloor = JPXLG_QUALITY_SATURATED_FLOOR,
  kStatusUnderTargetWorkCap = JPXLG_QUALITY_UNDER_TARGET_WORK_CAP,
  kStatusRescuedFreshStructure = JPXLG_QUALITY_RESCUED_FRESH_STRUCTURE,
};

// The navigator's per-solve caps (QualityBudget, quality.rs:395-420).
struct NavBudget {
  int32_t pixel_probes;
  int32_t exact_prices;
  int32_t structural_builds;
  double reserve;
};

struct ProbeResult {
  uint32_t rung;
  double score;
  double ms;
  bool feasible;
  bool fresh;  // the probe built its own cover + CfL at its rung
};

struct ControllerOutcome {
  uint32_t rung = 0;
  double score = 0.0;
  int met = 0;
  int status = kStatusMet;
  // The structure the chosen plan was scored with: the anchor's (the first
  // probe's fresh build) or one rebuilt at the chosen rung.
  bool chosen_fresh = false;
  uint32_t anchor_rung = 0;
  int pixel_probes = 0;
  int structural_builds = 0;
  int priced = 0;
  int exact_prices = 0;     // finalists priced by the price function
  uint64_t bytes = 0;       // the chosen finalist's exact size, when priced
  std::vector<ProbeResult> probes;
};

// probe(rung, fresh, &score, &ms) -> status.  fresh: build the cover and CfL
// at this rung's quantizer (the first probe's build is the anchor); otherwise
// reuse the anchor's structure.
using ProbeFn = std::function<int(uint32_t, bool, double*, double*)>;
// price(rung, fresh, &bytes) -> status: the exact codestream size of the plan
// at this rung on the anchor's structure (fresh: on its own).
using PriceFn = std::function<int(uint32_t, bool, uint64_t*)>;

// Bisection over the ladder for the coarsest rung meeting `target` (every
// probe on the first probe's structure).
int bisect_controller(double target, int max_probes, const ProbeFn& probe, ControllerOutcome* out);

// solve_baseline (quality.rs:1775-2050) with the reference's default levers:
// the predicted first probe builds the anchor structure; a first probe
// landing feasible inside the accept band ends the search
// (trusted-first-probe); otherwise one bounded direct correction and one
// slope-corrected retry (corrected-stop), then slope-extrapolated bracketing,
// log-loss crossing aims at target + reserve while the coarsest feasible
// probe lies outside the reserve-coupled accept band, one rescue probe when
// nothing met.  Finalists are the two coarsest feasible probes; one far from
// the anchor is rebuilt with its own structure and re-scored.  With a price
// function the feasible finalists are priced exactly and the smallest stream
// is chosen (the first of equal sizes), as the reference's default build
// does; a single feasible finalist is chosen without pricing (its size is
// the emitted stream's).  Without one the coarsest feasible rung is chosen.
// prior_beta: the crossing predictor's local loss exponent, which replaces
// PRIOR_LOSS_EXPONENT in the slope-extrapolated steps (NaN: none).
int navigate_controller(double target, int32_t effort, int max_probes, const ProbeFn& probe,
                        ControllerOutcome* out, uint32_t predicted_rung, double prior_beta = NAN,
                        const PriceFn* price = nullptr);
int navigate_with_budget(double target, const NavBudget& budget, const ProbeFn& probe, ControllerOutcome* out,
                         uint32_t predicted_rung, double prior_beta = NAN, const PriceFn* price = nullptr);

// The legacy initial-quantizer predictor (quality.rs:1021-1079 over the
// table of quality_predictor.rs): the effective scale the calibration
// predicts for `target` from a frame's median per-atom luma variance and its
// flat fraction.
double predicted_effective_scale(float luma_variance_q50, float flat_fraction, double target);
uint32_t rung_for_scale(double scale);

// Frame features of quality_features.rs:173-213 from per-atom variances
// (XYB Y, and X + B): the floor-indexed median and the fraction below
// FLAT_VARIANCE.
void source_features(std::vector<float> luma, float* luma_q50, float* flat_fraction);

// Exposed for the host-side tests of the navigator.
uint32_t log_loss_crossing(uint32_t lo_rung, double lo_score, uint32_t hi_rung, double hi_score,
                           double aim_score, int* ok);
uint32_t geometric_step(uint32_t from, int finer);
double aim_score_for(double threshold, double reserve);
double met_band(double threshold, double reserve);

// The exact rate loop's ladder search (search_ladder, rate.rs:387-921) with a
// cold bracket: price(rung, &bytes) -> status.  Returns 0 with the largest
// feasible size found (the finer rung of equal sizes), kRateUnreachable
// when even the coarsest rung is over the target (floor_bytes = its size),
// kRateBudgetExhausted when no feasible rung was priced within max_prices,
// or the price function's status.
enum { kRateBracket = 0, kRateBisect = 1, kRateFill = 2 };
constexpr int kRateUnreachable = 1;
constexpr int kRateBudgetExhausted = 2;
struct RateStep {
  uint32_t rung;
  uint64_t bytes;
  bool feasible;
  int phase;
};
struct RateOutcome {
  uint32_t rung = 0;
  uint64_t bytes = 0;
  bool saturated = false;  // the ladder's top was priced feasible
  uint64_t floor_bytes = 0;
  std::vector<RateStep> trace;
};
using RatePriceFn = std::function<int(uint32_t, uint64_t*)>;
// tol_bytes, tol_fraction: RateTolerance (the default is 8 bytes, 1 %).
int rate_search_ladder(uint32_t start, uint64_t target, uint64_t tol_bytes, double tol_fraction, uint32_t max_prices,
                       uint32_t fill_probes, const RatePriceFn& price, RateOutcome* out);
uint32_t rate_cold_step(uint32_t current, bool up);
bool rate_interpolated_rung(uint32_t lo_rung, uint64_t lo_bytes, uint32_t hi_rung, uint64_t hi_bytes,
                            uint64_t target, uint32_t* out);

}  // namespace jpxlg
