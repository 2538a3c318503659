// The quality controller: the quantizer ladder and the probe navigator.
#pragma once
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

// The per-effort request of EncodeRequest::for_quality (request.rs:955-972)
// and the budget of QualityBudget::for_preset (quality.rs:395-420).
struct QualityPolicy {
  uint32_t quant_lf;
  uint32_t x_qm, b_qm;
  int32_t truncate;
  float lambda_scale;
  float zero_token_bits;
  int32_t epf_iters;
  int32_t sharpness;
  int32_t cover_mode;
  int32_t pixel_probes;
  double reserve;
  jpxlg_quantizer quantizer_at(uint32_t rung) const;
};
QualityPolicy quality_policy(int32_t effort);

struct ProbeResult {
  uint32_t rung;
  double score;
  double ms;
  bool feasible;
};

struct ControllerOutcome {
  uint32_t rung = 0;
  double score = 0.0;
  int met = 0;
  std::vector<ProbeResult> probes;
};

// probe(rung, &score, &ms) -> status.
using ProbeFn = std::function<int(uint32_t, double*, double*)>;

// Bisection over the ladder for the coarsest rung meeting `target`.
int bisect_controller(double target, int max_probes, const ProbeFn& probe, ControllerOutcome* out);
// The JPXL navigator (quality.rs:1775-1990): predicted probe, slope-
// extrapolated bracketing, log-loss crossing aims with the effort's reserve,
// one rescue probe; picks the coarsest feasible probed rung.
int navigate_controller(double target, int32_t effort, int max_probes, const ProbeFn& probe,
                        ControllerOutcome* out, uint32_t predicted_rung = 0);

// Exposed for the host-side tests of the navigator.
uint32_t log_loss_crossing(uint32_t lo_rung, double lo_score, uint32_t hi_rung, double hi_score,
                           double aim_score, int* ok);
uint32_t geometric_step(uint32_t from, int finer);
double aim_score_for(double threshold, double reserve);

}  // namespace jpxlg
