// The quality search's first guess: the frame features the crossing
// predictor reads (source features, 32x32-cell preanalysis, DCT8 summary),
// the generated case table and knot model, and the routing of their
// prediction into the navigator's first rung and prior loss exponent.
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/jpxlg.h"

namespace jpxlg {

// The integer sums of one whole 32x32 preanalysis cell
// (preanalysis.rs:408-484), as the device writes them.
struct CellStats {
  int32_t sum, sum_sq;          // luma
  int32_t edge, flat;           // pixels with Sobel |gx|+|gy| > 48, < 12
  int32_t noise_sum, weak;      // |c - box mean| over pixels with gradient < 24, their count
  int32_t dir[4];               // edge-direction histogram
  int32_t cb_sum, cb_sq, cr_sum, cr_sq;
  int32_t pad[2];
};
static_assert(sizeof(CellStats) == 64, "device record layout");

// The DCT8 summary's partial sums over a run of blocks in raster order
// (lib.rs:2791-2874), as the device writes them.
struct SummaryPartial {
  double total_y, total_low, total_high, total_row, total_col, total_xb, dc_sum, dc_sumsq;
  uint32_t near_1e3, near_1e2, blocks, ac_cells;
};
static_assert(sizeof(SummaryPartial) == 80, "device record layout");
constexpr int kSummaryBins = 512;  // lib.rs:2776-2779 HIST_BINS

// quality_features.rs:19 FLAT_VARIANCE.
constexpr float kFlatVariance = 2e-5f;

// SourceFeatures without the preanalysis (quality_features.rs:173-213) from
// the per-atom variances' order statistics (floor-indexed q10/q50/q90 of
// luma, q50 of chroma) and the count of atoms below FLAT_VARIANCE among n.
void source_features_set(float q10, float q50, float q90, float chroma_q50, uint32_t flat, uint32_t n,
                         jpxlg_source_features* f);

// PreanalysisFeatures of the cell records (preanalysis.rs:486-540).
void preanalysis_from_cells(const CellStats* cells, int n, jpxlg_source_features* f);

// The frame-level summary from the partials and the histogram of per-block
// ln(Y AC energy) bins (lib.rs:2876-2955).
void transform_summary_from(const SummaryPartial* parts, int n, const uint32_t* hist, jpxlg_transform_summary* t);

// quality_prediction::predict (which = JPXLG_PREDICTOR_DISPATCH),
// predict_cases or predict_v2.  Returns false when no prediction exists.
bool predict_crossing(const jpxlg_source_features& f, const jpxlg_transform_summary& t, double target,
                      double reserve, int which, jpxlg_crossing_prediction* out);

// The first probe and the navigator's prior loss exponent
// (quality.rs:2485-2530, the default build): a confident prediction's
// candidate rung, an uncertain in-distribution one's median rung (both with
// its exponent), the median of a frame whose only out-of-distribution signal
// is its size, else the legacy table's rung and no prior (NaN).
struct QualityStart {
  uint32_t rung;
  double prior_beta;
};
QualityStart quality_start(const jpxlg_source_features& f, const jpxlg_transform_summary& t, double target,
                           double reserve, uint32_t legacy_rung);

}  // namespace jpxlg
