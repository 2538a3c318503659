// Host half of SSIMULACRA2: from pooled sums to the score.
#pragma once

namespace jpxlg {

struct MetricResult {
  double score;
  double raw_error;
  double terms[108];  // (component, scale, norm, map) order
};

// sums: per scale 18 doubles, [channel][sum d, sum d^4, sum art, sum art^4,
// sum lost, sum lost^4].  pool.rs:60-68 from_sums, ssimulacra2.rs:509-540.
MetricResult pool_terms(const double* sums, int nscales, const int* sw, const int* sh);

}  // namespace jpxlg
