// Exact 2-D total variation: host-side fixed-point setup and output mapping.
//
// Follows the reference's operation order exactly (sequential double sums in
// node order, round-half-away-from-zero conversions), so the device divide and
// conquer receives bit-identical integer networks.

#include "lf_tv.hpp"

#include <algorithm>
#include <cmath>

namespace lfamd {

static inline int64_t to_fv(double v) { return (int64_t)std::round(v * (double)((int64_t)1 << kTvFunctionBits)); }

TvBase tv_base(int64_t n, const double* f) {
  TvBase b;
  double v_min = f[0], v_max = f[0], v_sum = f[0];
  for (int64_t i = 1; i < n; ++i) {
    const double v = f[i];
    v_min = std::min(v_min, v);
    v_max = std::max(v_max, v);
    v_sum += v;
  }
  b.M = v_sum / (double)n;
  v_min -= b.M;
  v_max -= b.M;
  b.S = 2 * std::max(v_max, -v_min);
  b.uf.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) b.uf[i] = to_fv(0.5 + (f[i] - b.M) / b.S);
  return b;
}

int64_t tv_pwf(const TvBase& b, double lambda, double w) {
  const double gsv = 0.5 / (b.S * std::max(lambda, 1e-8));
  return to_fv((gsv * w) / b.S);
}

void tv_output_single(const TvBase& b, const int64_t* r, int64_t n, double* R) {
  double sum = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double v = ((double)(r[i] / 2) / (double)((int64_t)1 << kTvFunctionBits) - 0.5) * b.S + b.M;
    R[i] = v;
    sum += v;
  }
  for (int64_t i = 0; i < n; ++i) R[i] = (R[i] - sum / (double)n) * b.S + b.M;
}

void tv_output_path(const TvBase& b, const double* r_path, int64_t n, double* R) {
  double sum = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double v = (r_path[i] / (double)((int64_t)1 << kTvFunctionBits) - 0.5) * b.S + b.M;
    R[i] = v;
    sum += v;
  }
  const double mean = sum / (double)n;
  for (int64_t i = 0; i < n; ++i) R[i] = (R[i] - mean) * b.S + b.M;
}

}  // namespace lfamd
