// Exact 2-D total variation: host-side setup and output mapping.
//
// Single lambda: TVSolver::setup + runSingleLambda (solver/tv/tv_solver.hpp:71-112,
// 151-176) with the int64 fixed point of PFFlowNode (pf_flow_node.hpp:103-143):
// 40-bit function values, unary E1 = round((0.5 + (f - M)/S) 2^40), pairwise
// E01 = E10 = round((0.5 w_e / (S lambda)) / S 2^40), then the parametric
// divide and conquer on the device, then the reference's output mapping
// R = ((r/2) 2^-40 - 0.5) S + M, re-centred and scaled by S once more.
// The host parts follow the reference's operation order exactly (sequential
// double sums in node order), so R agrees to the last quantum.

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "lf_tv.hpp"

namespace lfamd {

static inline int64_t tv_to_fv(double v) {
  return (int64_t)std::round(v * (double)((int64_t)1 << kTvFunctionBits));
}

// TVSolver::setup (tv_solver.hpp:71-112)
TvSetup tv_setup(int64_t nx, int64_t ny, const double* f, double lambda, const std::vector<int>& offs,
                 const std::vector<double>& w, int k) {
  TvSetup t;
  const int64_t n = nx * ny;
  double v_min = f[0], v_max = f[0], v_sum = f[0];
  for (int64_t i = 1; i < n; ++i) {
    const double v = f[i];
    v_min = std::min(v_min, v);
    v_max = std::max(v_max, v);
    v_sum += v;
  }
  t.M = v_sum / (double)n;
  v_min -= t.M;
  v_max -= t.M;
  t.S = 2 * std::max(v_max, -v_min);
  const double max_lambda = std::max(lambda, 1e-8);
  t.red.assign((size_t)n, 0);
  t.alpha.assign((size_t)n * k, 0);
  for (int64_t i = 0; i < n; ++i) t.red[i] += 2 * tv_to_fv(0.5 + (f[i] - t.M) / t.S);
  const double gsv = 0.5 / (t.S * max_lambda);
  for (int64_t x = 0; x < nx; ++x)
    for (int64_t y = 0; y < ny; ++y) {
      const int64_t v = x * ny + y;
      for (int e = 0; e < k / 2; ++e) {
        const int64_t ux = x + offs[e * 2], uy = y + offs[e * 2 + 1];
        if (ux < 0 || ux >= nx || uy < 0 || uy >= ny) continue;
        const int64_t u = ux * ny + uy;
        const int64_t pwf = tv_to_fv((gsv * w[e]) / t.S);
        t.alpha[(size_t)v * k + e] += 2 * pwf;
        t.alpha[(size_t)u * k + (e + k / 2) % k] += 2 * pwf;
      }
    }
  return t;
}

// runSingleLambda's output mapping (tv_solver.hpp:158-175)
void tv_output(const TvSetup& t, const int64_t* r, int64_t n, double* R) {
  double sum = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double v = ((double)(r[i] / 2) / (double)((int64_t)1 << kTvFunctionBits) - 0.5) * t.S + t.M;
    R[i] = v;
    sum += v;
  }
  for (int64_t i = 0; i < n; ++i) R[i] = (R[i] - sum / (double)n) * t.S + t.M;
}

}  // namespace lfamd

