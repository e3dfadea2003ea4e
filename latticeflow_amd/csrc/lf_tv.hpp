// Exact 2-D total variation: host-side fixed-point setup and output mapping
// of TVSolver (solver/tv/tv_solver.hpp:71-112, 151-176, 181-244).
#pragma once

#include <cstdint>
#include <vector>

namespace lfamd {

constexpr int kTvFunctionBits = 40;  // int64: 7*8 - 2*8 bits (pf_flow_node.hpp:103-115)

// Per-image part of TVSolver::setup: mean M, scale S = 2 max|f - M| and the
// 40-bit function values uf = round((0.5 + (f - M) / S) 2^40).
struct TvBase {
  double M = 0, S = 0;
  std::vector<int64_t> uf;
};

TvBase tv_base(int64_t n, const double* f);
// Pairwise value of an edge of geocut weight w at lambda:
// round(((0.5 / (S max(lambda, 1e-8))) w / S) 2^40) (toUnshiftedFVDType).
int64_t tv_pwf(const TvBase& b, double lambda, double w);
// runSingleLambda's output mapping (tv_solver.hpp:158-175) of final raw r.
void tv_output_single(const TvBase& b, const int64_t* r, int64_t n, double* R);
// getRegularizationPath's output mapping (tv_solver.hpp:219-240) of a path
// value r_path (raw units, not halved).
void tv_output_path(const TvBase& b, const double* r_path, int64_t n, double* R);

}  // namespace lfamd
