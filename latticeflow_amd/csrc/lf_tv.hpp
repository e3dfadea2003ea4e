// Exact 2-D total variation: host-side fixed-point setup and output mapping
// of TVSolver (solver/tv/tv_solver.hpp:71-112, 151-176).
#pragma once

#include <cstdint>
#include <vector>

namespace lfamd {

constexpr int kTvFunctionBits = 40;  // int64: 7*8 - 2*8 bits (pf_flow_node.hpp:103-115)

struct TvSetup {
  double M = 0, S = 0;
  std::vector<int64_t> red, alpha;  // alpha node-major, lattice direction order
};

TvSetup tv_setup(int64_t nx, int64_t ny, const double* f, double lambda, const std::vector<int>& offs,
                 const std::vector<double>& w, int k);
void tv_output(const TvSetup& t, const int64_t* r, int64_t n, double* R);

}  // namespace lfamd
