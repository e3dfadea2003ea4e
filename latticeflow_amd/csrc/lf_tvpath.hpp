// Event-driven TV regularization path over lambda (TVSolver::run with
// _constructInitialRegPathsFromSolvedLattice and getRegularizationPath,
// solver/tv/tv_solver.hpp:114-244,300-642; TVRegPathSegment,
// solver/tv/tv_regpath_node.hpp:67-1028).
#pragma once

#include <cstdint>
#include <memory>

namespace lfamd {

struct TvPathStats {
  int64_t initial_regions = 0;    // level sets at lambda_max
  int64_t segments = 0;           // level sets of the path (each lives on an interval of lambda)
  int64_t joins = 0;              // join events
  int64_t splits = 0;             // split events
  int64_t checks = 0;             // level-set checks (one minimum cut of one level set each)
  int64_t host_checks = 0;        // of those, on the host
  int64_t device_calls = 0;       // batched device check launches
  int64_t dinkelbach_steps = 0;   // checks that located split points
  double initial_seconds = 0;     // exact solution at lambda_max
  double device_check_seconds = 0;
  double host_check_seconds = 0;
  double seconds = 0;             // whole construction
};

// The path of one image: built once, then evaluated at any lambda in
// [0, lambda_max].
class TvPath {
 public:
  // host_max: level sets of at most this many nodes are checked on the host,
  // larger ones on the device (0: the default; negative: all on the device).
  TvPath(const char* kernel, int64_t nx, int64_t ny, const double* f, double lambda_max, int64_t host_max);
  ~TvPath();
  void eval(int64_t n_lambda, const double* lambdas, double* R) const;
  const TvPathStats& stats() const;
  double lambda_max() const;
  int64_t nx() const;
  int64_t ny() const;

 private:
  struct Impl;
  std::unique_ptr<Impl> im_;
};

}  // namespace lfamd
