// Parametric divide-and-conquer on the device (placeholder until the keyed
// region solver lands): creation fails with LF_ERR_UNSUPPORTED.
#include "lf_levels.hpp"

namespace lfamd {

std::unique_ptr<LevelSolver> make_level_solver(const LatticeGeom&, MinCutEngine*) {
  fail(LF_ERR_UNSUPPORTED, "LatticeLevelReductions has no device path yet");
}

}  // namespace lfamd
