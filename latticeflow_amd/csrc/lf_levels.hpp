// Parametric divide-and-conquer ("LatticeLevelReductions") on the device.
#pragma once

#include <memory>

#include "lf_engine.hpp"

namespace lfamd {

class LevelSolver {
 public:
  virtual ~LevelSolver() {}
  virtual void invalidate() = 0;
  virtual void run(const lf_tuning& tuning) = 0;
  virtual bool solved() const = 0;
  // the level-reduction cut {r < 0} as int32 0/1 (pf_solver.hpp:181-182)
  virtual void cut_device(int32_t* d_out, hipStream_t s) = 0;
  lf_stats stats{};
};

std::unique_ptr<LevelSolver> make_level_solver(const LatticeGeom& g, MinCutEngine* eng);

}  // namespace lfamd
