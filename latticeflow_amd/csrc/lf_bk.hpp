// CPU cross-check solver: the Boykov-Kolmogorov max-flow algorithm on a lattice
// network (lf_bk.cpp).  Registered as the solver "BKGCEnergyMinimizer", the
// name of the reference's optional BK wrapper
// (solver/energy_minimization/bk_graphcuts.hpp:65-163).
#pragma once

#include <cstdint>
#include <vector>

namespace lfamd {

class BkSolver {
 public:
  // shape: ndim extents; offs: k * ndim lattice-order directions (entry i + k/2
  // is the reverse of entry i).
  BkSolver(int ndim, const std::vector<int64_t>& shape, int k, const std::vector<int>& offs);
  // Energies in the reference layout.  Like the reference's BK filler
  // (bk_graphcuts.hpp:41-59, which assigns), a call replaces the energies
  // given before instead of adding to them.
  void set_energies(const int64_t* unary, const int64_t* binary);
  void run();
  // 1 = label 1 ("on"); which = 0: minimal such set, 1: maximal
  void cut(int32_t* out, int which) const;
  int64_t flow() const { return flow_; }
  bool solved() const { return solved_; }
  int64_t augmentations() const { return augmentations_; }

 private:
  int nd_, k_;
  int64_t n_;
  std::vector<int64_t> shape_, stride_;
  std::vector<int> offs_;
  std::vector<int64_t> unary_, binary_;
  std::vector<int64_t> tr_;   // terminal residual: > 0 from the source, < 0 to the sink
  std::vector<int64_t> res_;  // res_[v * k + i]: residual of the arc v -> v + dir i
  std::vector<int64_t> nbr_;  // neighbour of v along i, -1 outside the lattice
  std::vector<uint8_t> smin_, smax_;
  int64_t flow_ = 0, augmentations_ = 0;
  bool solved_ = false;
};

}  // namespace lfamd
