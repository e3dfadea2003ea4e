// C++ template entry points of the reference's solvers.hpp that callers use
// directly: calculate2dTV for one lambda and for a regularization path
// (solver/tv/tv_solver.hpp:648-681), returning shared pointers to 2-D / 3-D
// double arrays with at(...) accessors (LatticeArray<double, 2|3>), with the
// kernels as template tags of the reference's names.  Implemented over the C
// ABI (lf_tv_single / lf_tv_path); the arithmetic runs on the device.
//
// The dtype template argument is accepted for source compatibility; the
// library's TV network always uses 64-bit integers (the reference's only
// instantiation in its Python layer, pylatticeflow.pyx:54-59).
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <vector>

#include "interface.hpp"

namespace latticeQBP {

// Row-major array of doubles with the shape of a lattice (the subset of the
// reference's LatticeArray that the TV results are read through).
template <typename T, int nd>
class LatticeArray {
 public:
  explicit LatticeArray(const std::array<size_t, nd>& shape) : shape_(shape) {
    size_t n = 1;
    for (size_t s : shape_) n *= s;
    data_.assign(n, T());
  }
  const std::array<size_t, nd>& shape() const { return shape_; }
  size_t size() const { return data_.size(); }
  T* data() { return data_.data(); }
  const T* data() const { return data_.data(); }
  template <typename... I>
  T& at(I... idx) {
    return data_[offset({(size_t)idx...})];
  }
  template <typename... I>
  const T& at(I... idx) const {
    return data_[offset({(size_t)idx...})];
  }

 private:
  size_t offset(const std::array<size_t, nd>& c) const {
    size_t o = 0;
    for (int d = 0; d < nd; ++d) {
      if (c[d] >= shape_[d]) throw std::out_of_range("LatticeArray index out of range");
      o = o * shape_[d] + c[d];
    }
    return o;
  }
  std::array<size_t, nd> shape_;
  std::vector<T> data_;
};

typedef std::shared_ptr<LatticeArray<double, 2> > FuncMapPtr;
typedef std::shared_ptr<LatticeArray<double, 3> > RegPathPtr;

// Kernel tags (the names of the reference's generated kernel types); the
// TV entry points accept the 2-D geocut kernels.
#define LF_AMD_KERNEL_TAG(NAME) \
  struct NAME {                 \
    static const char* name() { return #NAME; } \
  };
LF_AMD_KERNEL_TAG(Full2d_4)
LF_AMD_KERNEL_TAG(Full2d_8)
LF_AMD_KERNEL_TAG(Full2d_12)
LF_AMD_KERNEL_TAG(Full2d_20)
LF_AMD_KERNEL_TAG(Full2d_24)
LF_AMD_KERNEL_TAG(Full2d_28)
LF_AMD_KERNEL_TAG(Full2d_36)
LF_AMD_KERNEL_TAG(Full2d_48)
LF_AMD_KERNEL_TAG(Star2d_4)
LF_AMD_KERNEL_TAG(Star2d_8)
LF_AMD_KERNEL_TAG(Star2d_16)
LF_AMD_KERNEL_TAG(Star2d_24)
LF_AMD_KERNEL_TAG(Star2d_32)
#undef LF_AMD_KERNEL_TAG

// tv_solver.hpp:648-655: TV of the nx x ny image `function` (C order) at one lambda
template <typename Kernel, typename dtype = long>
FuncMapPtr calculate2dTV(size_t nx, size_t ny, double* function, double lambda) {
  FuncMapPtr R(new LatticeArray<double, 2>({nx, ny}));
  amd_detail::check(lf_tv_single(Kernel::name(), (int64_t)nx, (int64_t)ny, function, lambda, R->data(), nullptr));
  return R;
}

// tv_solver.hpp:657-681: the regularization path at the given lambdas (slice
// i of the result for lambda[i]); negative lambdas count as 0, and an empty
// list gives an empty path
template <typename Kernel, typename dtype = long>
RegPathPtr calculate2dTV(size_t nx, size_t ny, double* function, std::vector<double> lambda) {
  if (lambda.empty()) {
    std::fprintf(stderr, "Must supply list of regularization values at which to calculate the path.\n");
    return RegPathPtr(new LatticeArray<double, 3>({0, nx, ny}));
  }
  for (double& lm : lambda)
    if (lm < 0) {
      std::fprintf(stderr, "Ignoring reg value %g that is less than 0.\n", lm);
      lm = 0;
    }
  RegPathPtr R(new LatticeArray<double, 3>({lambda.size(), nx, ny}));
  amd_detail::check(lf_tv_path(Kernel::name(), (int64_t)nx, (int64_t)ny, function, (int64_t)lambda.size(),
                               lambda.data(), R->data(), nullptr));
  return R;
}

}  // namespace latticeQBP
