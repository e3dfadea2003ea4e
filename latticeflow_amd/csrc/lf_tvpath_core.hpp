// The event loop of the TV regularization path, independent of the device:
// level sets at s = 1 in, the whole path out.  Checks of large level sets go
// to a RegionChecker (the device); small ones run on the host.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "lf_tvpath.hpp"

namespace lfamd {

// One batch of level-set checks: every node v has a region reg[v] in
// [0, nreg) or -1 and a value r[v]; each region is cut alone at its mean
// (ceil of the mean, one below the maximum when they are equal; a region whose
// values span at most one unit is whole).  Writes the size of the low side
// of the minimal cut per region (0: the region is a level set) and whether
// each node is on it.
class RegionChecker {
 public:
  virtual ~RegionChecker() {}
  virtual void check(const int64_t* r, const int32_t* reg, int nreg, int64_t* cnt_low, uint8_t* low) = 0;
};

struct TvPathProblem {
  int64_t n = 0;
  int k = 0;
  std::vector<int32_t> nbr;   // n * k neighbours in lattice direction order, -1 outside
  std::vector<int64_t> acap;  // raw capacity per direction (direction i + k/2 reverses i)
  std::vector<int64_t> fv;    // raw function value per node at s = 1
  std::vector<int32_t> reg;   // level sets at s = 1 (any ids, >= 0)
  std::vector<int64_t> r;     // node values at s = 1 of a maximum flow with those level sets
};

// The segment table a trace needs: every pixel starts at initial[v]; segment
// g is alive on [s_lo[g], s_hi] with the value a[g] + s b[g]; below s_lo it
// continues in up[g] (a join) or, when up[g] < 0, in the part of split
// -(up[g] + 1) that holds the pixel (parts: part_begin/part_end into part_seg;
// members of part p: mem_begin[p]..mem_begin[p + 1] into members, sorted).
struct TvTrace {
  std::vector<int32_t> initial;
  std::vector<double> s_lo, a, b;
  std::vector<int32_t> up;
  std::vector<int32_t> part_begin, part_end, part_seg;
  std::vector<int64_t> mem_begin;
  std::vector<int32_t> members;
};

class TvPathCore {
 public:
  // host_max: level sets of at most host_max nodes are checked on the host,
  // larger ones by `device` (null: every check on the host).
  TvPathCore(TvPathProblem prob, RegionChecker* device, int64_t host_max, TvPathStats* stats);
  ~TvPathCore();
  // raw path values of every node at s in [0, 1] (host trace)
  void values(double s, double* rp) const;
  TvTrace trace_table() const;

 private:
  struct Impl;
  std::unique_ptr<Impl> im_;
};

}  // namespace lfamd
