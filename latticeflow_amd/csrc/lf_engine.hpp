// Device min-cut engine interface (host side).  The kernels live in lf_mincut.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/latticeflow_amd.h"
#include "lf_error.hpp"

namespace lfamd {

#define LF_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      ::lfamd::fail(LF_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Lattice geometry handed to the engine: every lattice is viewed as (D, H, W)
// (1-D: (1,1,n); 2-D: (1,n0,n1); 3-D: (n0,n1,n2)) with directions in the
// reference lattice order, entry i + k/2 the reverse of entry i.
struct LatticeGeom {
  int ndim = 0;
  int64_t D = 1, H = 1, W = 1;
  int k = 0;
  std::vector<int> dz, dy, dx;
  // the lattice's own shape and directions (k * ndim, lattice order): the
  // node-parallel engine of lattices with more than three dimensions uses these
  std::vector<int64_t> dims;
  std::vector<int> offsets;
  int64_t n() const { return D * H * W; }
};

// One lattice of a batch: a (h, w) 2-D lattice placed at (y0, x0) of the
// engine's canvas; its nodes are entries [off, off + h*w) of the batch's
// concatenated node order.  perm[e] = this lattice's edge index of the
// canvas's positive direction e (the lattice order depends on the shape).
struct BatchRect {
  int64_t y0, x0, h, w, off;
  int8_t perm[16];
};

class MinCutEngine {
 public:
  virtual ~MinCutEngine() {}
  // Accumulate energies (reference layout, device pointers) into the base network.
  virtual void add_energies_device(const int64_t* d_unary, const int64_t* d_binary, hipStream_t s) = 0;
  virtual void reset_energies() = 0;
  // Solve the binary min-cut of the accumulated network.
  virtual void run(const lf_tuning& tuning) = 0;
  // Cut of the last solve (LF_CUT_SMIN / LF_CUT_SMAX), as int32 0/1 into device memory.
  virtual void cut_device(int32_t* d_out, int which, hipStream_t s) = 0;
  virtual hipStream_t stream() const = 0;
  // Parametric divide and conquer over the accumulated network
  // (ParametricFlowSolver::run, pf_solver.hpp:33-208): final r per node.
  virtual void pf_run(const lf_tuning& tuning) = 0;
  // The same over a stack of `layers` equal lattices along the slowest axis
  // (no arcs between layers): one initial region per layer.
  virtual void pf_run_layers(const lf_tuning& tuning, int layers) = 0;
  // The same from an initial partition: regions of neighbouring nodes whose
  // predicted values (d_pred, one float per node) differ by at most delta,
  // with the cut arcs checked afterwards and failing regions merged (at most
  // max_repairs times; then whole layers).  The result is the unique solution.
  // Predicted values also choose thresholds: a region cuts at the emptiest gap
  // of its predicted values (in r units: pred * pscale + poff) while it has
  // one, and cuts at its mean (the plain rule) once all gaps are used.
  virtual void pf_run_layers_guided(const lf_tuning& tuning, int layers, const float* d_pred, float delta,
                                    int max_repairs, double pscale, double poff, lf_guided_stats* gs) = 0;
  // TV network of a stack of images: uf[H*W] per-pixel function values, pwf[layers][k/2]
  // per-direction pairwise values (tv_solver.hpp:71-112).
  virtual void set_tv_network(const int64_t* uf, const int64_t* pwf, int layers) = 0;
  virtual void pf_values_device(int64_t* d_out, hipStream_t s) = 0;
  virtual void pf_cut_device(int32_t* d_out, hipStream_t s) = 0;  // r < 0
  // Region id and flow-consistent value r of every node after a divide and
  // conquer (device pointers).
  virtual void pf_regions_device(int32_t* d_reg, int64_t* d_r, hipStream_t s) = 0;
  // Level-set check of many regions at once (host arrays of N entries): reg[v]
  // the region of node v (0..nreg-1, -1 = not checked), r[v] its value; each
  // region is cut at ceil(mean r) with the arcs to other regions cut, and
  // cnt_low[k] receives the number of region k's nodes on the low side of the
  // minimal cut (low[v] = 1 for them).  Region k is a level set of its own
  // problem when cnt_low[k] is 0 or its size.
  virtual void pf_check_regions(const int64_t* r, const int32_t* reg, int nreg, int64_t* cnt_low, uint8_t* low) = 0;
  // Replace the base network (host arrays: red[N], alpha[N*k] node-major, lattice order).
  virtual void set_base_network(const int64_t* red, const int64_t* alpha) = 0;
  // Batched lattices on a 2-D canvas (arcs between lattices stay zero).
  virtual void batch_set_layout(const std::vector<BatchRect>& rects) = 0;
  // Energies of lattice i in its own layout (device pointers).
  virtual void batch_add_energies_device(int i, const int64_t* d_unary, const int64_t* d_binary, hipStream_t s) = 0;
  virtual void batch_flows(int64_t* host_flows) = 0;  // raw flow of every lattice of the last run
  // Cuts of every lattice, concatenated in batch node order.
  virtual void batch_cut_device(int32_t* d_out, int which, hipStream_t s) = 0;
  // Invariant checks of the working state on host copies after every push
  // chunk and exact relabel (lf_verify.hpp); a violation fails the solve with
  // LF_ERR_STATE.  Engines without the checks ignore the switch.
  virtual void set_debug_checks(bool on) { (void)on; }
  int64_t flow = 0;
  bool solved = false;
  bool pf_solved = false;
  int64_t pf_levels = 0, pf_sections = 0;
  lf_stats stats{};
};

std::unique_ptr<MinCutEngine> make_mincut_engine(const LatticeGeom& g, int cap_bytes);
// Node-parallel engine for lattices of up to four dimensions (lf_mincut_generic.hip):
// the 4-D kernels.
std::unique_ptr<MinCutEngine> make_generic_engine(const LatticeGeom& g, int cap_bytes);
// Engines of the wider 1-D / 2-D neighbourhoods (radius 2 to 4) that the main
// factory does not build itself (lf_mincut_wide.hip); nullptr when none fits.
std::unique_ptr<MinCutEngine> make_wide_engine(const LatticeGeom& g, int cap_bytes);
// ... of the radius-3 neighbourhoods of 32 and 36 directions (lf_mincut_wide32.hip)
std::unique_ptr<MinCutEngine> make_wide32_engine(const LatticeGeom& g, int cap_bytes);
// ... and of Full2d_48 and Full3d_32 (lf_mincut_wide48.hip).
std::unique_ptr<MinCutEngine> make_wide48_engine(const LatticeGeom& g, int cap_bytes);
// Engines of the radius-1 neighbourhoods (2-D 4/8, 3-D 6/26 directions) whose
// direction table is the canonical lattice order, compiled into the kernels
// (lf_mincut_canon.hip); nullptr for any other kernel or order.
std::unique_ptr<MinCutEngine> make_canon_engine(const LatticeGeom& g, int cap_bytes);

// Learned TV predictor (lf_tvnet.hip): parameter count, workspace size,
// forward pass, loss gradients and one Adam update, all on device pointers.
int tvnet_param_count();
size_t tvnet_work_bytes(int n, int64_t H, int64_t W);
void tvnet_forward(const float* params, int n, int64_t H, int64_t W, const float* f, const float* lam, float* pred,
                   void* work, size_t work_size, hipStream_t s);
// The forward pass with its scratch in a static device array (no allocation).
void tvnet_forward_noalloc(const float* params, int n, int64_t H, int64_t W, const float* f, const float* lam,
                           float* pred, hipStream_t s);
void tvnet_gradients(const float* params, int n, int64_t H, int64_t W, const float* f, const float* lam,
                     const float* target, float* grad, float* loss, void* work, size_t work_size, hipStream_t s);
void tvnet_adam(float* params, const float* grad, float* m, float* v, int step, float lr, hipStream_t s);

}  // namespace lfamd
