// C ABI of the library (declared in include/latticeflow_amd.h).
//
// Stands in for the type-erased reference API: the factory
// getLatticeFlowInterface and name helpers (solver/interface.hpp:171-190,
// generated factory install/templates/factory/file.cpp:83-124) and
// LatticeFlowDirectInterface (solver/interfacing/direct_interface.hpp:11-108).
// Errors are returned as LF_ERR_* codes; the C++ and Python hosts turn them
// back into BadSolverOrKernelName / ValueError / std::out_of_range.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/latticeflow_amd.h"
#include "lf_bk.hpp"
#include "lf_engine.hpp"
#include "lf_kernels.hpp"
#include "lf_tv.hpp"
#include "lf_tvpath.hpp"
#include "lf_verify.hpp"

using namespace lfamd;

namespace {

thread_local std::string g_last_error;

// LatticeEnergyMinimizer / LatticeLevelReductions run on the device; the
// reference's optional BKGCEnergyMinimizer (wscript:21-23,187-190) is the CPU
// cross-check solver of lf_bk.cpp, used only when a caller names it.
const char* const kSolvers[] = {"LatticeEnergyMinimizer", "LatticeLevelReductions", "BKGCEnergyMinimizer"};

bool valid_solver(const std::string& s) {
  for (const char* n : kSolvers)
    if (s == n) return true;
  return false;
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return LF_OK;
  } catch (const LfError& e) {
    g_last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "out of host memory";
    return LF_ERR_DEVICE;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return LF_ERR_DEVICE;
  }
}

int copy_list(const std::vector<std::string>& names, char* buf, size_t buflen) {
  std::string s;
  for (size_t i = 0; i < names.size(); ++i) {
    if (i) s += '\n';
    s += names[i];
  }
  if (buf && buflen) {
    size_t n = std::min(buflen - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return (int)s.size();
}

}  // namespace

struct lf_handle {
  std::string solver, kernel;
  int ndim = 0;
  std::vector<int64_t> shape;
  int64_t n = 0;
  int k = 0;
  int cap_bytes = 8;
  std::vector<int> offs;      // k * ndim, lattice order
  std::vector<double> geocut; // k
  LatticeGeom geom;
  std::unique_ptr<MinCutEngine> eng;
  std::unique_ptr<BkSolver> bk;  // BKGCEnergyMinimizer (CPU); eng is then null
  bool levels = false;  // LatticeLevelReductions (parametric divide and conquer)
  bool batch = false;   // lattices packed on one canvas
  std::vector<BatchRect> rects;
  std::vector<std::vector<int>> batch_offs;  // per lattice: k/2 * 2 edge deltas
  int64_t canvas[2] = {0, 0};
  lf_tuning tuning{};
  double run_time = 0;
};

extern "C" {

const char* lf_last_error(void) { return g_last_error.c_str(); }

int lf_is_valid_solver(const char* name) { return name && valid_solver(name) ? 1 : 0; }

int lf_is_valid_kernel(const char* name) { return name && find_kernel(name) ? 1 : 0; }

int lf_kernel_dimension(const char* kernel) {
  const KernelDef* k = kernel ? find_kernel(kernel) : nullptr;
  if (!k) {
    g_last_error = std::string("Bad Kernel or Solver Name: ") + (kernel ? kernel : "(null)");
    return LF_ERR_BAD_NAME;
  }
  return k->ndim;
}

int lf_valid_kernel_names(char* buf, size_t buflen) {
  std::vector<std::string> names;
  for (const auto& kv : kernel_registry()) names.push_back(kv.first);
  return copy_list(names, buf, buflen);
}

int lf_valid_solver_names(char* buf, size_t buflen) {
  std::vector<std::string> names(std::begin(kSolvers), std::end(kSolvers));
  return copy_list(names, buf, buflen);
}

int lf_kernel_size(const char* kernel) {
  const KernelDef* k = kernel ? find_kernel(kernel) : nullptr;
  if (!k) {
    g_last_error = std::string("Bad Kernel or Solver Name: ") + (kernel ? kernel : "(null)");
    return LF_ERR_BAD_NAME;
  }
  return (int)k->edges.size();
}

int lf_kernel_offsets(const char* kernel, int32_t* offsets, double* w) {
  return guarded([&] {
    const KernelDef* k = kernel ? find_kernel(kernel) : nullptr;
    if (!k) fail(LF_ERR_BAD_NAME, std::string("Bad Kernel or Solver Name: ") + (kernel ? kernel : "(null)"));
    for (size_t i = 0; i < k->edges.size(); ++i) {
      for (int d = 0; d < k->ndim; ++d)
        if (offsets) offsets[i * k->ndim + d] = k->edges[i][d];
      if (w) w[i] = k->geocut_w[i];
    }
  });
}

int lf_create(lf_handle** out, const char* solver, const char* kernel, int ndim, const int64_t* shape,
              int cap_bytes) {
  if (out) *out = nullptr;
  return guarded([&] {
    if (!out || !solver || !kernel || !shape) fail(LF_ERR_BAD_ARG, "null argument");
    if (!valid_solver(solver)) fail(LF_ERR_BAD_NAME, std::string("Bad Kernel or Solver Name: ") + solver);
    const KernelDef* kd = find_kernel(kernel);
    if (!kd) fail(LF_ERR_BAD_NAME, std::string("Bad Kernel or Solver Name: ") + kernel);
    if (ndim != kd->ndim)
      fail(LF_ERR_BAD_ARG, "Kernel '" + std::string(kernel) + "' is " + std::to_string(kd->ndim) +
                               "-dimensional; a shape of " + std::to_string(ndim) + " dimensions was provided.");
    if (cap_bytes != 4 && cap_bytes != 8) fail(LF_ERR_BAD_ARG, "cap_bytes must be 4 or 8");
    std::unique_ptr<lf_handle> h(new lf_handle);
    h->solver = solver;
    h->kernel = kernel;
    h->ndim = ndim;
    h->shape.assign(shape, shape + ndim);
    h->n = 1;
    for (int d = 0; d < ndim; ++d) {
      if (shape[d] <= 0) fail(LF_ERR_BAD_ARG, "every dimension must be positive");
      if (shape[d] > (1LL << 30)) fail(LF_ERR_BAD_ARG, "dimension too large");
      h->n *= shape[d];
    }
    h->k = (int)kd->edges.size();
    h->cap_bytes = cap_bytes;
    if (!lattice_edge_order(*kd, h->shape, h->offs, h->geocut))
      fail(LF_ERR_BAD_ARG, "kernel is not symmetric on this lattice");
    LatticeGeom& g = h->geom;
    g.ndim = ndim;
    g.k = h->k;
    g.dims = h->shape;
    g.offsets = h->offs;
    g.D = ndim == 3 ? shape[0] : 1;
    g.H = ndim == 3 ? shape[1] : (ndim == 2 ? shape[0] : 1);
    g.W = shape[ndim - 1];
    for (int i = 0; i < h->k; ++i) {
      const int* o = &h->offs[i * ndim];
      g.dz.push_back(ndim == 3 ? o[0] : 0);
      g.dy.push_back(ndim == 3 ? o[1] : (ndim == 2 ? o[0] : 0));
      g.dx.push_back(o[ndim - 1]);
    }
    if (h->solver == "BKGCEnergyMinimizer") {
      h->bk.reset(new BkSolver(ndim, h->shape, h->k, h->offs));
      *out = h.release();
      return;
    }
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
      fail(LF_ERR_DEVICE, "no HIP device available");
    h->eng = make_mincut_engine(g, cap_bytes);
    h->levels = h->solver == "LatticeLevelReductions";
    *out = h.release();
  });
}

int lf_destroy(lf_handle* h) {
  return guarded([&] { delete h; });
}

int lf_n_dimensions(const lf_handle* h) { return h ? h->ndim : LF_ERR_BAD_ARG; }
int64_t lf_n_nodes(const lf_handle* h) { return h ? h->n : LF_ERR_BAD_ARG; }
int lf_n_filling_edges(const lf_handle* h) { return h ? h->k / 2 : LF_ERR_BAD_ARG; }

int lf_edge_deltas(const lf_handle* h, int32_t* ed) {
  return guarded([&] {
    if (!h || !ed) fail(LF_ERR_BAD_ARG, "null argument");
    if (h->batch) fail(LF_ERR_STATE, "a batch handle has one edge order per lattice (lf_batch_edge_deltas)");
    for (int i = 0; i < h->k / 2 * h->ndim; ++i) ed[i] = h->offs[i];
  });
}

int lf_directions(const lf_handle* h, int32_t* offs) {
  return guarded([&] {
    if (!h || !offs) fail(LF_ERR_BAD_ARG, "null argument");
    for (int i = 0; i < h->k * h->ndim; ++i) offs[i] = h->offs[i];
  });
}

int lf_add_energy_potentials_device(lf_handle* h, const int64_t* du, const int64_t* db, void* stream) {
  return guarded([&] {
    if (!h || !du || !db) fail(LF_ERR_BAD_ARG, "null argument");
    if (h->batch) fail(LF_ERR_STATE, "a batch handle takes energies per lattice (lf_batch_add_energy_potentials)");
    if (h->bk) {
      // the CPU solver reads host copies
      std::vector<int64_t> u((size_t)h->n * 2), b((size_t)h->n * (h->k / 2) * 4);
      hipStream_t s = reinterpret_cast<hipStream_t>(stream);
      LF_HIP(hipMemcpyAsync(u.data(), du, u.size() * sizeof(int64_t), hipMemcpyDeviceToHost, s));
      LF_HIP(hipMemcpyAsync(b.data(), db, b.size() * sizeof(int64_t), hipMemcpyDeviceToHost, s));
      LF_HIP(hipStreamSynchronize(s));
      h->bk->set_energies(u.data(), b.data());
      return;
    }
    h->eng->add_energies_device(du, db, reinterpret_cast<hipStream_t>(stream));
  });
}

int lf_add_energy_potentials(lf_handle* h, const int64_t* unary, const int64_t* binary) {
  return guarded([&] {
    if (!h || !unary || !binary) fail(LF_ERR_BAD_ARG, "null argument");
    if (h->batch) fail(LF_ERR_STATE, "a batch handle takes energies per lattice (lf_batch_add_energy_potentials)");
    if (h->bk) {
      h->bk->set_energies(unary, binary);
      return;
    }
    const size_t nu = (size_t)h->n * 2, nb = (size_t)h->n * (h->k / 2) * 4;
    int64_t *du = nullptr, *db = nullptr;
    hipStream_t s = h->eng->stream();
    LF_HIP(hipMallocAsync((void**)&du, nu * sizeof(int64_t), s));
    LF_HIP(hipMallocAsync((void**)&db, nb * sizeof(int64_t), s));
    struct Free {
      int64_t *a, *b;
      hipStream_t s;
      ~Free() {
        hipFreeAsync(a, s);
        hipFreeAsync(b, s);
        hipStreamSynchronize(s);
      }
    } fr{du, db, s};
    LF_HIP(hipMemcpyAsync(du, unary, nu * sizeof(int64_t), hipMemcpyHostToDevice, s));
    LF_HIP(hipMemcpyAsync(db, binary, nb * sizeof(int64_t), hipMemcpyHostToDevice, s));
    h->eng->add_energies_device(du, db, s);
  });
}

int lf_reset_energies(lf_handle* h) {
  return guarded([&] {
    if (!h) fail(LF_ERR_BAD_ARG, "null argument");
    if (h->bk) {
      std::vector<int64_t> u((size_t)h->n * 2, 0), b((size_t)h->n * (h->k / 2) * 4, 0);
      h->bk->set_energies(u.data(), b.data());
      return;
    }
    h->eng->reset_energies();
  });
}

int lf_set_tuning(lf_handle* h, const lf_tuning* t) {
  return guarded([&] {
    if (!h) fail(LF_ERR_BAD_ARG, "null argument");
    h->tuning = t ? *t : lf_tuning{};
  });
}

int lf_set_debug_checks(lf_handle* h, int on) {
  return guarded([&] {
    if (!h) fail(LF_ERR_BAD_ARG, "null argument");
    if (h->eng) h->eng->set_debug_checks(on != 0);
  });
}

int lf_verify_state(const int64_t* dhw, int k, const int32_t* dirs, const int64_t* red0, const int64_t* alpha0,
                    int cap_bytes, const void* res, const void* ex, const void* mbox, int mbox_planes,
                    const void* mask, int mask_bytes, const int32_t* h, int labels_exact) {
  return guarded([&] {
    if (!dhw || !dirs || !red0 || !alpha0 || !res || !ex || k <= 0 || k % 2 || k > 64)
      fail(LF_ERR_BAD_ARG, "bad state arrays");
    if (cap_bytes != 4 && cap_bytes != 8) fail(LF_ERR_BAD_ARG, "cap_bytes must be 4 or 8");
    if (mask && mask_bytes != 1 && mask_bytes != 2 && mask_bytes != 4 && mask_bytes != 8)
      fail(LF_ERR_BAD_ARG, "mask_bytes must be 1, 2, 4 or 8");
    if ((mbox_planes != 0 && !mbox) || mbox_planes < 0) fail(LF_ERR_BAD_ARG, "bad mailbox planes");
    std::vector<int> dz(k), dy(k), dx(k);
    for (int i = 0; i < k; ++i) {
      dz[i] = dirs[3 * i];
      dy[i] = dirs[3 * i + 1];
      dx[i] = dirs[3 * i + 2];
    }
    VerifyView v;
    v.D = (int)dhw[0]; v.H = (int)dhw[1]; v.W = (int)dhw[2]; v.K = k;
    v.dz = dz.data(); v.dy = dy.data(); v.dx = dx.data();
    v.red0 = red0; v.alpha0 = alpha0; v.cap_bytes = cap_bytes;
    v.res = res; v.ex = ex; v.mbox = mbox; v.mbox_planes = mbox_planes;
    v.mask = mask; v.mask_bytes = mask_bytes; v.h = h; v.labels_exact = labels_exact != 0;
    const std::string err = verify_state(v);
    if (!err.empty()) fail(LF_ERR_STATE, err);
  });
}

int lf_get_tuning(const lf_handle* h, lf_tuning* t) {
  return guarded([&] {
    if (!h || !t) fail(LF_ERR_BAD_ARG, "null argument");
    *t = h->tuning;
  });
}

int lf_autotune(lf_handle* h, int reps, lf_tuning* best) {
  return guarded([&] {
    if (!h) fail(LF_ERR_BAD_ARG, "null argument");
    if (h->bk) {  // the CPU solver has nothing to tune
      if (best) *best = h->tuning;
      return;
    }
    if (reps < 1) reps = 1;
    // one untimed warm-up, then the mean of `reps` timed solves (the
    // schedule of a push-relabel run is not deterministic, so a candidate
    // that is only sometimes fast must not win on its best run)
    auto timed = [&](const lf_tuning& t) {
      if (h->levels) h->eng->pf_run(t);
      else h->eng->run(t);
      double tot = 0;
      for (int r = 0; r < reps; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        if (h->levels) h->eng->pf_run(t);
        else h->eng->run(t);
        tot += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      }
      return tot;
    };
    // stage 1: the adaptive schedule and seven fixed ones, on every tile shape
    // (one round of one iteration per relabel suits networks that are nearly
    // solved at the start, 12 rounds of 4 small lattices whose rounds are
    // latency)
    const lf_tuning sched[] = {{0, 0, 0, 0, 0}, {4, 4, 0, 0, 0}, {4, 8, 0, 0, 0},  {6, 4, 0, 0, 0},
                               {4, 2, 0, 0, 0}, {8, 2, 0, 0, 0}, {1, 1, 0, 0, 0}, {12, 4, 0, 0, 0}};
    lf_tuning best_t = h->tuning;
    double best_s = 1e300;
    for (int shape = 1; shape <= 3; ++shape)
      for (const lf_tuning& c : sched) {
        lf_tuning t = c;
        t.tile_shape = shape;
        const double s = timed(t);
        if (s < best_s) {
          best_s = s;
          best_t = t;
        }
      }
    // stage 2: how far the relabels between push rounds reach
    const lf_tuning base = best_t;
    for (int cap : {2, 4, 8, 16}) {
      lf_tuning t = base;
      t.relabel_launches = cap;
      const double s = timed(t);
      if (s < best_s) {
        best_s = s;
        best_t = t;
      }
    }
    h->tuning = best_t;
    if (best) *best = best_t;
  });
}

int lf_run(lf_handle* h) {
  return guarded([&] {
    if (!h) fail(LF_ERR_BAD_ARG, "null argument");
    const auto t0 = std::chrono::steady_clock::now();
    if (h->bk) h->bk->run();
    else if (h->levels) h->eng->pf_run(h->tuning);
    else h->eng->run(h->tuning);
    h->run_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

int lf_run_time(const lf_handle* h, double* seconds) {
  return guarded([&] {
    if (!h || !seconds) fail(LF_ERR_BAD_ARG, "null argument");
    *seconds = h->run_time;
  });
}

int lf_get_stats(const lf_handle* h, lf_stats* st) {
  return guarded([&] {
    if (!h || !st) fail(LF_ERR_BAD_ARG, "null argument");
    if (h->bk) {
      *st = lf_stats{};
      st->cycles = h->bk->augmentations();
      st->seconds = h->run_time;
      return;
    }
    *st = h->eng->stats;
  });
}

int lf_get_cut_device(const lf_handle* h, int32_t* d_out, int which, void* stream) {
  return guarded([&] {
    if (!h || !d_out) fail(LF_ERR_BAD_ARG, "null argument");
    if (h->bk) {
      if (which != LF_CUT_SMIN && which != LF_CUT_SMAX) fail(LF_ERR_BAD_ARG, "bad cut selector");
      std::vector<int32_t> c((size_t)h->n);
      h->bk->cut(c.data(), which);
      hipStream_t s = reinterpret_cast<hipStream_t>(stream);
      LF_HIP(hipMemcpyAsync(d_out, c.data(), c.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
      LF_HIP(hipStreamSynchronize(s));
      return;
    }
    if (h->levels) {
      if (!h->eng->pf_solved) fail(LF_ERR_STATE, "run() has not been called");
      h->eng->pf_cut_device(d_out, reinterpret_cast<hipStream_t>(stream));
      return;
    }
    if (!h->eng->solved) fail(LF_ERR_STATE, "run() has not been called");
    if (which != LF_CUT_SMIN && which != LF_CUT_SMAX) fail(LF_ERR_BAD_ARG, "bad cut selector");
    if (h->batch) h->eng->batch_cut_device(d_out, which, reinterpret_cast<hipStream_t>(stream));
    else h->eng->cut_device(d_out, which, reinterpret_cast<hipStream_t>(stream));
  });
}

int lf_get_cut(const lf_handle* h, int32_t* out, int which) {
  return guarded([&] {
    if (!h || !out) fail(LF_ERR_BAD_ARG, "null argument");
    if (h->bk) {
      if (which != LF_CUT_SMIN && which != LF_CUT_SMAX) fail(LF_ERR_BAD_ARG, "bad cut selector");
      h->bk->cut(out, which);
      return;
    }
    hipStream_t s = h->eng->stream();
    int32_t* d = nullptr;
    LF_HIP(hipMallocAsync((void**)&d, sizeof(int32_t) * h->n, s));
    struct Free {
      int32_t* d;
      hipStream_t s;
      ~Free() {
        hipFreeAsync(d, s);
        hipStreamSynchronize(s);
      }
    } fr{d, s};
    if (h->levels) {
      if (!h->eng->pf_solved) fail(LF_ERR_STATE, "run() has not been called");
      h->eng->pf_cut_device(d, s);
    } else {
      if (!h->eng->solved) fail(LF_ERR_STATE, "run() has not been called");
      if (which != LF_CUT_SMIN && which != LF_CUT_SMAX) fail(LF_ERR_BAD_ARG, "bad cut selector");
      if (h->batch) h->eng->batch_cut_device(d, which, s);
      else h->eng->cut_device(d, which, s);
    }
    LF_HIP(hipMemcpyAsync(out, d, sizeof(int32_t) * h->n, hipMemcpyDeviceToHost, s));
    LF_HIP(hipStreamSynchronize(s));
  });
}

int lf_get_levels(const lf_handle* h, int64_t* r_out) {
  return guarded([&] {
    if (!h || !r_out) fail(LF_ERR_BAD_ARG, "null argument");
    if (h->bk || !h->levels) fail(LF_ERR_STATE, "level values exist only for LatticeLevelReductions");
    if (!h->eng->pf_solved) fail(LF_ERR_STATE, "run() has not been called");
    hipStream_t s = h->eng->stream();
    int64_t* d = nullptr;
    LF_HIP(hipMallocAsync((void**)&d, sizeof(int64_t) * h->n, s));
    struct Free {
      int64_t* d;
      hipStream_t s;
      ~Free() {
        hipFreeAsync(d, s);
        hipStreamSynchronize(s);
      }
    } fr{d, s};
    h->eng->pf_values_device(d, s);
    LF_HIP(hipMemcpyAsync(r_out, d, sizeof(int64_t) * h->n, hipMemcpyDeviceToHost, s));
    LF_HIP(hipStreamSynchronize(s));
  });
}

int lf_get_flow(const lf_handle* h, int64_t* flow) {
  return guarded([&] {
    if (!h || !flow) fail(LF_ERR_BAD_ARG, "null argument");
    if (h->bk) {
      if (!h->bk->solved()) fail(LF_ERR_STATE, "run() has not been called");
      *flow = h->bk->flow();
      return;
    }
    if (!h->eng->solved) fail(LF_ERR_STATE, "run() has not been called");
    *flow = h->eng->flow;
  });
}

// ---- batched lattices --------------------------------------------------------

// Shelf packing: lattices sorted by height (tallest first) fill rows of a
// canvas whose width is about the square root of the total area (at least the
// widest lattice), left to right.
static void pack_batch(lf_handle* h, const int64_t* shapes, int n) {
  std::vector<int> idx(n);
  double area = 0;
  int64_t wmax = 1;
  for (int i = 0; i < n; ++i) {
    idx[i] = i;
    area += (double)shapes[2 * i] * (double)shapes[2 * i + 1];
    wmax = std::max(wmax, shapes[2 * i + 1]);
  }
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    if (shapes[2 * a] != shapes[2 * b]) return shapes[2 * a] > shapes[2 * b];
    return shapes[2 * a + 1] > shapes[2 * b + 1];
  });
  const int64_t Wc = std::max<int64_t>(wmax, ((int64_t)std::ceil(std::sqrt(area)) + 63) / 64 * 64);
  int64_t y = 0, x = 0, shelf = 0;
  h->rects.assign(n, BatchRect{});
  for (int i : idx) {
    const int64_t hh = shapes[2 * i], ww = shapes[2 * i + 1];
    if (x + ww > Wc) {
      y += shelf;
      x = 0;
      shelf = 0;
    }
    BatchRect& R = h->rects[i];
    R.y0 = y;
    R.x0 = x;
    R.h = hh;
    R.w = ww;
    shelf = std::max(shelf, hh);
    x += ww;
  }
  int64_t off = 0;
  for (int i = 0; i < n; ++i) {
    h->rects[i].off = off;
    off += shapes[2 * i] * shapes[2 * i + 1];
  }
  h->canvas[0] = y + shelf;
  h->canvas[1] = Wc;
}

int lf_batch_create(lf_handle** out, const char* kernel, int n, const int64_t* shapes, int cap_bytes) {
  if (out) *out = nullptr;
  return guarded([&] {
    if (!out || !kernel || !shapes) fail(LF_ERR_BAD_ARG, "null argument");
    const KernelDef* kd = find_kernel(kernel);
    if (!kd) fail(LF_ERR_BAD_NAME, std::string("Bad Kernel or Solver Name: ") + kernel);
    if (kd->ndim != 2) fail(LF_ERR_UNSUPPORTED, "batches take 2-D kernels");
    if (n < 1) fail(LF_ERR_BAD_ARG, "a batch needs at least one lattice");
    if (cap_bytes != 4 && cap_bytes != 8) fail(LF_ERR_BAD_ARG, "cap_bytes must be 4 or 8");
    std::unique_ptr<lf_handle> h(new lf_handle);
    h->solver = "LatticeEnergyMinimizer";
    h->kernel = kernel;
    h->ndim = 2;
    h->k = (int)kd->edges.size();
    h->cap_bytes = cap_bytes;
    h->batch = true;
    if (h->k / 2 > 16) fail(LF_ERR_UNSUPPORTED, "kernel too large for a batch");
    h->n = 0;
    for (int i = 0; i < n; ++i) {
      if (shapes[2 * i] <= 0 || shapes[2 * i + 1] <= 0) fail(LF_ERR_BAD_ARG, "every dimension must be positive");
      if (shapes[2 * i] > (1LL << 20) || shapes[2 * i + 1] > (1LL << 20)) fail(LF_ERR_BAD_ARG, "dimension too large");
      h->n += shapes[2 * i] * shapes[2 * i + 1];
    }
    pack_batch(h.get(), shapes, n);
    h->shape = {h->canvas[0], h->canvas[1]};
    if (!lattice_edge_order(*kd, h->shape, h->offs, h->geocut))
      fail(LF_ERR_BAD_ARG, "kernel is not symmetric on this lattice");
    const int kp = h->k / 2;
    h->batch_offs.resize(n);
    for (int i = 0; i < n; ++i) {
      std::vector<int> o;
      std::vector<double> w;
      if (!lattice_edge_order(*kd, {shapes[2 * i], shapes[2 * i + 1]}, o, w))
        fail(LF_ERR_BAD_ARG, "kernel is not symmetric on this lattice");
      h->batch_offs[i].assign(o.begin(), o.begin() + 2 * kp);
      for (int e = 0; e < kp; ++e) {
        int found = -1;
        for (int f = 0; f < kp; ++f)
          if (o[2 * f] == h->offs[2 * e] && o[2 * f + 1] == h->offs[2 * e + 1]) found = f;
        if (found < 0) fail(LF_ERR_UNSUPPORTED, "lattice with a different set of positive directions");
        h->rects[i].perm[e] = (int8_t)found;
      }
    }
    LatticeGeom& g = h->geom;
    g.ndim = 2;
    g.k = h->k;
    g.D = 1;
    g.H = h->canvas[0];
    g.W = h->canvas[1];
    for (int i = 0; i < h->k; ++i) {
      g.dz.push_back(0);
      g.dy.push_back(h->offs[i * 2]);
      g.dx.push_back(h->offs[i * 2 + 1]);
    }
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
      fail(LF_ERR_DEVICE, "no HIP device available");
    h->eng = make_mincut_engine(g, cap_bytes);
    h->eng->batch_set_layout(h->rects);
    *out = h.release();
  });
}

int lf_batch_size(const lf_handle* h) {
  if (!h) return LF_ERR_BAD_ARG;
  if (!h->batch) {
    g_last_error = "not a batch handle";
    return LF_ERR_STATE;
  }
  return (int)h->rects.size();
}

int lf_batch_canvas(const lf_handle* h, int64_t* rc) {
  return guarded([&] {
    if (!h || !rc) fail(LF_ERR_BAD_ARG, "null argument");
    if (!h->batch) fail(LF_ERR_STATE, "not a batch handle");
    rc[0] = h->canvas[0];
    rc[1] = h->canvas[1];
  });
}

int lf_batch_edge_deltas(const lf_handle* h, int i, int32_t* ed) {
  return guarded([&] {
    if (!h || !ed) fail(LF_ERR_BAD_ARG, "null argument");
    if (!h->batch) fail(LF_ERR_STATE, "not a batch handle");
    if (i < 0 || i >= (int)h->rects.size()) fail(LF_ERR_OUT_OF_RANGE, "batch lattice index out of range");
    for (size_t j = 0; j < h->batch_offs[i].size(); ++j) ed[j] = h->batch_offs[i][j];
  });
}

int lf_batch_add_energy_potentials_device(lf_handle* h, int i, const int64_t* du, const int64_t* db, void* stream) {
  return guarded([&] {
    if (!h || !du || !db) fail(LF_ERR_BAD_ARG, "null argument");
    if (!h->batch) fail(LF_ERR_STATE, "not a batch handle");
    h->eng->batch_add_energies_device(i, du, db, reinterpret_cast<hipStream_t>(stream));
  });
}

int lf_batch_add_energy_potentials(lf_handle* h, int i, const int64_t* unary, const int64_t* binary) {
  return guarded([&] {
    if (!h || !unary || !binary) fail(LF_ERR_BAD_ARG, "null argument");
    if (!h->batch) fail(LF_ERR_STATE, "not a batch handle");
    if (i < 0 || i >= (int)h->rects.size()) fail(LF_ERR_OUT_OF_RANGE, "batch lattice index out of range");
    const int64_t m = h->rects[i].h * h->rects[i].w;
    const size_t nu = (size_t)m * 2, nb = (size_t)m * (h->k / 2) * 4;
    int64_t *du = nullptr, *db = nullptr;
    hipStream_t s = h->eng->stream();
    LF_HIP(hipMallocAsync((void**)&du, nu * sizeof(int64_t), s));
    LF_HIP(hipMallocAsync((void**)&db, nb * sizeof(int64_t), s));
    struct Free {
      int64_t *a, *b;
      hipStream_t s;
      ~Free() {
        hipFreeAsync(a, s);
        hipFreeAsync(b, s);
        hipStreamSynchronize(s);
      }
    } fr{du, db, s};
    LF_HIP(hipMemcpyAsync(du, unary, nu * sizeof(int64_t), hipMemcpyHostToDevice, s));
    LF_HIP(hipMemcpyAsync(db, binary, nb * sizeof(int64_t), hipMemcpyHostToDevice, s));
    h->eng->batch_add_energies_device(i, du, db, s);
  });
}

int lf_batch_get_flows(const lf_handle* h, int64_t* flows) {
  return guarded([&] {
    if (!h || !flows) fail(LF_ERR_BAD_ARG, "null argument");
    if (!h->batch) fail(LF_ERR_STATE, "not a batch handle");
    if (!h->eng->solved) fail(LF_ERR_STATE, "run() has not been called");
    h->eng->batch_flows(flows);
  });
}

// Exact TV of one image at several lambdas: the lambdas are the layers of one
// stacked lattice (no arcs between layers), solved by one device divide and
// conquer.  r_out[l*n + i] receives the final raw r of pixel i at lambda l.
// Guidance for tv_solve_layers: predictions of the output per solved layer.
struct TvGuide {
  const float* d_pred = nullptr;  // layers * nx * ny floats on the device
  float delta = 0;
  int max_repairs = 0;
  lf_guided_stats* stats = nullptr;
};

static void tv_solve_layers(const KernelDef& kd, int64_t nx, int64_t ny, const TvBase& b,
                            const std::vector<double>& lams, const lf_tuning& tun, std::vector<int64_t>& r_out,
                            lf_stats* stats, const TvGuide* guide = nullptr) {
  const int layers = (int)lams.size();
  std::vector<int> offs;
  std::vector<double> w;
  lattice_edge_order(kd, {nx, ny}, offs, w);
  const int k = (int)kd.edges.size();
  LatticeGeom g;
  g.ndim = 2;
  g.k = k;
  g.D = layers;
  g.H = nx;
  g.W = ny;
  for (int i = 0; i < k; ++i) {
    g.dz.push_back(0);
    g.dy.push_back(offs[i * 2]);
    g.dx.push_back(offs[i * 2 + 1]);
  }
  std::vector<int64_t> pwf((size_t)layers * (k / 2));
  for (int l = 0; l < layers; ++l)
    for (int e = 0; e < k / 2; ++e) {
      const int64_t p = tv_pwf(b, lams[l], w[e]);
      if (p > ((int64_t)1 << 56) / k)
        fail(LF_ERR_OVERFLOW, "lambda too small for the 64-bit fixed point of the TV network");
      pwf[(size_t)l * (k / 2) + e] = p;
    }
  std::unique_ptr<MinCutEngine> eng = make_mincut_engine(g, 8);
  eng->set_tv_network(b.uf.data(), pwf.data(), layers);
  if (guide && b.S > 0)
    // predicted output values x map to r units as the function values do:
    // r = 2^41 ((x - M) / S + 0.5)
    eng->pf_run_layers_guided(tun, layers, guide->d_pred, guide->delta, guide->max_repairs,
                              std::ldexp(1.0, kTvFunctionBits + 1) / b.S,
                              std::ldexp(1.0, kTvFunctionBits + 1) * (0.5 - b.M / b.S), guide->stats);
  else
    eng->pf_run_layers(tun, layers);
  r_out.resize((size_t)layers * nx * ny);
  hipStream_t s = eng->stream();
  int64_t* d = nullptr;
  LF_HIP(hipMallocAsync((void**)&d, sizeof(int64_t) * r_out.size(), s));
  eng->pf_values_device(d, s);
  LF_HIP(hipMemcpyAsync(r_out.data(), d, sizeof(int64_t) * r_out.size(), hipMemcpyDeviceToHost, s));
  LF_HIP(hipFreeAsync(d, s));
  LF_HIP(hipStreamSynchronize(s));
  if (stats) *stats = eng->stats;
}

static const KernelDef& tv_kernel(const char* kernel) {
  if (!kernel) fail(LF_ERR_BAD_ARG, "null argument");
  const KernelDef* kd = find_kernel(kernel);
  if (!kd) fail(LF_ERR_BAD_NAME, std::string("Bad Kernel or Solver Name: ") + kernel);
  if (kd->ndim != 2 || !kd->geocut)
    fail(LF_ERR_BAD_ARG, "Kernel is not valid for GeoCuts or TV Minimization (2-D geocut kernels only).");
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) fail(LF_ERR_DEVICE, "no HIP device available");
  return *kd;
}

int lf_tv_single(const char* kernel, int64_t nx, int64_t ny, const double* f, double lambda, double* R,
                 lf_stats* stats) {
  return guarded([&] {
    const KernelDef& kd = tv_kernel(kernel);
    if (!f || !R) fail(LF_ERR_BAD_ARG, "null argument");
    if (nx <= 0 || ny <= 0) fail(LF_ERR_BAD_ARG, "empty image");
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t n = nx * ny;
    TvBase b = tv_base(n, f);
    std::vector<int64_t> r;
    tv_solve_layers(kd, nx, ny, b, {lambda}, lf_tuning{}, r, stats);
    tv_output_single(b, r.data(), n, R);
    if (stats) stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

static void tv_path_impl(const char* kernel, int64_t nx, int64_t ny, const double* f, int64_t n_lambda,
                         const double* lambdas, double* R, lf_stats* stats, const float* d_pred, double delta,
                         int max_repairs, lf_guided_stats* gstats) {
  const KernelDef& kd = tv_kernel(kernel);
  if (!f || !R || (n_lambda > 0 && !lambdas)) fail(LF_ERR_BAD_ARG, "null argument");
  if (nx <= 0 || ny <= 0) fail(LF_ERR_BAD_ARG, "empty image");
  if (gstats) *gstats = lf_guided_stats{};
  if (n_lambda <= 0) return;  // the reference returns an empty path (tv_solver.hpp:664-667)
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t n = nx * ny;
  // negative values count as 0 (tv_solver.hpp:669-674)
  std::vector<double> lam(lambdas, lambdas + n_lambda);
  for (double& l : lam) l = std::max(l, 0.0);
  const double lmax = std::max(*std::max_element(lam.begin(), lam.end()), 1e-8);
  TvBase b = tv_base(n, f);
  // each positive lambda is an exact single-lambda problem; lambda = 0 is the
  // fully merged solution (constant M)
  std::vector<double> solve;
  std::vector<int64_t> which(n_lambda, -1);
  for (int64_t l = 0; l < n_lambda; ++l)
    if (lam[l] > 0) {
      which[l] = (int64_t)solve.size();
      solve.push_back(lam[l]);
    }
  std::vector<int64_t> r;
  if (!solve.empty()) {
    if (d_pred) {
      // predictions of the solved lambdas only, in solve order
      float* dp = nullptr;
      hipStream_t s = nullptr;
      const bool all = (int64_t)solve.size() == n_lambda;
      if (!all) {
        LF_HIP(hipMalloc((void**)&dp, sizeof(float) * n * solve.size()));
        for (int64_t l = 0; l < n_lambda; ++l)
          if (which[l] >= 0)
            LF_HIP(hipMemcpyAsync(dp + which[l] * n, d_pred + l * n, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
        LF_HIP(hipStreamSynchronize(s));
      }
      struct Free {
        float* p;
        ~Free() { hipFree(p); }
      } fr{dp};
      TvGuide gd;
      gd.d_pred = all ? d_pred : dp;
      gd.delta = (float)delta;
      gd.max_repairs = max_repairs;
      gd.stats = gstats;
      tv_solve_layers(kd, nx, ny, b, solve, lf_tuning{}, r, stats, &gd);
    } else {
      tv_solve_layers(kd, nx, ny, b, solve, lf_tuning{}, r, stats);
    }
  }
  std::vector<double> rp((size_t)n);
  for (int64_t l = 0; l < n_lambda; ++l) {
    double* out = R + l * n;
    if (which[l] < 0) {
      for (int64_t i = 0; i < n; ++i) out[i] = b.M;
      continue;
    }
    // path value at s = lambda / lambda_max: r_path(s) = s r(lambda) by
    // homogeneity of the TV problem (TVRegPathSegment::get_r_AtLambda,
    // tv_regpath_node.hpp:146-150, is linear in s over each segment)
    const double sc = lam[l] / lmax;
    const int64_t* rl = r.data() + which[l] * n;
    for (int64_t i = 0; i < n; ++i) rp[i] = sc * (double)rl[i];
    tv_output_path(b, rp.data(), n, out);
  }
  if (stats) stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int lf_tv_path(const char* kernel, int64_t nx, int64_t ny, const double* f, int64_t n_lambda,
               const double* lambdas, double* R, lf_stats* stats) {
  return guarded([&] { tv_path_impl(kernel, nx, ny, f, n_lambda, lambdas, R, stats, nullptr, 0, 0, nullptr); });
}

int lf_tv_path_guided(const char* kernel, int64_t nx, int64_t ny, const double* f, int64_t n_lambda,
                      const double* lambdas, const float* d_pred, double delta, int max_repairs, double* R,
                      lf_stats* stats, lf_guided_stats* gstats, void* stream) {
  return guarded([&] {
    if (!d_pred) fail(LF_ERR_BAD_ARG, "null prediction");
    // the predictions are complete before any stream of this call reads them
    LF_HIP(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
    if (!(delta >= 0)) fail(LF_ERR_BAD_ARG, "delta must be >= 0");
    if (max_repairs < 0) fail(LF_ERR_BAD_ARG, "max_repairs must be >= 0");
    tv_path_impl(kernel, nx, ny, f, n_lambda, lambdas, R, stats, d_pred, delta, max_repairs, gstats);
  });
}

// ---- event-driven regularization path --------------------------------------------

struct lf_tvpath {
  std::unique_ptr<lfamd::TvPath> path;
};

int lf_tvpath_create(lf_tvpath** out, const char* kernel, int64_t nx, int64_t ny, const double* f,
                     double lambda_max, int64_t host_max) {
  return guarded([&] {
    if (!out) fail(LF_ERR_BAD_ARG, "null argument");
    *out = nullptr;
    std::unique_ptr<lf_tvpath> p(new lf_tvpath);
    p->path.reset(new TvPath(kernel, nx, ny, f, lambda_max, host_max));
    *out = p.release();
  });
}

int lf_tvpath_eval(const lf_tvpath* p, int64_t n_lambda, const double* lambdas, double* R) {
  return guarded([&] {
    if (!p) fail(LF_ERR_BAD_ARG, "null handle");
    p->path->eval(n_lambda, lambdas, R);
  });
}

int lf_tvpath_get_stats(const lf_tvpath* p, lf_tvpath_stats* out) {
  return guarded([&] {
    if (!p || !out) fail(LF_ERR_BAD_ARG, "null argument");
    const TvPathStats& s = p->path->stats();
    out->initial_regions = s.initial_regions;
    out->segments = s.segments;
    out->joins = s.joins;
    out->splits = s.splits;
    out->checks = s.checks;
    out->host_checks = s.host_checks;
    out->device_calls = s.device_calls;
    out->dinkelbach_steps = s.dinkelbach_steps;
    out->initial_seconds = s.initial_seconds;
    out->device_check_seconds = s.device_check_seconds;
    out->host_check_seconds = s.host_check_seconds;
    out->seconds = s.seconds;
  });
}

int lf_tvpath_destroy(lf_tvpath* p) {
  return guarded([&] { delete p; });
}

// ---- learned TV predictor ------------------------------------------------------

int lf_tvnet_param_count(void) { return tvnet_param_count(); }

size_t lf_tvnet_work_bytes(int n, int64_t H, int64_t W) { return tvnet_work_bytes(n, H, W); }

int lf_tvnet_forward(const float* params, int n, int64_t H, int64_t W, const float* f, const float* lam, float* pred,
                     void* work, size_t work_size, void* stream) {
  return guarded([&] {
    if (!params || !f || !lam || !pred || !work) fail(LF_ERR_BAD_ARG, "null argument");
    tvnet_forward(params, n, H, W, f, lam, pred, work, work_size, reinterpret_cast<hipStream_t>(stream));
  });
}

int lf_tvnet_forward_noalloc(const float* params, int n, int64_t H, int64_t W, const float* f, const float* lam,
                             float* pred, void* stream) {
  return guarded([&] {
    if (!params || !f || !lam || !pred) fail(LF_ERR_BAD_ARG, "null argument");
    tvnet_forward_noalloc(params, n, H, W, f, lam, pred, reinterpret_cast<hipStream_t>(stream));
  });
}

int lf_tvnet_gradients(const float* params, int n, int64_t H, int64_t W, const float* f, const float* lam,
                       const float* target, float* grad, float* loss, void* work, size_t work_size, void* stream) {
  return guarded([&] {
    if (!params || !f || !lam || !target || !grad || !loss || !work) fail(LF_ERR_BAD_ARG, "null argument");
    tvnet_gradients(params, n, H, W, f, lam, target, grad, loss, work, work_size,
                    reinterpret_cast<hipStream_t>(stream));
  });
}

int lf_tvnet_adam(float* params, const float* grad, float* m, float* v, int step, float lr, void* stream) {
  return guarded([&] {
    if (!params || !grad || !m || !v) fail(LF_ERR_BAD_ARG, "null argument");
    tvnet_adam(params, grad, m, v, step, lr, reinterpret_cast<hipStream_t>(stream));
  });
}

}  // extern "C"
