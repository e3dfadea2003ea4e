// The TV regularization path on the device: the exact solution at lambda_max
// by the device divide and conquer (its final regions are the level sets at
// s = 1), level-set checks of large level sets by the device min-cut engine
// (MinCutEngine::pf_check_regions), the event loop in lf_tvpath_core.cpp, and
// the trace of every pixel through the path at the requested lambdas as one
// kernel (getRegularizationPath / traceRegularizationPath,
// solver/tv/tv_solver.hpp:181-244,566-642).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/latticeflow_amd.h"
#include "lf_engine.hpp"
#include "lf_kernels.hpp"
#include "lf_tv.hpp"
#include "lf_tvpath.hpp"
#include "lf_tvpath_core.hpp"

namespace lfamd {

namespace {

constexpr int kTraceThreads = 256;
// Level sets of at most this many nodes are checked on the host: a device
// check costs 5-8 ms (a whole-lattice region solve), a host cut of a level
// set of a few thousand nodes as much (128^2 / 256^2 Voronoi paths: 2.7 / 30.2 s
// at 32768, 2.2 / 13.8 s at 2048, 2.9 / 18.6 s at 512)
constexpr int64_t kDefaultHostMax = 2048;

class DeviceChecker : public RegionChecker {
 public:
  explicit DeviceChecker(MinCutEngine* e) : e_(e) {}
  void check(const int64_t* r, const int32_t* reg, int nreg, int64_t* cnt_low, uint8_t* low) override {
    e_->pf_check_regions(r, reg, nreg, cnt_low, low);
  }

 private:
  MinCutEngine* e_;
};

struct TraceDev {
  const int32_t* initial;
  const double* s_lo;
  const double* a;
  const double* b;
  const int32_t* up;
  const int32_t* part_begin;
  const int32_t* part_end;
  const int32_t* part_seg;
  const long long* mem_begin;
  const int32_t* members;
};

__device__ inline bool sorted_contains(const int32_t* m, long long lo, long long hi, int32_t v) {
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    const int32_t x = m[mid];
    if (x == v) return true;
    if (x < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return false;
}

// One pixel per thread: from its level set at s = 1 through joins (up) and
// splits (the part holding the pixel) to the level set alive at each s, the
// values written for the lambdas in decreasing order.
__global__ __launch_bounds__(kTraceThreads) void k_trace(TraceDev t, long long n, int nl,
                                                         const double* __restrict__ s_desc,
                                                         const int* __restrict__ order, double* __restrict__ out,
                                                         int* __restrict__ err) {
  const long long v = (long long)blockIdx.x * kTraceThreads + threadIdx.x;
  if (v >= n) return;
  int id = t.initial[v];
  for (int l = 0; l < nl; ++l) {
    const double s = s_desc[l];
    while (t.s_lo[id] > s) {
      const int u = t.up[id];
      if (u >= 0) {
        id = u;
        continue;
      }
      const int si = -u - 1;
      int next = -1;
      for (int p = t.part_begin[si]; p < t.part_end[si]; ++p)
        if (sorted_contains(t.members, t.mem_begin[p], t.mem_begin[p + 1], (int32_t)v)) {
          next = t.part_seg[p];
          break;
        }
      if (next < 0) {
        err[0] = 1;  // every pixel of a split level set is in one of its parts
        return;
      }
      id = next;
    }
    out[(long long)order[l] * n + v] = t.a[id] + s * t.b[id];
  }
}

template <typename T>
T* upload(const std::vector<T>& h, hipStream_t s) {
  T* d = nullptr;
  LF_HIP(hipMalloc((void**)&d, sizeof(T) * std::max<size_t>(h.size(), 1)));
  if (!h.empty()) LF_HIP(hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
  return d;
}
}  // namespace

struct TvPath::Impl {
  int64_t nx = 0, ny = 0, n = 0;
  double lmax = 0;
  bool constant = false;  // a constant image is its own path
  TvBase b;
  TvPathStats st;
  hipStream_t stream = nullptr;
  TraceDev td{};
  std::vector<void*> bufs;

  ~Impl() {
    for (void* p : bufs) hipFree(p);
    if (stream) hipStreamDestroy(stream);
  }
};

TvPath::TvPath(const char* kernel, int64_t nx, int64_t ny, const double* f, double lambda_max, int64_t host_max)
    : im_(new Impl) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  Impl& P = *im_;
  if (!kernel || !f) fail(LF_ERR_BAD_ARG, "null argument");
  const KernelDef* kd = find_kernel(kernel);
  if (!kd) fail(LF_ERR_BAD_NAME, std::string("Bad Kernel or Solver Name: ") + kernel);
  if (kd->ndim != 2 || !kd->geocut)
    fail(LF_ERR_BAD_ARG, "Kernel is not valid for GeoCuts or TV Minimization (2-D geocut kernels only).");
  if (nx <= 0 || ny <= 0) fail(LF_ERR_BAD_ARG, "empty image");
  if (nx * ny >= (1LL << 31)) fail(LF_ERR_UNSUPPORTED, "the event path handles fewer than 2^31 pixels");
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) fail(LF_ERR_DEVICE, "no HIP device available");
  P.nx = nx;
  P.ny = ny;
  P.n = nx * ny;
  // tv_solver.hpp:91: the network is built at max(lambda, 1e-8)
  P.lmax = std::max(lambda_max, 1e-8);
  P.b = tv_base(P.n, f);
  LF_HIP(hipStreamCreate(&P.stream));
  if (!(P.b.S > 0)) {
    P.constant = true;
    P.st.seconds = std::chrono::duration<double>(clk::now() - t0).count();
    return;
  }
  std::vector<int> offs;
  std::vector<double> w;
  lattice_edge_order(*kd, {nx, ny}, offs, w);
  const int k = (int)kd->edges.size();
  TvPathProblem prob;
  prob.n = P.n;
  prob.k = k;
  prob.nbr.assign((size_t)P.n * k, -1);
  for (int64_t v = 0; v < P.n; ++v) {
    const int64_t y = v / ny, x = v % ny;
    for (int i = 0; i < k; ++i) {
      const int64_t yy = y + offs[2 * i], xx = x + offs[2 * i + 1];
      if (yy >= 0 && yy < nx && xx >= 0 && xx < ny) prob.nbr[(size_t)v * k + i] = (int32_t)(yy * ny + xx);
    }
  }
  std::vector<int64_t> pwf((size_t)k / 2);
  prob.acap.assign((size_t)k, 0);
  for (int e = 0; e < k / 2; ++e) {
    pwf[e] = tv_pwf(P.b, P.lmax, w[e]);
    if (pwf[e] > ((int64_t)1 << 56) / k) fail(LF_ERR_OVERFLOW, "lambda too small for the 64-bit fixed point");
    prob.acap[e] = prob.acap[e + k / 2] = 2 * pwf[e];
  }
  prob.fv.resize((size_t)P.n);
  for (int64_t v = 0; v < P.n; ++v) prob.fv[v] = 2 * P.b.uf[v];
  LatticeGeom g;
  g.ndim = 2;
  g.k = k;
  g.D = 1;
  g.H = nx;
  g.W = ny;
  g.dims = {nx, ny};
  g.offsets = offs;
  for (int i = 0; i < k; ++i) {
    g.dz.push_back(0);
    g.dy.push_back(offs[i * 2]);
    g.dx.push_back(offs[i * 2 + 1]);
  }
  std::unique_ptr<MinCutEngine> eng = make_mincut_engine(g, 8);
  eng->set_tv_network(P.b.uf.data(), pwf.data(), 1);
  // the exact solution at lambda_max (s = 1)
  eng->pf_run_layers(lf_tuning{}, 1);
  prob.reg.resize((size_t)P.n);
  prob.r.resize((size_t)P.n);
  {
    hipStream_t s = eng->stream();
    int32_t* dreg = nullptr;
    int64_t* dr = nullptr;
    LF_HIP(hipMallocAsync((void**)&dreg, sizeof(int32_t) * P.n, s));
    LF_HIP(hipMallocAsync((void**)&dr, sizeof(int64_t) * P.n, s));
    eng->pf_regions_device(dreg, dr, s);
    LF_HIP(hipMemcpyAsync(prob.reg.data(), dreg, sizeof(int32_t) * P.n, hipMemcpyDeviceToHost, s));
    LF_HIP(hipMemcpyAsync(prob.r.data(), dr, sizeof(int64_t) * P.n, hipMemcpyDeviceToHost, s));
    LF_HIP(hipFreeAsync(dreg, s));
    LF_HIP(hipFreeAsync(dr, s));
    LF_HIP(hipStreamSynchronize(s));
  }
  P.st.initial_seconds = std::chrono::duration<double>(clk::now() - t0).count();
  DeviceChecker checker(eng.get());
  TvPathCore core(std::move(prob), &checker, host_max == 0 ? kDefaultHostMax : host_max, &P.st);
  eng.reset();
  // the segment table for the trace kernel
  const TvTrace t = core.trace_table();
  auto keep = [&](void* p) {
    P.bufs.push_back(p);
    return p;
  };
  P.td.initial = (const int32_t*)keep(upload(t.initial, P.stream));
  P.td.s_lo = (const double*)keep(upload(t.s_lo, P.stream));
  P.td.a = (const double*)keep(upload(t.a, P.stream));
  P.td.b = (const double*)keep(upload(t.b, P.stream));
  P.td.up = (const int32_t*)keep(upload(t.up, P.stream));
  P.td.part_begin = (const int32_t*)keep(upload(t.part_begin, P.stream));
  P.td.part_end = (const int32_t*)keep(upload(t.part_end, P.stream));
  P.td.part_seg = (const int32_t*)keep(upload(t.part_seg, P.stream));
  std::vector<long long> mb(t.mem_begin.begin(), t.mem_begin.end());
  P.td.mem_begin = (const long long*)keep(upload(mb, P.stream));
  P.td.members = (const int32_t*)keep(upload(t.members, P.stream));
  LF_HIP(hipStreamSynchronize(P.stream));
  P.st.seconds = std::chrono::duration<double>(clk::now() - t0).count();
}

TvPath::~TvPath() {}

const TvPathStats& TvPath::stats() const { return im_->st; }
double TvPath::lambda_max() const { return im_->lmax; }
int64_t TvPath::nx() const { return im_->nx; }
int64_t TvPath::ny() const { return im_->ny; }

// getRegularizationPath (tv_solver.hpp:181-244): the values of the path at
// s = lambda / lambda_max, mapped to the output as the reference does.
void TvPath::eval(int64_t n_lambda, const double* lambdas, double* R) const {
  const Impl& P = *im_;
  if (n_lambda < 0) fail(LF_ERR_BAD_ARG, "negative lambda count");
  if (n_lambda > 0 && (!lambdas || !R)) fail(LF_ERR_BAD_ARG, "null argument");
  for (int64_t l = 0; l < n_lambda; ++l)
    if (lambdas[l] > P.lmax * (1 + 1e-12))
      fail(LF_ERR_OUT_OF_RANGE, "Given lambda values not in calculated range.");
  if (n_lambda == 0) return;
  if (P.constant) {
    std::fill(R, R + n_lambda * P.n, P.b.M);
    return;
  }
  if (n_lambda > (int64_t)1 << 30) fail(LF_ERR_BAD_ARG, "too many lambdas");
  const int nl = (int)n_lambda;
  std::vector<int> order(nl);
  std::iota(order.begin(), order.end(), 0);
  std::vector<double> s(nl);
  for (int l = 0; l < nl; ++l) s[l] = std::min(std::max(lambdas[l], 0.0) / P.lmax, 1.0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return s[x] > s[y]; });
  std::vector<double> s_desc(nl);
  for (int l = 0; l < nl; ++l) s_desc[l] = s[order[l]];
  hipStream_t st = P.stream;
  double* d_s = upload(s_desc, st);
  int* d_order = upload(order, st);
  double* d_out = nullptr;
  int* d_err = nullptr;
  struct Free {
    std::vector<void*> p;
    ~Free() {
      for (void* x : p) hipFree(x);
    }
  } fr{{d_s, d_order}};
  LF_HIP(hipMalloc((void**)&d_out, sizeof(double) * (size_t)nl * P.n));
  fr.p.push_back(d_out);
  LF_HIP(hipMalloc((void**)&d_err, sizeof(int)));
  fr.p.push_back(d_err);
  LF_HIP(hipMemsetAsync(d_err, 0, sizeof(int), st));
  const long long blocks = (P.n + kTraceThreads - 1) / kTraceThreads;
  hipLaunchKernelGGL(k_trace, dim3((unsigned)blocks), dim3(kTraceThreads), 0, st, P.td, (long long)P.n, nl, d_s,
                     d_order, d_out, d_err);
  LF_HIP(hipGetLastError());
  std::vector<double> rp((size_t)nl * P.n);
  int err = 0;
  LF_HIP(hipMemcpyAsync(rp.data(), d_out, sizeof(double) * rp.size(), hipMemcpyDeviceToHost, st));
  LF_HIP(hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
  LF_HIP(hipStreamSynchronize(st));
  if (err) fail(LF_ERR_STATE, "TV path: pixel lost at a split");
  for (int l = 0; l < nl; ++l) tv_output_path(P.b, rp.data() + (size_t)l * P.n, P.n, R + (size_t)l * P.n);
}

}  // namespace lfamd
