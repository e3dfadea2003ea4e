// Node-parallel device min-cut for lattices the tiled engine does not cover:
// the 4-D kernels (Full4d_8 / Star4d_8, install/code_setup.py:141-153, built
// by the reference with --4d) and, in general, any lattice of up to four
// dimensions.
//
// Same network and outputs as the tiled engine (lf_mincut.hip): residual
// planes res[i][v], signed excess ex[v] (> 0 excess, < 0 remaining sink
// capacity), exact distance labels h[v] to the sink-capable nodes (INF_H =
// cannot reach a sink).  No tiles and no LDS staging: a node's neighbour is
// found from its coordinates in up to four dimensions, so one kernel serves
// every shape and offset set.
//
//   push pulse    = one launch over all nodes: an active node (ex > 0, h <
//                   INF) pushes along its residual arcs to lower neighbours
//                   while it has excess, else relabels to 1 + the lowest
//                   residual neighbour.  Only a node's own thread lowers its
//                   excess and its residuals; other threads only add to them
//                   (atomics), so every unit of flow is conserved.
//   global relabel = level-synchronous BFS from the sink-capable nodes over
//                   reverse residual arcs (one launch per level).
//   termination   = an exact global relabel that leaves no active node: the
//                   remaining excess cannot reach a sink (maximum preflow).
//   outputs       = flow from the remaining sink capacities, S_max = {h = INF},
//                   S_min = flood of the excess nodes over residual arcs.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "lf_engine.hpp"

namespace lfamd {
namespace {

constexpr int GT = 256;
constexpr int GINF = 1 << 30;
constexpr int GMAXK = 48;

struct GGeo {
  long long shape[4];   // leading dimensions padded with 1
  long long stride[4];
  signed char off[GMAXK][4];
  int k;
  long long N;
};

__device__ inline long long g_nbr(const GGeo& g, long long v, int i) {
  long long w = 0, rem = v;
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    const long long c = rem / g.stride[d];
    rem -= c * g.stride[d];
    const long long cn = c + g.off[i][d];
    if (cn < 0 || cn >= g.shape[d]) return -1;
    w += cn * g.stride[d];
  }
  return w;
}

inline unsigned gblocks(long long n) { return (unsigned)((n + GT - 1) / GT); }

// energies (reference layout, gather form) -> red0, alpha0 (accumulated)
__global__ __launch_bounds__(GT) void k_g_reparam(GGeo g, const long long* __restrict__ unary,
                                                  const long long* __restrict__ binary, long long* __restrict__ red0,
                                                  long long* __restrict__ alpha0, int* err) {
  const long long v = (long long)blockIdx.x * GT + threadIdx.x;
  if (v >= g.N) return;
  const int kp = g.k / 2;
  long long r = 2 * (unary[2 * v + 1] - unary[2 * v]);
  bool bad = false;
  for (int e = 0; e < kp; ++e) {
    if (g_nbr(g, v, e) >= 0) {
      const long long* E = binary + (v * kp + e) * 4;
      const long long q1 = E[2] - E[0], q12 = E[0] + E[3] - E[1] - E[2];
      bad |= q12 > 0;
      r += 2 * q1 + q12;
      alpha0[(long long)e * g.N + v] += -q12;
    }
    const long long u = g_nbr(g, v, e + kp);  // u + dir[e] = v
    if (u >= 0) {
      const long long* E = binary + (u * kp + e) * 4;
      const long long q2 = E[1] - E[0], q12 = E[0] + E[3] - E[1] - E[2];
      bad |= q12 > 0;
      r += 2 * q2 + q12;
      alpha0[(long long)(e + kp) * g.N + v] += -q12;
    }
  }
  red0[v] += r;
  if (bad) atomicOr(err, 1);
}

template <typename T>
__global__ __launch_bounds__(GT) void k_g_load(GGeo g, const long long* __restrict__ red0,
                                               const long long* __restrict__ alpha0, T* __restrict__ res,
                                               T* __restrict__ ex, int* err) {
  const long long v = (long long)blockIdx.x * GT + threadIdx.x;
  if (v >= g.N) return;
  const long long lim = sizeof(T) == 4 ? 0x3fffffffLL : 0x3fffffffffffffffLL;
  bool bad = false;
  for (int i = 0; i < g.k; ++i) {
    const long long a = alpha0[(long long)i * g.N + v];
    bad |= a < 0 || a > lim;
    res[(long long)i * g.N + v] = (T)a;
  }
  const long long r = red0[v];
  bad |= r > lim || r < -lim;
  ex[v] = (T)(-r);
  if (bad) atomicOr(err, 2);
}

template <typename T>
__global__ __launch_bounds__(GT) void k_g_label_init(long long N, const T* __restrict__ ex, int* __restrict__ h) {
  const long long v = (long long)blockIdx.x * GT + threadIdx.x;
  if (v < N) h[v] = ex[v] < 0 ? 0 : GINF;
}

// one BFS level: in-neighbours w of the nodes at `level` with a residual arc
// w -> v get level + 1
template <typename T>
__global__ __launch_bounds__(GT) void k_g_bfs_level(GGeo g, const T* __restrict__ res, int* __restrict__ h,
                                                    int level, int* changed) {
  const long long v = (long long)blockIdx.x * GT + threadIdx.x;
  if (v >= g.N || h[v] != level) return;
  const int kp = g.k / 2;
  bool any = false;
  for (int i = 0; i < g.k; ++i) {
    const long long w = g_nbr(g, v, i);
    if (w < 0 || h[w] != GINF) continue;
    const int r = i < kp ? i + kp : i - kp;  // w's direction back to v
    if (res[(long long)r * g.N + w] > 0) {
      h[w] = level + 1;
      any = true;
    }
  }
  if (any) atomicOr(changed, 1);
}

template <typename T>
__global__ __launch_bounds__(GT) void k_g_count_active(long long N, const T* __restrict__ ex,
                                                       const int* __restrict__ h, unsigned long long* cnt) {
  const long long v = (long long)blockIdx.x * GT + threadIdx.x;
  const bool act = v < N && ex[v] > 0 && h[v] < GINF;
  const unsigned long long b = __ballot(act);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(cnt, (unsigned long long)__popcll(b));
}

template <typename T>
__device__ inline void g_atomic_add(T* a, T v) {
  if (sizeof(T) == 4)
    atomicAdd(reinterpret_cast<int*>(a), (int)v);
  else
    atomicAdd(reinterpret_cast<unsigned long long*>(a), (unsigned long long)(long long)v);
}

template <typename T>
__global__ __launch_bounds__(GT) void k_g_pulse(GGeo g, T* __restrict__ res, T* __restrict__ ex, int* __restrict__ h) {
  const long long v = (long long)blockIdx.x * GT + threadIdx.x;
  if (v >= g.N) return;
  const int hv = __atomic_load_n(&h[v], __ATOMIC_RELAXED);
  T e = __atomic_load_n(&ex[v], __ATOMIC_RELAXED);
  if (e <= 0 || hv >= GINF) return;
  const int kp = g.k / 2;
  T pushed = 0;
  int lowest = GINF;
  for (int i = 0; i < g.k && e > pushed; ++i) {
    const long long w = g_nbr(g, v, i);
    if (w < 0) continue;
    T* rp = &res[(long long)i * g.N + v];
    const T r = __atomic_load_n(rp, __ATOMIC_RELAXED);
    if (r <= 0) continue;
    const int hw = __atomic_load_n(&h[w], __ATOMIC_RELAXED);
    if (hw < hv) {
      const T d = r < e - pushed ? r : e - pushed;
      g_atomic_add(rp, (T)-d);
      g_atomic_add(&res[(long long)(i < kp ? i + kp : i - kp) * g.N + w], d);
      g_atomic_add(&ex[w], d);
      pushed += d;
    } else if (hw < lowest) {
      lowest = hw;
    }
  }
  if (pushed > 0) {
    g_atomic_add(&ex[v], (T)-pushed);
  } else {
    // no admissible arc: relabel (labels only grow between global relabels)
    const int nh = lowest >= GINF - 1 ? GINF : lowest + 1;
    if (nh > hv) __atomic_store_n(&h[v], nh, __ATOMIC_RELAXED);
  }
}

template <typename T>
__global__ __launch_bounds__(GT) void k_g_flood_init(long long N, const T* __restrict__ ex, int* __restrict__ reach) {
  const long long v = (long long)blockIdx.x * GT + threadIdx.x;
  if (v < N) reach[v] = ex[v] > 0 ? 1 : 0;
}

// nodes reached in the last step (reach == step) reach their residual out-neighbours
template <typename T>
__global__ __launch_bounds__(GT) void k_g_flood_step(GGeo g, const T* __restrict__ res, int* __restrict__ reach,
                                                     int step, int* changed) {
  const long long v = (long long)blockIdx.x * GT + threadIdx.x;
  if (v >= g.N || reach[v] != step) return;
  bool any = false;
  for (int i = 0; i < g.k; ++i) {
    const long long w = g_nbr(g, v, i);
    if (w < 0 || reach[w] != 0) continue;
    if (res[(long long)i * g.N + v] > 0) {
      reach[w] = step + 1;
      any = true;
    }
  }
  if (any) atomicOr(changed, 1);
}

template <typename T>
__global__ __launch_bounds__(GT) void k_g_flow(long long N, const long long* __restrict__ red0, const T* __restrict__ ex,
                                               unsigned long long* out) {
  const long long v = (long long)blockIdx.x * GT + threadIdx.x;
  long long f = 0;
  if (v < N) {
    const long long r0 = red0[v], rem = -(long long)ex[v];
    f = (r0 > 0 ? r0 : 0) - (rem > 0 ? rem : 0);
  }
  for (int o = 32; o > 0; o >>= 1) f += __shfl_down(f, o, 64);
  if ((threadIdx.x & 63) == 0 && f != 0) atomicAdd(out, (unsigned long long)f);
}

template <typename T>
__global__ __launch_bounds__(GT) void k_g_cut(long long N, const int* __restrict__ reach, const int* __restrict__ h,
                                              int which, int* __restrict__ out) {
  const long long v = (long long)blockIdx.x * GT + threadIdx.x;
  if (v >= N) return;
  out[v] = which == LF_CUT_SMAX ? (h[v] >= GINF ? 1 : 0) : (reach[v] != 0 ? 1 : 0);
}

template <typename T>
class GenericEngine final : public MinCutEngine {
 public:
  explicit GenericEngine(const LatticeGeom& geo) {
    const int nd = (int)geo.dims.size();
    if (nd < 1 || nd > 4) fail(LF_ERR_UNSUPPORTED, "lattices of 1 to 4 dimensions");
    if (geo.k > GMAXK) fail(LF_ERR_UNSUPPORTED, "kernel too large");
    std::memset(&g_, 0, sizeof(g_));
    for (int d = 0; d < 4; ++d) g_.shape[d] = 1;
    for (int d = 0; d < nd; ++d) g_.shape[4 - nd + d] = geo.dims[d];
    long long s = 1;
    for (int d = 3; d >= 0; --d) {
      g_.stride[d] = s;
      s *= g_.shape[d];
    }
    g_.N = s;
    g_.k = geo.k;
    for (int i = 0; i < geo.k; ++i)
      for (int d = 0; d < nd; ++d) g_.off[i][4 - nd + d] = (signed char)geo.offsets[(size_t)i * nd + d];
    N_ = s;
    LF_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    LF_HIP(hipMalloc(&red0_, sizeof(long long) * N_));
    LF_HIP(hipMalloc(&alpha0_, sizeof(long long) * N_ * g_.k));
    LF_HIP(hipMalloc(&res_, sizeof(T) * N_ * g_.k));
    LF_HIP(hipMalloc(&ex_, sizeof(T) * N_));
    LF_HIP(hipMalloc(&h_, sizeof(int) * N_));
    LF_HIP(hipMalloc(&reach_, sizeof(int) * N_));
    LF_HIP(hipMalloc(&small_, 64));
    LF_HIP(hipHostMalloc(&host_, 64, hipHostMallocDefault));
    reset_energies();
  }
  ~GenericEngine() override {
    hipStreamSynchronize(s_);
    hipFree(red0_); hipFree(alpha0_); hipFree(res_); hipFree(ex_); hipFree(h_); hipFree(reach_); hipFree(small_);
    hipHostFree(host_);
    hipStreamDestroy(s_);
  }
  hipStream_t stream() const override { return s_; }

  void reset_energies() override {
    LF_HIP(hipMemsetAsync(red0_, 0, sizeof(long long) * N_, s_));
    LF_HIP(hipMemsetAsync(alpha0_, 0, sizeof(long long) * N_ * g_.k, s_));
    solved = false;
  }

  void add_energies_device(const int64_t* du, const int64_t* db, hipStream_t s) override {
    if (s != s_) {
      hipEvent_t ev;
      LF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      LF_HIP(hipEventRecord(ev, s));
      LF_HIP(hipStreamWaitEvent(s_, ev, 0));
      LF_HIP(hipEventDestroy(ev));
    }
    int* err = reinterpret_cast<int*>(small_);
    LF_HIP(hipMemsetAsync(err, 0, sizeof(int), s_));
    hipLaunchKernelGGL(k_g_reparam, dim3(gblocks(N_)), dim3(GT), 0, s_, g_, reinterpret_cast<const long long*>(du),
                       reinterpret_cast<const long long*>(db), red0_, alpha0_, err);
    LF_HIP(hipGetLastError());
    if (read_int(err) & 1) fail(LF_ERR_NONSUBMODULAR, "pairwise term with E00 + E11 > E01 + E10 (not submodular)");
    solved = false;
  }

  void run(const lf_tuning& tun) override {
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    stats = lf_stats{};
    int* err = reinterpret_cast<int*>(small_);
    LF_HIP(hipMemsetAsync(err, 0, sizeof(int), s_));
    hipLaunchKernelGGL((k_g_load<T>), dim3(gblocks(N_)), dim3(GT), 0, s_, g_, red0_, alpha0_, res_, ex_, err);
    LF_HIP(hipGetLastError());
    if (read_int(err) & 2)
      fail(LF_ERR_OVERFLOW, sizeof(T) == 4 ? "capacities exceed the int32 range; use cap_bytes=8"
                                           : "capacities exceed the int64 range");
    const int R = tun.rounds_per_relabel > 0 ? tun.rounds_per_relabel : 16;
    // every cycle starts from exact labels, so each active node pushes in its
    // first pulse; the default bound only turns a defect into an error
    const long long max_cycles = tun.max_cycles > 0 ? tun.max_cycles : 64 * (N_ + 16);
    for (long long cyc = 0;; ++cyc) {
      const auto tb = clk::now();
      const long long active = relabel();
      stats.bfs_seconds += std::chrono::duration<double>(clk::now() - tb).count();
      ++stats.cycles;
      if (active == 0) break;
      if (cyc >= max_cycles) fail(LF_ERR_STATE, "max_cycles reached before the flow became maximum");
      for (int r = 0; r < R; ++r)
        hipLaunchKernelGGL((k_g_pulse<T>), dim3(gblocks(N_)), dim3(GT), 0, s_, g_, res_, ex_, h_);
      LF_HIP(hipGetLastError());
      stats.rounds += R;
    }
    // S_min: flood of the excess nodes over residual arcs
    const auto tf = clk::now();
    hipLaunchKernelGGL((k_g_flood_init<T>), dim3(gblocks(N_)), dim3(GT), 0, s_, N_, ex_, reach_);
    for (int step = 1;; ++step) {
      int* ch = reinterpret_cast<int*>(small_);
      LF_HIP(hipMemsetAsync(ch, 0, sizeof(int), s_));
      hipLaunchKernelGGL((k_g_flood_step<T>), dim3(gblocks(N_)), dim3(GT), 0, s_, g_, res_, reach_, step, ch);
      ++stats.flood_launches;
      if (!read_int(ch)) break;
    }
    stats.flood_seconds = std::chrono::duration<double>(clk::now() - tf).count();
    unsigned long long* fl = reinterpret_cast<unsigned long long*>(small_ + 8);
    LF_HIP(hipMemsetAsync(fl, 0, sizeof(unsigned long long), s_));
    hipLaunchKernelGGL((k_g_flow<T>), dim3(gblocks(N_)), dim3(GT), 0, s_, N_, red0_, ex_, fl);
    LF_HIP(hipGetLastError());
    LF_HIP(hipMemcpyAsync(host_, fl, sizeof(unsigned long long), hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    flow = (int64_t)*reinterpret_cast<unsigned long long*>(host_);
    stats.seconds = std::chrono::duration<double>(clk::now() - t0).count();
    solved = true;
  }

  void cut_device(int32_t* d_out, int which, hipStream_t s) override {
    hipLaunchKernelGGL((k_g_cut<T>), dim3(gblocks(N_)), dim3(GT), 0, s_, N_, reach_, h_, which, d_out);
    LF_HIP(hipGetLastError());
    if (s == s_) return;
    hipEvent_t ev;
    LF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    LF_HIP(hipEventRecord(ev, s_));
    LF_HIP(hipStreamWaitEvent(s, ev, 0));
    LF_HIP(hipEventDestroy(ev));
  }

  void pf_run(const lf_tuning&) override { unsupported(); }
  void pf_run_layers(const lf_tuning&, int) override { unsupported(); }
  void pf_run_layers_guided(const lf_tuning&, int, const float*, float, int, double, double,
                            lf_guided_stats*) override {
    unsupported();
  }
  void set_tv_network(const int64_t*, const int64_t*, int) override { unsupported(); }
  void pf_values_device(int64_t*, hipStream_t) override { unsupported(); }
  void pf_cut_device(int32_t*, hipStream_t) override { unsupported(); }
  void pf_regions_device(int32_t*, int64_t*, hipStream_t) override { unsupported(); }
  void pf_check_regions(const int64_t*, const int32_t*, int, int64_t*, uint8_t*) override { unsupported(); }
  void set_base_network(const int64_t*, const int64_t*) override { unsupported(); }
  void batch_set_layout(const std::vector<BatchRect>&) override { unsupported(); }
  void batch_add_energies_device(int, const int64_t*, const int64_t*, hipStream_t) override { unsupported(); }
  void batch_flows(int64_t*) override { unsupported(); }
  void batch_cut_device(int32_t*, int, hipStream_t) override { unsupported(); }

 private:
  [[noreturn]] static void unsupported() {
    fail(LF_ERR_UNSUPPORTED, "parametric and batched solves need a lattice of 1 to 3 dimensions");
  }
  int read_int(const int* d) {
    LF_HIP(hipMemcpyAsync(host_, d, sizeof(int), hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    return *reinterpret_cast<int*>(host_);
  }
  // exact distance labels; returns the number of active nodes
  long long relabel() {
    hipLaunchKernelGGL((k_g_label_init<T>), dim3(gblocks(N_)), dim3(GT), 0, s_, N_, ex_, h_);
    int* ch = reinterpret_cast<int*>(small_);
    for (int level = 0;; ++level) {
      LF_HIP(hipMemsetAsync(ch, 0, sizeof(int), s_));
      hipLaunchKernelGGL((k_g_bfs_level<T>), dim3(gblocks(N_)), dim3(GT), 0, s_, g_, res_, h_, level, ch);
      ++stats.bfs_launches;
      if (!read_int(ch)) break;
    }
    unsigned long long* cnt = reinterpret_cast<unsigned long long*>(small_ + 16);
    LF_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), s_));
    hipLaunchKernelGGL((k_g_count_active<T>), dim3(gblocks(N_)), dim3(GT), 0, s_, N_, ex_, h_, cnt);
    LF_HIP(hipGetLastError());
    LF_HIP(hipMemcpyAsync(host_, cnt, sizeof(unsigned long long), hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    return (long long)*reinterpret_cast<unsigned long long*>(host_);
  }

  GGeo g_{};
  long long N_ = 0;
  hipStream_t s_ = nullptr;
  long long* red0_ = nullptr;
  long long* alpha0_ = nullptr;
  T* res_ = nullptr;
  T* ex_ = nullptr;
  int* h_ = nullptr;
  int* reach_ = nullptr;
  char* small_ = nullptr;
  void* host_ = nullptr;
};

}  // namespace

std::unique_ptr<MinCutEngine> make_generic_engine(const LatticeGeom& g, int cap_bytes) {
  if (cap_bytes == 4) return std::unique_ptr<MinCutEngine>(new GenericEngine<int>(g));
  return std::unique_ptr<MinCutEngine>(new GenericEngine<long long>(g));
}

}  // namespace lfamd
