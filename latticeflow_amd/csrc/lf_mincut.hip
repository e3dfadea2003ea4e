// lf_mincut.hip -- exact lattice min-cut on MI355X (gfx950).
//
// Replaces the sequential highest-label PRFlow engine of the reference
// (solver/network_flow/push_relabel.hpp:1069-1354; hot start BFS :305-415,
// gap/flip :677-711) with a tiled, synchronous push-relabel:
//
//   state (structure of arrays in HBM, C order over the unpadded lattice):
//     res[i][v]  residual of the arc v -> v+dir[i]      (T = int32 / int64)
//     ex[v]      >0 excess, <0 remaining sink capacity (= -reduction of the
//                reference node, network_flow_node.hpp:171-180)
//     h[v]       distance label; INF_H = cannot reach a sink
//     mbox[i][v] flow v pushed along dir i into a node of another tile, not
//                yet applied there (the "mailbox")
//
//   one push round  = one launch of k_push: a workgroup owns one tile, stages
//     its residuals / excess / labels (+ a halo of labels) in LDS, collects its
//     mailbox, runs `local_iters` synchronous push / relabel iterations in LDS
//     (pushes along downhill residual arcs with labels frozen, then relabels
//     with residuals frozen), and writes the tile back.  Flow pushed out of the
//     tile goes to the mailbox with a device-scope atomic add; the receiver
//     takes it with an atomic exchange, so every unit of flow is applied once.
//   global relabel  = exact BFS distance to the sink-capable nodes over
//     residual arcs (the reference's hot start), computed as tile-local
//     Bellman-Ford sweeps in LDS over a 1-bit-per-arc residual mask, repeated
//     until no tile border changes.  Nodes it cannot reach get INF_H (the
//     reference's gap / flip to the source side).
//   termination      = no node with ex > 0 and h < INF_H right after an exact
//     global relabel: the remaining excess cannot reach a sink, so the flow is
//     maximum (a maximum preflow).
//   outputs          = flow = sum(initial sink capacity) - sum(remaining sink
//     capacity); S_max = {h == INF_H}; S_min = flood fill from the excess
//     nodes over residual arcs (tile-local, iterated like the BFS).
//
// Correctness does not depend on the labels between global relabels (they only
// steer pushes); flow conservation and capacities are kept exactly by
// construction, and the stopping test uses exact labels.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "lf_engine.hpp"

namespace lfamd {
namespace {

constexpr int NTHREADS = 256;
constexpr int INF_H = 1 << 30;
constexpr int MAXK = 32;

struct DirTab {
  int8_t dz[MAXK], dy[MAXK], dx[MAXK];
};

template <int TD_, int TH_, int TW_, int RZ_, int RY_, int RX_>
struct Geo {
  static constexpr int TD = TD_, TH = TH_, TW = TW_, RZ = RZ_, RY = RY_, RX = RX_;
  static constexpr int NT = TD * TH * TW;
  static constexpr int HD = TD + 2 * RZ, HH = TH + 2 * RY, HW = TW + 2 * RX;
  static constexpr int NH = HD * HH * HW;
  static constexpr int NPT = NT / NTHREADS;
  static_assert(NT % NTHREADS == 0, "tile must be a multiple of the block");
};

template <typename T>
struct KP {
  T* res;
  T* ex;
  int* h;
  T* mbox;
  int* flag_in;
  int* flag_out;
  unsigned long long* counter;
  int D, H, W;
  long long N;
  int ntx, nty, ntz, ntiles, nblocks;
  int iters;
  DirTab dirs;
};

// ---- small device helpers -------------------------------------------------

__device__ inline void lds_add(int* a, int v) { atomicAdd(a, v); }
__device__ inline void lds_add(long long* a, long long v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(a), static_cast<unsigned long long>(v));
}
__device__ inline void gadd(int* a, int v) { atomicAdd(a, v); }
__device__ inline void gadd(long long* a, long long v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(a), static_cast<unsigned long long>(v));
}
__device__ inline int gtake(int* a) { return atomicExch(a, 0); }
__device__ inline long long gtake(long long* a) {
  return static_cast<long long>(atomicExch(reinterpret_cast<unsigned long long*>(a), 0ull));
}

// Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous run
// of tiles so that halo reads of neighbouring tiles hit the same L2.
__device__ inline int tile_of_block(int nblocks) {
  const int per = nblocks >> 3;
  const int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}

template <class G>
struct TileCtx {
  int x0, y0, z0;
  long long g[G::NPT];   // global index of owned node j, -1 outside the lattice
  int hi[G::NPT];        // LDS halo index of owned node j
  unsigned m_lat[G::NPT];  // bit i: neighbour i inside the lattice
  unsigned m_tile[G::NPT]; // bit i: neighbour i inside this tile
};

template <int K, class G, typename T>
__device__ inline void tile_setup(const KP<T>& p, int tile, TileCtx<G>& c) {
  const int tx = tile % p.ntx, ty = (tile / p.ntx) % p.nty, tz = tile / (p.ntx * p.nty);
  c.x0 = tx * G::TW; c.y0 = ty * G::TH; c.z0 = tz * G::TD;
#pragma unroll
  for (int j = 0; j < G::NPT; ++j) {
    const int t = threadIdx.x + j * NTHREADS;
    const int z = t / (G::TH * G::TW), y = (t / G::TW) % G::TH, x = t % G::TW;
    const int gz = c.z0 + z, gy = c.y0 + y, gx = c.x0 + x;
    const bool valid = gz < p.D && gy < p.H && gx < p.W;
    c.g[j] = valid ? ((long long)gz * p.H + gy) * p.W + gx : -1;
    c.hi[j] = ((z + G::RZ) * G::HH + (y + G::RY)) * G::HW + (x + G::RX);
    unsigned ml = 0, mt = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int dz = p.dirs.dz[i], dy = p.dirs.dy[i], dx = p.dirs.dx[i];
      const int nz = gz + dz, ny = gy + dy, nx = gx + dx;
      if (valid && nz >= 0 && nz < p.D && ny >= 0 && ny < p.H && nx >= 0 && nx < p.W) ml |= 1u << i;
      const int lz = z + dz, ly = y + dy, lx = x + dx;
      if (lz >= 0 && lz < G::TD && ly >= 0 && ly < G::TH && lx >= 0 && lx < G::TW) mt |= 1u << i;
    }
    c.m_lat[j] = ml;
    c.m_tile[j] = mt;
  }
}

template <class G, typename T>
__device__ inline void load_halo_h(const KP<T>& p, const TileCtx<G>& c, const int* __restrict__ src, int* s_h) {
  for (int q = threadIdx.x; q < G::NH; q += NTHREADS) {
    const int hz = q / (G::HH * G::HW), hy = (q / G::HW) % G::HH, hx = q % G::HW;
    const int gz = c.z0 + hz - G::RZ, gy = c.y0 + hy - G::RY, gx = c.x0 + hx - G::RX;
    int v = INF_H;
    if (gz >= 0 && gz < p.D && gy >= 0 && gy < p.H && gx >= 0 && gx < p.W)
      v = src[((long long)gz * p.H + gy) * p.W + gx];
    s_h[q] = v;
  }
}

// Flag the neighbouring tile that holds neighbour i of a node at global (gz,gy,gx).
template <class G, typename T>
__device__ inline void flag_neighbor_tile(const KP<T>& p, int gz, int gy, int gx, int* flags) {
  const int tz = gz / G::TD, ty = gy / G::TH, tx = gx / G::TW;
  flags[(tz * p.nty + ty) * p.ntx + tx] = 1;
}

// ---- the push / relabel round -----------------------------------------------

template <int K, class G, typename T>
__global__ __launch_bounds__(NTHREADS) void k_push(KP<T> p) {
  constexpr int NT = G::NT;
  __shared__ T s_res[K * NT];
  __shared__ T s_ex[NT];
  __shared__ T s_in[NT];
  __shared__ int s_h[G::NH];
  __shared__ int s_flag;

  const int tile = tile_of_block(p.nblocks);
  if (tile >= p.ntiles) return;
  if (threadIdx.x == 0) {
    s_flag = p.flag_in[tile];
    p.flag_in[tile] = 0;
  }
  __syncthreads();
  if (!s_flag) return;

  TileCtx<G> c;
  tile_setup<K, G, T>(p, tile, c);
  const long long N = p.N;

  int hj[K], tj[K];
  long long gj[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    hj[i] = (p.dirs.dz[i] * G::HH + p.dirs.dy[i]) * G::HW + p.dirs.dx[i];
    tj[i] = (p.dirs.dz[i] * G::TH + p.dirs.dy[i]) * G::TW + p.dirs.dx[i];
    gj[i] = ((long long)p.dirs.dz[i] * p.H + p.dirs.dy[i]) * p.W + p.dirs.dx[i];
  }

  // stage the tile
#pragma unroll
  for (int j = 0; j < G::NPT; ++j) {
    const int t = threadIdx.x + j * NTHREADS;
    const long long g = c.g[j];
    if (g >= 0) {
#pragma unroll
      for (int i = 0; i < K; ++i) s_res[i * NT + t] = p.res[i * N + g];
      s_ex[t] = p.ex[g];
    } else {
#pragma unroll
      for (int i = 0; i < K; ++i) s_res[i * NT + t] = 0;
      s_ex[t] = 0;
    }
    s_in[t] = 0;
  }
  load_halo_h<G, T>(p, c, p.h, s_h);

  // collect flow that neighbouring tiles pushed into this tile
#pragma unroll
  for (int j = 0; j < G::NPT; ++j) {
    const int t = threadIdx.x + j * NTHREADS;
    const long long g = c.g[j];
    const unsigned outside = c.m_lat[j] & ~c.m_tile[j];
    if (g >= 0 && outside) {
#pragma unroll
      for (int i = 0; i < K; ++i) {
        if (outside & (1u << i)) {
          T* slot = &p.mbox[(long long)((i + K / 2) % K) * N + g + gj[i]];
          if (__builtin_nontemporal_load(slot) != 0) {
            const T a = gtake(slot);
            s_ex[t] += a;
            s_res[i * NT + t] += a;
          }
        }
      }
    }
  }
  __syncthreads();

  for (int it = 0; it < p.iters; ++it) {
    // push: labels frozen
#pragma unroll
    for (int j = 0; j < G::NPT; ++j) {
      const int t = threadIdx.x + j * NTHREADS;
      const long long g = c.g[j];
      if (g < 0) continue;
      T e = s_ex[t];
      const int hv = s_h[c.hi[j]];
      if (e <= 0 || hv >= INF_H) continue;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        if (e > 0 && (c.m_lat[j] & (1u << i))) {
          const int hn = s_h[c.hi[j] + hj[i]];
          if (hn < hv) {
            const T r = s_res[i * NT + t];
            if (r > 0) {
              const T d = r < e ? r : e;
              s_res[i * NT + t] = r - d;
              e -= d;
              if (c.m_tile[j] & (1u << i)) {
                const int tn = t + tj[i];
                s_res[((i + K / 2) % K) * NT + tn] += d;
                lds_add(&s_in[tn], d);
              } else {
                gadd(&p.mbox[(long long)i * N + g], d);
                const int z = t / (G::TH * G::TW), y = (t / G::TW) % G::TH, x = t % G::TW;
                flag_neighbor_tile<G, T>(p, c.z0 + z + p.dirs.dz[i], c.y0 + y + p.dirs.dy[i],
                                         c.x0 + x + p.dirs.dx[i], p.flag_out);
              }
            }
          }
        }
      }
      s_ex[t] = e;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < G::NPT; ++j) {
      const int t = threadIdx.x + j * NTHREADS;
      const T a = s_in[t];
      if (a != 0) {
        s_ex[t] += a;
        s_in[t] = 0;
      }
    }
    __syncthreads();
    // relabel: residuals frozen; labels only grow, so concurrent reads are safe
    bool act = false;
#pragma unroll
    for (int j = 0; j < G::NPT; ++j) {
      const int t = threadIdx.x + j * NTHREADS;
      if (c.g[j] < 0) continue;
      const T e = s_ex[t];
      const int hv = s_h[c.hi[j]];
      if (e <= 0 || hv >= INF_H) continue;
      int m = INF_H;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        if ((c.m_lat[j] & (1u << i)) && s_res[i * NT + t] > 0) {
          const int hn = s_h[c.hi[j] + hj[i]];
          m = hn < m ? hn : m;
        }
      }
      if (m >= hv) {
        const int nh = (m >= INF_H - 1) ? INF_H : m + 1;
        s_h[c.hi[j]] = nh;
        if (nh < INF_H) act = true;
      } else {
        act = true;
      }
    }
    if (!__syncthreads_or(act)) break;
  }
  __syncthreads();

  bool act = false;
#pragma unroll
  for (int j = 0; j < G::NPT; ++j) {
    const int t = threadIdx.x + j * NTHREADS;
    const long long g = c.g[j];
    if (g < 0) continue;
#pragma unroll
    for (int i = 0; i < K; ++i) p.res[i * N + g] = s_res[i * NT + t];
    const T e = s_ex[t];
    const int hv = s_h[c.hi[j]];
    p.ex[g] = e;
    p.h[g] = hv;
    if (e > 0 && hv < INF_H) act = true;
  }
  if (__syncthreads_or(act)) {
    if (threadIdx.x == 0) p.flag_out[tile] = 1;
  }
}

// ---- mailbox delivery (before any exact pass) --------------------------------

template <int K, class G, typename T>
__global__ __launch_bounds__(NTHREADS) void k_deliver(KP<T> p) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= p.N) return;
  const int x = (int)(v % p.W), y = (int)((v / p.W) % p.H), z = (int)(v / ((long long)p.W * p.H));
  const int lx = x % G::TW, ly = y % G::TH, lz = z % G::TD;
  if (lx >= G::RX && lx < G::TW - G::RX && ly >= G::RY && ly < G::TH - G::RY && lz >= G::RZ &&
      lz < G::TD - G::RZ)
    return;
  T add_ex = 0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int dz = p.dirs.dz[i], dy = p.dirs.dy[i], dx = p.dirs.dx[i];
    const int nz = z + dz, ny = y + dy, nx = x + dx;
    if (nz < 0 || nz >= p.D || ny < 0 || ny >= p.H || nx < 0 || nx >= p.W) continue;
    const int qz = lz + dz, qy = ly + dy, qx = lx + dx;
    if (qz >= 0 && qz < G::TD && qy >= 0 && qy < G::TH && qx >= 0 && qx < G::TW) continue;
    const long long w = ((long long)nz * p.H + ny) * p.W + nx;
    T* slot = &p.mbox[(long long)((i + K / 2) % K) * p.N + w];
    const T a = *slot;
    if (a != 0) {
      *slot = 0;
      add_ex += a;
      p.res[(long long)i * p.N + v] += a;
    }
  }
  if (add_ex != 0) p.ex[v] += add_ex;
}

// ---- exact global relabel -----------------------------------------------------

// mask[v] bit i = (res[i][v] > 0); h[v] = 0 on sink-capable nodes, INF_H elsewhere.
template <int K, typename T, typename MT>
__global__ __launch_bounds__(NTHREADS) void k_bfs_init(KP<T> p, MT* __restrict__ mask) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= p.N) return;
  MT m = 0;
#pragma unroll
  for (int i = 0; i < K; ++i)
    if (p.res[(long long)i * p.N + v] > 0) m |= (MT)1 << i;
  mask[v] = m;
  p.h[v] = p.ex[v] < 0 ? 0 : INF_H;
}

template <int K, class G, typename T, typename MT>
__global__ __launch_bounds__(NTHREADS) void k_bfs_tile(KP<T> p, const MT* __restrict__ mask) {
  __shared__ int s_d[G::NH];
  __shared__ int s_flag;
  const int tile = tile_of_block(p.nblocks);
  if (tile >= p.ntiles) return;
  if (threadIdx.x == 0) {
    s_flag = p.flag_in[tile];
    p.flag_in[tile] = 0;
  }
  __syncthreads();
  if (!s_flag) return;
  TileCtx<G> c;
  tile_setup<K, G, T>(p, tile, c);
  int hj[K];
#pragma unroll
  for (int i = 0; i < K; ++i) hj[i] = (p.dirs.dz[i] * G::HH + p.dirs.dy[i]) * G::HW + p.dirs.dx[i];
  MT mk[G::NPT];
  int d0[G::NPT];
#pragma unroll
  for (int j = 0; j < G::NPT; ++j) mk[j] = c.g[j] >= 0 ? mask[c.g[j]] : (MT)0;
  load_halo_h<G, T>(p, c, p.h, s_d);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < G::NPT; ++j) d0[j] = s_d[c.hi[j]];
  for (;;) {
    bool ch = false;
#pragma unroll
    for (int j = 0; j < G::NPT; ++j) {
      if (c.g[j] < 0) continue;
      const int dv = s_d[c.hi[j]];
      if (dv == 0) continue;
      int best = dv;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        if (mk[j] & ((MT)1 << i)) {
          const int dn = s_d[c.hi[j] + hj[i]] + 1;
          best = dn < best ? dn : best;
        }
      }
      if (best < dv) {
        s_d[c.hi[j]] = best;
        ch = true;
      }
    }
    if (!__syncthreads_or(ch)) break;
  }
  bool border = false;
#pragma unroll
  for (int j = 0; j < G::NPT; ++j) {
    if (c.g[j] < 0) continue;
    const int dv = s_d[c.hi[j]];
    if (dv != d0[j]) {
      p.h[c.g[j]] = dv;
      if (c.m_lat[j] & ~c.m_tile[j]) border = true;
    }
  }
  if (__syncthreads_or(border)) {
    if (threadIdx.x == 0) {
      atomicAdd(p.counter, 1ull);
      const int tx = tile % p.ntx, ty = (tile / p.ntx) % p.nty, tz = tile / (p.ntx * p.nty);
      for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
          for (int e = -1; e <= 1; ++e) {
            const int z = tz + a, y = ty + b, x = tx + e;
            if ((a | b | e) == 0 || z < 0 || z >= p.ntz || y < 0 || y >= p.nty || x < 0 || x >= p.ntx) continue;
            p.flag_out[(z * p.nty + y) * p.ntx + x] = 1;
          }
    }
  }
}

// count active nodes (ex > 0, h < INF) and flag their tiles for the push rounds
template <class G, typename T>
__global__ __launch_bounds__(NTHREADS) void k_count_active(KP<T> p) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  bool a = false;
  if (v < p.N) {
    a = p.ex[v] > 0 && p.h[v] < INF_H;
    if (a) {
      const int x = (int)(v % p.W), y = (int)((v / p.W) % p.H), z = (int)(v / ((long long)p.W * p.H));
      flag_neighbor_tile<G, T>(p, z, y, x, p.flag_out);
    }
  }
  const unsigned long long b = __ballot(a);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(p.counter, (unsigned long long)__popcll(b));
}

// ---- S_min: flood fill from excess nodes over residual arcs --------------------

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_flood_init(KP<T> p, int* __restrict__ reach) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= p.N) return;
  reach[v] = p.ex[v] > 0 ? 1 : 0;
}

template <int K, class G, typename T, typename MT>
__global__ __launch_bounds__(NTHREADS) void k_flood_tile(KP<T> p, const MT* __restrict__ mask, int* __restrict__ reach) {
  __shared__ int s_r[G::NH];
  __shared__ MT s_m[G::NH];
  __shared__ int s_flag;
  const int tile = tile_of_block(p.nblocks);
  if (tile >= p.ntiles) return;
  if (threadIdx.x == 0) {
    s_flag = p.flag_in[tile];
    p.flag_in[tile] = 0;
  }
  __syncthreads();
  if (!s_flag) return;
  TileCtx<G> c;
  tile_setup<K, G, T>(p, tile, c);
  int hj[K];
#pragma unroll
  for (int i = 0; i < K; ++i) hj[i] = (p.dirs.dz[i] * G::HH + p.dirs.dy[i]) * G::HW + p.dirs.dx[i];
  for (int q = threadIdx.x; q < G::NH; q += NTHREADS) {
    const int hz = q / (G::HH * G::HW), hy = (q / G::HW) % G::HH, hx = q % G::HW;
    const int gz = c.z0 + hz - G::RZ, gy = c.y0 + hy - G::RY, gx = c.x0 + hx - G::RX;
    int r = 0;
    MT m = 0;
    if (gz >= 0 && gz < p.D && gy >= 0 && gy < p.H && gx >= 0 && gx < p.W) {
      const long long g = ((long long)gz * p.H + gy) * p.W + gx;
      r = reach[g];
      m = mask[g];
    }
    s_r[q] = r;
    s_m[q] = m;
  }
  __syncthreads();
  int r0[G::NPT];
#pragma unroll
  for (int j = 0; j < G::NPT; ++j) r0[j] = s_r[c.hi[j]];
  for (;;) {
    bool ch = false;
#pragma unroll
    for (int j = 0; j < G::NPT; ++j) {
      if (c.g[j] < 0 || s_r[c.hi[j]]) continue;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        // in-arc from v = w + dir[i]: v's arc along rev(i) points at w
        if (c.m_lat[j] & (1u << i)) {
          const int q = c.hi[j] + hj[i];
          if (s_r[q] && (s_m[q] & ((MT)1 << ((i + K / 2) % K)))) {
            s_r[c.hi[j]] = 1;
            ch = true;
            break;
          }
        }
      }
    }
    if (!__syncthreads_or(ch)) break;
  }
  bool border = false;
#pragma unroll
  for (int j = 0; j < G::NPT; ++j) {
    if (c.g[j] < 0) continue;
    if (s_r[c.hi[j]] != r0[j]) {
      reach[c.g[j]] = 1;
      if (c.m_lat[j] & ~c.m_tile[j]) border = true;
    }
  }
  if (__syncthreads_or(border)) {
    if (threadIdx.x == 0) {
      atomicAdd(p.counter, 1ull);
      const int tx = tile % p.ntx, ty = (tile / p.ntx) % p.nty, tz = tile / (p.ntx * p.nty);
      for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
          for (int e = -1; e <= 1; ++e) {
            const int z = tz + a, y = ty + b, x = tx + e;
            if ((a | b | e) == 0 || z < 0 || z >= p.ntz || y < 0 || y >= p.nty || x < 0 || x >= p.ntx) continue;
            p.flag_out[(z * p.nty + y) * p.ntx + x] = 1;
          }
    }
  }
}

// ---- outputs -------------------------------------------------------------------

// flow = sum(max(red0,0)) - sum(max(-ex,0)), exact in int64
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_flow(KP<T> p, const long long* __restrict__ red0,
                                                   unsigned long long* out) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  long long f = 0;
  if (v < p.N) {
    const long long r = red0[v];
    const long long e = (long long)p.ex[v];
    f = (r > 0 ? r : 0) - (e < 0 ? -e : 0);
  }
  for (int o = 32; o > 0; o >>= 1) f += __shfl_down(f, o, 64);
  if ((threadIdx.x & 63) == 0 && f != 0) atomicAdd(out, (unsigned long long)f);
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_cut(KP<T> p, const int* __restrict__ reach, int which,
                                                  int* __restrict__ out) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= p.N) return;
  out[v] = which == LF_CUT_SMAX ? (p.h[v] >= INF_H ? 1 : 0) : reach[v];
}

// ---- energies -> network (NodeFiller::addE1/addE2, network_flow_node.hpp:419-485) ----

// Gather form: every node reads the pairwise terms of the edges it belongs to,
// so no two threads write the same word.  err[0] |= 1 on a non-submodular term.
template <int K>
__global__ __launch_bounds__(NTHREADS) void k_reparam(int D, int H, int W, long long N, DirTab dirs,
                                                      const long long* __restrict__ unary,
                                                      const long long* __restrict__ binary,
                                                      long long* __restrict__ red0,
                                                      long long* __restrict__ alpha0, int* err) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= N) return;
  constexpr int KP2 = K / 2;
  const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((long long)W * H));
  long long r = 2 * (unary[2 * v + 1] - unary[2 * v + 0]);
  bool bad = false;
#pragma unroll
  for (int e = 0; e < KP2; ++e) {
    const int dz = dirs.dz[e], dy = dirs.dy[e], dx = dirs.dx[e];
    // v as node 1 of edge (v, e)
    {
      const int nz = z + dz, ny = y + dy, nx = x + dx;
      if (nz >= 0 && nz < D && ny >= 0 && ny < H && nx >= 0 && nx < W) {
        const long long* E = binary + (v * KP2 + e) * 4;
        const long long e00 = E[0], e01 = E[1], e10 = E[2], e11 = E[3];
        const long long q1 = e10 - e00, q12 = e00 + e11 - e01 - e10;
        bad |= q12 > 0;
        r += 2 * q1 + q12;
        alpha0[(long long)e * N + v] += -q12;
      }
    }
    // v as node 2 of edge (w, e), w = v - dir[e]
    {
      const int nz = z - dz, ny = y - dy, nx = x - dx;
      if (nz >= 0 && nz < D && ny >= 0 && ny < H && nx >= 0 && nx < W) {
        const long long w = ((long long)nz * H + ny) * W + nx;
        const long long* E = binary + (w * KP2 + e) * 4;
        const long long e00 = E[0], e01 = E[1], e10 = E[2], e11 = E[3];
        const long long q2 = e01 - e00, q12 = e00 + e11 - e01 - e10;
        bad |= q12 > 0;
        r += 2 * q2 + q12;
        alpha0[(long long)(e + KP2) * N + v] += -q12;
      }
    }
  }
  red0[v] += r;
  if (bad) atomicOr(err, 1);
}

// working state from the base network.  err[0] |= 2 on a negative arc
// capacity; mx[0] / mx[1] receive max |red0| and max alpha0, which bound every
// excess (|red0| + k * max alpha) and residual (2 * max alpha) of the solve.
template <int K, typename T>
__global__ __launch_bounds__(NTHREADS) void k_load(KP<T> p, const long long* __restrict__ red0,
                                                   const long long* __restrict__ alpha0, int* err,
                                                   unsigned long long* mx) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  long long mr = 0, ma = 0;
  bool bad = false;
  if (v < p.N) {
    const long long r = red0[v];
    mr = r < 0 ? -r : r;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const long long a = alpha0[(long long)i * p.N + v];
      bad |= a < 0;
      ma = a > ma ? a : ma;
      p.res[(long long)i * p.N + v] = (T)a;
      p.mbox[(long long)i * p.N + v] = 0;
    }
    p.ex[v] = (T)(-r);
  }
  for (int o = 32; o > 0; o >>= 1) {
    const long long a = __shfl_down(mr, o, 64), b = __shfl_down(ma, o, 64);
    mr = a > mr ? a : mr;
    ma = b > ma ? b : ma;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&mx[0], (unsigned long long)mr);
    atomicMax(&mx[1], (unsigned long long)ma);
  }
  if (bad) atomicOr(err, 2);
}

template <typename V>
__global__ void k_fill(V* a, V val, long long n) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v < n) a[v] = val;
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + NTHREADS - 1) / NTHREADS); }

// ---- engine -----------------------------------------------------------------------

template <int K, class G, typename T, typename MT>
class Engine final : public MinCutEngine {
 public:
  Engine(const LatticeGeom& geo) : g_(geo) {
    LF_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    N_ = geo.n();
    p_.D = (int)geo.D; p_.H = (int)geo.H; p_.W = (int)geo.W; p_.N = N_;
    p_.ntx = (int)((geo.W + G::TW - 1) / G::TW);
    p_.nty = (int)((geo.H + G::TH - 1) / G::TH);
    p_.ntz = (int)((geo.D + G::TD - 1) / G::TD);
    ntiles_ = (long long)p_.ntx * p_.nty * p_.ntz;
    if (ntiles_ > (1LL << 30)) fail(LF_ERR_BAD_ARG, "lattice too large");
    p_.ntiles = (int)ntiles_;
    p_.nblocks = (int)(((ntiles_ + 7) / 8) * 8);
    for (int i = 0; i < K; ++i) {
      p_.dirs.dz[i] = (int8_t)geo.dz[i];
      p_.dirs.dy[i] = (int8_t)geo.dy[i];
      p_.dirs.dx[i] = (int8_t)geo.dx[i];
    }
    LF_HIP(hipMalloc(&red0_, sizeof(long long) * N_));
    LF_HIP(hipMalloc(&alpha0_, sizeof(long long) * N_ * K));
    LF_HIP(hipMalloc(&res_, sizeof(T) * N_ * K));
    LF_HIP(hipMalloc(&mbox_, sizeof(T) * N_ * K));
    LF_HIP(hipMalloc(&ex_, sizeof(T) * N_));
    LF_HIP(hipMalloc(&h_, sizeof(int) * N_));
    LF_HIP(hipMalloc(&mask_, sizeof(MT) * N_));
    LF_HIP(hipMalloc(&reach_, sizeof(int) * N_));
    LF_HIP(hipMalloc(&flags_[0], sizeof(int) * ntiles_));
    LF_HIP(hipMalloc(&flags_[1], sizeof(int) * ntiles_));
    LF_HIP(hipMalloc(&dev_small_, 64));
    LF_HIP(hipHostMalloc(&host_small_, 64, hipHostMallocDefault));
    reset_energies();
  }
  ~Engine() override {
    hipStreamSynchronize(s_);
    hipFree(red0_); hipFree(alpha0_); hipFree(res_); hipFree(mbox_); hipFree(ex_); hipFree(h_);
    hipFree(mask_); hipFree(reach_); hipFree(flags_[0]); hipFree(flags_[1]); hipFree(dev_small_);
    hipHostFree(host_small_);
    hipStreamDestroy(s_);
  }

  hipStream_t stream() const override { return s_; }

  void reset_energies() override {
    LF_HIP(hipMemsetAsync(red0_, 0, sizeof(long long) * N_, s_));
    LF_HIP(hipMemsetAsync(alpha0_, 0, sizeof(long long) * N_ * K, s_));
    solved = false;
  }

  void add_energies_device(const int64_t* du, const int64_t* db, hipStream_t s) override {
    if (s && s != s_) {
      hipEvent_t ev;
      LF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      LF_HIP(hipEventRecord(ev, s));
      LF_HIP(hipStreamWaitEvent(s_, ev, 0));
      LF_HIP(hipEventDestroy(ev));
    }
    int* err = reinterpret_cast<int*>(dev_small_);
    LF_HIP(hipMemsetAsync(err, 0, sizeof(int), s_));
    hipLaunchKernelGGL((k_reparam<K>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, p_.D, p_.H, p_.W, N_,
                       p_.dirs, reinterpret_cast<const long long*>(du), reinterpret_cast<const long long*>(db),
                       red0_, alpha0_, err);
    LF_HIP(hipGetLastError());
    LF_HIP(hipMemcpyAsync(host_small_, err, sizeof(int), hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    if (reinterpret_cast<int*>(host_small_)[0] & 1)
      fail(LF_ERR_NONSUBMODULAR, "pairwise term with E00 + E11 > E01 + E10 (not submodular)");
    solved = false;
  }

  void run(const lf_tuning& tun) override {
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    stats = lf_stats{};
    KP<T> p = params();
    p.iters = tun.local_iters > 0 ? tun.local_iters : 8;
    const int R = tun.rounds_per_relabel > 0 ? tun.rounds_per_relabel : 16;
    int* err = reinterpret_cast<int*>(dev_small_);
    unsigned long long* mx = reinterpret_cast<unsigned long long*>(dev_small_ + 32);
    LF_HIP(hipMemsetAsync(dev_small_, 0, 64, s_));
    hipLaunchKernelGGL((k_load<K, T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, p, red0_, alpha0_, err, mx);
    LF_HIP(hipGetLastError());
    LF_HIP(hipMemcpyAsync(host_small_, dev_small_, 64, hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    {
      const int e = reinterpret_cast<int*>(host_small_)[0];
      const unsigned long long* m = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(host_small_) + 32);
      if (e & 2) fail(LF_ERR_NONSUBMODULAR, "negative pairwise capacity");
      const double bound = (double)m[0] + (double)K * (double)m[1] + 2.0 * (double)m[1];
      const double lim = sizeof(T) == 4 ? 2147483647.0 : 9.2e18;
      if (bound >= lim)
        fail(LF_ERR_OVERFLOW, sizeof(T) == 4 ? "capacities exceed the int32 range; use cap_bytes=8"
                                             : "capacities exceed the int64 range");
    }
    double t_push = 0, t_bfs = 0;
    bool pushed = false;
    for (;;) {
      const auto tb = clk::now();
      const long long active = global_relabel(p, pushed);
      t_bfs += std::chrono::duration<double>(clk::now() - tb).count();
      ++stats.cycles;
      if (active == 0) break;
      if (tun.max_cycles > 0 && stats.cycles > tun.max_cycles)
        fail(LF_ERR_STATE, "max_cycles reached before the flow became maximum");
      const auto tp = clk::now();
      for (int r = 0; r < R; ++r) {
        KP<T> q = p;
        q.flag_in = flags_[cur_];
        q.flag_out = flags_[cur_ ^ 1];
        hipLaunchKernelGGL((k_push<K, G, T>), dim3(p_.nblocks), dim3(NTHREADS), 0, s_, q);
        cur_ ^= 1;
        ++stats.rounds;
      }
      LF_HIP(hipGetLastError());
      pushed = true;
      LF_HIP(hipStreamSynchronize(s_));
      t_push += std::chrono::duration<double>(clk::now() - tp).count();
    }
    // S_min flood
    hipLaunchKernelGGL((k_flood_init<T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, p, reach_);
    fill_flags(flags_[cur_], 1);
    fill_flags(flags_[cur_ ^ 1], 0);
    for (;;) {
      zero_counter();
      KP<T> q = p;
      q.flag_in = flags_[cur_];
      q.flag_out = flags_[cur_ ^ 1];
      hipLaunchKernelGGL((k_flood_tile<K, G, T, MT>), dim3(p_.nblocks), dim3(NTHREADS), 0, s_, q, mask_, reach_);
      cur_ ^= 1;
      ++stats.flood_launches;
      if (read_counter() == 0) break;
    }
    // flow value
    zero_counter();
    hipLaunchKernelGGL((k_flow<T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, p, red0_, counter_ptr());
    flow = (long long)read_counter();
    LF_HIP(hipStreamSynchronize(s_));
    stats.push_seconds = t_push;
    stats.bfs_seconds = t_bfs;
    stats.seconds = std::chrono::duration<double>(clk::now() - t0).count();
    solved = true;
  }

  void cut_device(int32_t* d_out, int which, hipStream_t s) override {
    KP<T> p = params();
    hipLaunchKernelGGL((k_cut<T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, p, reach_, which, d_out);
    LF_HIP(hipGetLastError());
    if (s && s != s_) {
      hipEvent_t ev;
      LF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      LF_HIP(hipEventRecord(ev, s_));
      LF_HIP(hipStreamWaitEvent(s, ev, 0));
      LF_HIP(hipEventDestroy(ev));
    }
  }

 private:
  KP<T> params() const {
    KP<T> p = p_;
    p.res = res_; p.ex = ex_; p.h = h_; p.mbox = mbox_;
    p.counter = counter_ptr();
    p.flag_in = flags_[cur_];
    p.flag_out = flags_[cur_ ^ 1];
    return p;
  }
  unsigned long long* counter_ptr() const { return reinterpret_cast<unsigned long long*>(dev_small_ + 16); }
  void zero_counter() {
    hipLaunchKernelGGL((k_fill<unsigned long long>), dim3(1), dim3(NTHREADS), 0, s_, counter_ptr(), 0ull, 1LL);
  }
  unsigned long long read_counter() {
    LF_HIP(hipMemcpyAsync(host_small_, counter_ptr(), sizeof(unsigned long long), hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    return *reinterpret_cast<unsigned long long*>(host_small_);
  }
  void fill_flags(int* f, int v) {
    hipLaunchKernelGGL((k_fill<int>), dim3(blocks_for(ntiles_)), dim3(NTHREADS), 0, s_, f, v, ntiles_);
  }

  // deliver mailboxes, exact BFS labels, flag active tiles; returns #active nodes
  long long global_relabel(KP<T> p, bool pushed) {
    if (pushed)
      hipLaunchKernelGGL((k_deliver<K, G, T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, p);
    hipLaunchKernelGGL((k_bfs_init<K, T, MT>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, p, mask_);
    fill_flags(flags_[cur_], 1);
    fill_flags(flags_[cur_ ^ 1], 0);
    for (;;) {
      zero_counter();
      KP<T> q = p;
      q.flag_in = flags_[cur_];
      q.flag_out = flags_[cur_ ^ 1];
      hipLaunchKernelGGL((k_bfs_tile<K, G, T, MT>), dim3(p_.nblocks), dim3(NTHREADS), 0, s_, q, mask_);
      cur_ ^= 1;
      ++stats.bfs_launches;
      if (read_counter() == 0) break;
    }
    // push rounds start with exactly the tiles holding active nodes
    fill_flags(flags_[cur_], 0);
    fill_flags(flags_[cur_ ^ 1], 0);
    zero_counter();
    KP<T> q = p;
    q.flag_out = flags_[cur_];
    hipLaunchKernelGGL((k_count_active<G, T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, q);
    LF_HIP(hipGetLastError());
    return (long long)read_counter();
  }

  LatticeGeom g_;
  hipStream_t s_ = nullptr;
  long long N_ = 0, ntiles_ = 0;
  KP<T> p_{};
  long long* red0_ = nullptr;
  long long* alpha0_ = nullptr;
  T* res_ = nullptr;
  T* mbox_ = nullptr;
  T* ex_ = nullptr;
  int* h_ = nullptr;
  MT* mask_ = nullptr;
  int* reach_ = nullptr;
  int* flags_[2] = {nullptr, nullptr};
  int cur_ = 0;
  char* dev_small_ = nullptr;
  void* host_small_ = nullptr;
};

template <int K, class G>
std::unique_ptr<MinCutEngine> make_k(const LatticeGeom& g, int cap_bytes) {
  using MT = typename std::conditional<(K <= 8), unsigned char,
             typename std::conditional<(K <= 16), unsigned short, unsigned int>::type>::type;
  if (cap_bytes == 4) return std::unique_ptr<MinCutEngine>(new Engine<K, G, int, MT>(g));
  return std::unique_ptr<MinCutEngine>(new Engine<K, G, long long, MT>(g));
}

int max_abs(const std::vector<int>& v) {
  int m = 0;
  for (int x : v) m = std::max(m, std::abs(x));
  return m;
}

}  // namespace

std::unique_ptr<MinCutEngine> make_mincut_engine(const LatticeGeom& g, int cap_bytes) {
  if (cap_bytes != 4 && cap_bytes != 8) fail(LF_ERR_BAD_ARG, "cap_bytes must be 4 or 8");
  const int rz = max_abs(g.dz), ry = max_abs(g.dy), rx = max_abs(g.dx);
  const bool is3d = g.ndim == 3;
  if (!is3d && rz == 0 && ry <= 1 && rx <= 1) {
    using G = Geo<1, 16, 64, 0, 1, 1>;
    if (g.k == 2) return make_k<2, G>(g, cap_bytes);
    if (g.k == 4) return make_k<4, G>(g, cap_bytes);
    if (g.k == 8) return make_k<8, G>(g, cap_bytes);
  }
  if (!is3d && rz == 0 && ry <= 2 && rx <= 2) {
    using G = Geo<1, 16, 64, 0, 2, 2>;
    if (g.k == 12) return make_k<12, G>(g, cap_bytes);
    if (g.k == 16) return make_k<16, G>(g, cap_bytes);
  }
  if (is3d && rz <= 1 && ry <= 1 && rx <= 1) {
    if (g.k == 6) return make_k<6, Geo<4, 8, 32, 1, 1, 1>>(g, cap_bytes);
    if (g.k == 18) return make_k<18, Geo<4, 4, 16, 1, 1, 1>>(g, cap_bytes);
    if (g.k == 26) return make_k<26, Geo<4, 4, 16, 1, 1, 1>>(g, cap_bytes);
  }
  fail(LF_ERR_UNSUPPORTED, "no device path for this kernel size / radius yet");
}

}  // namespace lfamd
