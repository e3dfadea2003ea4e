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
#include <cmath>
#include <initializer_list>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "lf_engine.hpp"
#include "lf_verify.hpp"

namespace lfamd {
namespace {

constexpr int NTHREADS = 256;
// directions whose LDS operands the push round reads in one batch
constexpr int HOIST = 8;
constexpr int INF_H = 1 << 30;
constexpr int MAXK = 48;

struct DirTab {
  int8_t dz[MAXK], dy[MAXK], dx[MAXK];
};

// Direction tables of the hot kernels.  RtDir reads the engine's table (a
// kernel argument).  CDir<K, ND> is the lattice edge order of the radius-1
// neighbourhoods (2-D: 4 and 8 directions, 3-D: 6 and 26) known at compile
// time: the lexicographically positive offsets (dz, dy, dx) in ascending order,
// then their negations in the same order.  With the unrolled direction loops
// every offset, LDS address and tile-border test becomes a constant.
struct RtDir {
  __device__ static int z(const DirTab& t, int i) { return t.dz[i]; }
  __device__ static int y(const DirTab& t, int i) { return t.dy[i]; }
  __device__ static int x(const DirTab& t, int i) { return t.dx[i]; }
};

template <int K, int ND>
__host__ __device__ constexpr bool cdir_member(int l1) {
  return ND == 2 ? (K == 8 || (K == 4 && l1 == 1)) : (K == 26 || (K == 6 && l1 == 1));
}

template <int K, int ND>
__host__ __device__ constexpr int cdir(int i, int axis) {
  const int e = i < K / 2 ? i : i - K / 2;
  const int sg = i < K / 2 ? 1 : -1;
  int c = 0;
  for (int z = (ND == 3 ? -1 : 0); z <= (ND == 3 ? 1 : 0); ++z)
    for (int y = -1; y <= 1; ++y)
      for (int x = -1; x <= 1; ++x) {
        if (!(z > 0 || (z == 0 && (y > 0 || (y == 0 && x > 0))))) continue;
        const int l1 = (z < 0 ? -z : z) + (y < 0 ? -y : y) + (x < 0 ? -x : x);
        if (!cdir_member<K, ND>(l1)) continue;
        if (c == e) return sg * (axis == 0 ? z : (axis == 1 ? y : x));
        ++c;
      }
  return 0;
}

template <int K, int ND>
struct CDir {
  __device__ static constexpr int z(const DirTab&, int i) { return cdir<K, ND>(i, 0); }
  __device__ static constexpr int y(const DirTab&, int i) { return cdir<K, ND>(i, 1); }
  __device__ static constexpr int x(const DirTab&, int i) { return cdir<K, ND>(i, 2); }
  // true when the engine's table is this order
  static bool matches(const int* dz, const int* dy, const int* dx, int k) {
    if (k != K) return false;
    for (int i = 0; i < K; ++i)
      if (dz[i] != cdir<K, ND>(i, 0) || dy[i] != cdir<K, ND>(i, 1) || dx[i] != cdir<K, ND>(i, 2)) return false;
    return true;
  }
};
static_assert(cdir<8, 2>(0, 2) == 1 && cdir<8, 2>(1, 1) == 1 && cdir<8, 2>(1, 2) == -1 && cdir<8, 2>(4, 2) == -1,
              "2-D 8-neighbourhood order");
static_assert(cdir<26, 3>(4, 0) == 1 && cdir<26, 3>(4, 1) == -1 && cdir<26, 3>(25, 0) == -1, "3-D 26 order");

template <int TD_, int TH_, int TW_, int RZ_, int RY_, int RX_>
struct Geo {
  static constexpr int TD = TD_, TH = TH_, TW = TW_, RZ = RZ_, RY = RY_, RX = RX_;
  static constexpr int NT = TD * TH * TW;
  static constexpr int HD = TD + 2 * RZ, HH = TH + 2 * RY, HW = TW + 2 * RX;
  static constexpr int NH = HD * HH * HW;
  static constexpr int NPT = NT / NTHREADS;
  static_assert(NT % NTHREADS == 0, "tile must be a multiple of the block");
};

// 1 bit per direction: the residual-arc mask of a node
template <int K>
using MaskT = typename std::conditional<(K <= 8), unsigned char,
              typename std::conditional<(K <= 16), unsigned short,
              typename std::conditional<(K <= 32), unsigned int, unsigned long long>::type>::type>::type;

template <typename T>
struct KP {
  T* res;
  T* ex;
  int* h;
  // Mailboxes, one per round parity.  mbox[i][v]: flow that node v pushed along
  // direction i into a node of another tile in a round, not yet applied there.
  // A round adds what its tiles send to mbox (no other round writes it
  // meanwhile) and drains mbox_in, which the round before filled and nobody
  // writes during this round, with plain loads and stores: every tile that
  // received flow is on the next round's list, so after a round the only
  // undelivered flow is in the mailbox that round filled.
  T* mbox;
  T* mbox_in;
  // mail[t] bit 9 (sz+1) + 3 (sy+1) + (sx+1): the tile at offset (sz,sy,sx)
  // from t put flow into t's mailbox slots in this round (mail_in: in the
  // round before, read and cleared by t's block)
  unsigned* mail;
  unsigned* mail_in;
  // dmail[t]: the same side bits, kept until the delivery pass, which reads
  // only the slots of those sides and then clears them
  unsigned* dmail;
  // MaskT<K> per node, bit i = (res[i][v] > 0), kept current by every kernel
  // that changes residuals so that a relabel need not rebuild it
  void* mask;
  int* flag_in;
  int* flag_out;
  unsigned long long* counter;
  unsigned long long* tiles;  // processed-tile counter of the push rounds
  // push work lists: the round takes its tiles from wl_in (wl_cnt_in of
  // them) and appends the tiles of the next round to wl_out; wl_cnt_zero is
  // the counter the round after next appends to, cleared here
  const int* wl_in;
  int* wl_out;
  const unsigned* wl_cnt_in;
  unsigned* wl_cnt_out;
  unsigned* wl_cnt_zero;
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

// Buffer addressing: the base of an array (a residual plane, or a plane
// shifted by a neighbour offset) lives in scalar registers as a buffer
// resource, and a node's 32-bit byte offset in one vector register serves
// every plane.  With 64-bit flat addresses every load in flight held its own
// register pair.  Engines using it need N * sizeof(element) < 2^32.
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void* base) {
  // stride 0, the whole 32-bit offset range, dword data format (gfx9 word 3)
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, -1, 0x00020000);
}
template <typename V>
__device__ inline V bld(const void* base, unsigned off) {
  static_assert(sizeof(V) == 1 || sizeof(V) == 2 || sizeof(V) == 4 || sizeof(V) == 8, "element size");
  if constexpr (sizeof(V) == 1) {
    return (V)__builtin_amdgcn_raw_buffer_load_b8(buf_rsrc(base), (int)off, 0, 0);
  } else if constexpr (sizeof(V) == 2) {
    return (V)__builtin_amdgcn_raw_buffer_load_b16(buf_rsrc(base), (int)off, 0, 0);
  } else if constexpr (sizeof(V) == 4) {
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(buf_rsrc(base), (int)off, 0, 0));
  } else {
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(buf_rsrc(base), (int)off, 0, 0));
  }
}
template <typename V>
__device__ inline void bst(void* base, unsigned off, V v) {
  static_assert(sizeof(V) == 1 || sizeof(V) == 2 || sizeof(V) == 4 || sizeof(V) == 8, "element size");
  if constexpr (sizeof(V) == 1) {
    __builtin_amdgcn_raw_buffer_store_b8((unsigned char)v, buf_rsrc(base), (int)off, 0, 0);
  } else if constexpr (sizeof(V) == 2) {
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, buf_rsrc(base), (int)off, 0, 0);
  } else if constexpr (sizeof(V) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), buf_rsrc(base), (int)off, 0, 0);
  } else {
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), buf_rsrc(base), (int)off, 0, 0);
  }
}

// Streaming kernels that reduce run a fixed grid of RED_BLOCKS blocks with a
// grid-stride loop and leave one partial per block; k_reduce folds them.  No
// two blocks ever update the same word.
constexpr int RED_BLOCKS = 2048;

template <bool IS_MAX>
__device__ inline long long block_reduce(long long v) {
  __shared__ long long s_part[NTHREADS / 64];
  for (int o = 32; o > 0; o >>= 1) {
    const long long w = __shfl_down(v, o, 64);
    v = IS_MAX ? (w > v ? w : v) : v + w;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) s_part[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = s_part[0];
    for (int i = 1; i < NTHREADS / 64; ++i) v = IS_MAX ? (s_part[i] > v ? s_part[i] : v) : v + s_part[i];
  }
  __syncthreads();
  return v;  // valid in thread 0
}

// out[j] = fold over b < nparts of part[b * stride + j], one block
template <bool IS_MAX>
__global__ __launch_bounds__(NTHREADS) void k_reduce(const long long* __restrict__ part, int nparts, int stride,
                                                     int j, unsigned long long* out) {
  long long v = 0;
  for (int b = threadIdx.x; b < nparts; b += NTHREADS) {
    const long long w = part[(long long)b * stride + j];
    v = IS_MAX ? (w > v ? w : v) : v + w;
  }
  v = block_reduce<IS_MAX>(v);
  if (threadIdx.x == 0) out[0] = (unsigned long long)v;
}

// Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous run
// of tiles so that halo reads of neighbouring tiles hit the same L2.
__device__ inline int tile_of_block(int nblocks) {
  const int per = nblocks >> 3;
  const int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}

// Flag the neighbouring tile that holds neighbour i of a node at global (gz,gy,gx).
template <class G, typename T>
__device__ inline void flag_neighbor_tile(const KP<T>& p, int gz, int gy, int gx, int* flags) {
  const int tz = gz / G::TD, ty = gy / G::TH, tx = gx / G::TW;
  flags[(tz * p.nty + ty) * p.ntx + tx] = 1;
}

// ---- the push / relabel round -----------------------------------------------

// Local coordinates of tile slot t.
template <class G>
__device__ inline void tile_coords(int t, int& z, int& y, int& x) {
  z = t / (G::TH * G::TW);
  y = (t / G::TW) % G::TH;
  x = t % G::TW;
}

// One push / relabel round over the tiles of the work list; a fixed grid of
// resident blocks walks the list, so a round over few active tiles costs
// little more than those tiles.
// Narrow int32 networks run 5 waves per SIMD (at most 96 VGPRs, a few spilled
// registers): 2048^2 segmentation at 8 local iterations per round 147 -> 134 us
// per round against 4 waves.
template <int K, class G, typename T, class DIR>
__global__ __launch_bounds__(NTHREADS, (K <= 8 && sizeof(T) == 4) ? 5 : 1) void k_push(KP<T> p) {
  constexpr int NT = G::NT;
  __shared__ T s_res[K * NT];
  __shared__ T s_ex[NT];
  __shared__ int s_h[G::NH];
  __shared__ unsigned short s_list[NT];
  __shared__ int s_cnt[2];
  __shared__ int s_tile;
  __shared__ unsigned int s_nbr;  // bit 9 (sz+1) + 3 (sy+1) + (sx+1): neighbouring tile received flow
  __shared__ unsigned int s_mail;  // sides of this tile whose neighbours sent flow (p.mail)

  if (blockIdx.x == 0 && threadIdx.x == 0) *p.wl_cnt_zero = 0;
  const int ntodo = (int)*p.wl_cnt_in;
  // thread 0 loads the id of its block's next tile, and then that tile's mail
  // bits (final since the round before), while the current one is processed
  int next_tile = (threadIdx.x == 0 && (int)blockIdx.x < ntodo) ? p.wl_in[blockIdx.x] : -1;
  unsigned next_mail = next_tile >= 0 ? p.mail_in[next_tile] : 0u;
  unsigned long long ntiles_done = 0;
  for (int li = blockIdx.x; li < ntodo; li += gridDim.x) {
    if (threadIdx.x == 0) {
      const int t = next_tile;
      next_tile = li + (int)gridDim.x < ntodo ? p.wl_in[li + gridDim.x] : -1;
      s_tile = t;
      p.flag_in[t] = 0;
      s_cnt[0] = 0;
      s_cnt[1] = 0;
      s_nbr = 0;
      s_mail = next_mail;
      if (next_mail) p.mail_in[t] = 0u;
      ++ntiles_done;
    }
    __syncthreads();
    const int tile = s_tile;

    const int tx = tile % p.ntx, ty = (tile / p.ntx) % p.nty, tz = tile / (p.ntx * p.nty);
    const int x0 = tx * G::TW, y0 = ty * G::TH, z0 = tz * G::TD;
    const long long N = p.N;
    // lattice extent of this tile (nodes past the lattice edge are never active)
    const int lw = min(G::TW, p.W - x0), lh = min(G::TH, p.H - y0), ld = min(G::TD, p.D - z0);

    // stage residuals and excess of the tile.  Slots past the lattice edge
    // load node 0 and discard it: no load waits behind a branch, so a
    // thread's whole batch of loads is in flight at once.  Every plane is
    // addressed by the node's byte offset (buffer loads).
    {
      T r[G::NPT][K + 1];
      bool in[G::NPT];
      unsigned off[G::NPT];
      // labels of the tile and its halo, loaded in the same batch
      constexpr int NHI = (G::NH + NTHREADS - 1) / NTHREADS;
      int hv[NHI];
      bool hin[NHI];
#pragma unroll
      for (int k = 0; k < NHI; ++k) {
        const int q = threadIdx.x + k * NTHREADS;
        const int hz = q / (G::HH * G::HW), hy = (q / G::HW) % G::HH, hx = q % G::HW;
        const int gz = z0 + hz - G::RZ, gy = y0 + hy - G::RY, gx = x0 + hx - G::RX;
        hin[k] = q < G::NH && gz >= 0 && gz < p.D && gy >= 0 && gy < p.H && gx >= 0 && gx < p.W;
        hv[k] = bld<int>(p.h, hin[k] ? (unsigned)(((long long)gz * p.H + gy) * p.W + gx) * 4u : 0u);
      }
#pragma unroll
      for (int j = 0; j < G::NPT; ++j) {
        int z, y, x;
        tile_coords<G>(threadIdx.x + j * NTHREADS, z, y, x);
        in[j] = z < ld && y < lh && x < lw;
        const unsigned g = in[j] ? (unsigned)(((long long)(z0 + z) * p.H + (y0 + y)) * p.W + (x0 + x)) : 0u;
        off[j] = g * (unsigned)sizeof(T);
#pragma unroll
        for (int i = 0; i < K; ++i) r[j][i] = bld<T>(p.res + (long long)i * N, off[j]);
        r[j][K] = bld<T>(p.ex, off[j]);
      }
      // kernels of at most HOIST directions read their mailbox slots (the
      // sides the mail bits name) in the same batch as the staging loads
      T a[G::NPT][K <= HOIST ? K : 1];
      if constexpr (K <= HOIST) {
        const unsigned mail = s_mail;
#pragma unroll
        for (int j = 0; j < G::NPT; ++j) {
          int z, y, x;
          tile_coords<G>(threadIdx.x + j * NTHREADS, z, y, x);
          const int gz = z0 + z, gy = y0 + y, gx = x0 + x;
#pragma unroll
          for (int i = 0; i < K; ++i) {
            const int dz = DIR::z(p.dirs, i), dy = DIR::y(p.dirs, i), dx = DIR::x(p.dirs, i);
            const int lz = z + dz, ly = y + dy, lx = x + dx;
            const int nz = gz + dz, ny = gy + dy, nx = gx + dx;
            const int side = (lz < 0 ? 0 : (lz >= G::TD ? 2 : 1)) * 9 + (ly < 0 ? 0 : (ly >= G::TH ? 2 : 1)) * 3 +
                             (lx < 0 ? 0 : (lx >= G::TW ? 2 : 1));
            a[j][i] = 0;
            // the neighbour's slot of the reverse direction: the plane shifted
            // by the neighbour offset, at this node's byte offset
            if (side != 13 && ((mail >> side) & 1u) && in[j] && nz >= 0 && nz < p.D && ny >= 0 && ny < p.H &&
                nx >= 0 && nx < p.W)
              a[j][i] = bld<T>(p.mbox_in + (long long)((i + K / 2) % K) * N + ((long long)dz * p.H + dy) * p.W + dx,
                               off[j]);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < NHI; ++k) {
        const int q = threadIdx.x + k * NTHREADS;
        if (q < G::NH) s_h[q] = hin[k] ? hv[k] : INF_H;
      }
#pragma unroll
      for (int j = 0; j < G::NPT; ++j) {
        const int t = threadIdx.x + j * NTHREADS;
        T add = 0;
#pragma unroll
        for (int i = 0; i < K; ++i) {
          T v = in[j] ? r[j][i] : (T)0;
          if constexpr (K <= HOIST) {
            v += a[j][i];
            add += a[j][i];
            // a slot read with flow is cleared for the round after next
            if (a[j][i] != 0)
              bst<T>(p.mbox_in + (long long)((i + K / 2) % K) * N +
                         ((long long)DIR::z(p.dirs, i) * p.H + DIR::y(p.dirs, i)) * p.W + DIR::x(p.dirs, i),
                     off[j], (T)0);
          }
          s_res[i * NT + t] = v;
        }
        s_ex[t] = (in[j] ? r[j][K] : (T)0) + add;
      }
    }
    // collect the flow that neighbouring tiles pushed into this tile in the
    // round before, from the sides the mail bits name, all loads of a thread
    // in flight at once; slots read with flow are cleared
    if (K > HOIST && s_mail) {
      const unsigned mail = s_mail;
      T add_e[G::NPT];
#pragma unroll
      for (int j = 0; j < G::NPT; ++j) add_e[j] = 0;
#pragma unroll
      for (int c = 0; c < K; c += HOIST) {
        T a[G::NPT][HOIST];
#pragma unroll
        for (int j = 0; j < G::NPT; ++j) {
          int z, y, x;
          tile_coords<G>(threadIdx.x + j * NTHREADS, z, y, x);
          const int gz = z0 + z, gy = y0 + y, gx = x0 + x;
#pragma unroll
          for (int u = 0; u < HOIST && c + u < K; ++u) {
            const int i = c + u;
            const int dz = DIR::z(p.dirs, i), dy = DIR::y(p.dirs, i), dx = DIR::x(p.dirs, i);
            const int lz = z + dz, ly = y + dy, lx = x + dx;
            const int nz = gz + dz, ny = gy + dy, nx = gx + dx;
            const int side = (lz < 0 ? 0 : (lz >= G::TD ? 2 : 1)) * 9 + (ly < 0 ? 0 : (ly >= G::TH ? 2 : 1)) * 3 +
                             (lx < 0 ? 0 : (lx >= G::TW ? 2 : 1));
            a[j][u] = 0;
            const unsigned o = (unsigned)(((long long)gz * p.H + gy) * p.W + gx) * (unsigned)sizeof(T);
            if (side != 13 && ((mail >> side) & 1u) && z < ld && y < lh && x < lw && nz >= 0 && nz < p.D &&
                ny >= 0 && ny < p.H && nx >= 0 && nx < p.W)
              a[j][u] = bld<T>(p.mbox_in + (long long)((i + K / 2) % K) * N + ((long long)dz * p.H + dy) * p.W + dx, o);
          }
        }
#pragma unroll
        for (int j = 0; j < G::NPT; ++j) {
          const int t = threadIdx.x + j * NTHREADS;
          int z, y, x;
          tile_coords<G>(t, z, y, x);
#pragma unroll
          for (int u = 0; u < HOIST && c + u < K; ++u) {
            if (a[j][u] == 0) continue;
            const int i = c + u;
            add_e[j] += a[j][u];
            s_res[i * NT + t] += a[j][u];  // only this thread touches node t's planes while staging
            const unsigned o = (unsigned)(((long long)(z0 + z) * p.H + (y0 + y)) * p.W + (x0 + x)) * (unsigned)sizeof(T);
            bst<T>(p.mbox_in + (long long)((i + K / 2) % K) * N +
                       ((long long)DIR::z(p.dirs, i) * p.H + DIR::y(p.dirs, i)) * p.W + DIR::x(p.dirs, i),
                   o, (T)0);
          }
        }
      }
#pragma unroll
      for (int j = 0; j < G::NPT; ++j) s_ex[threadIdx.x + j * NTHREADS] += add_e[j];
    }
    __syncthreads();
    if (threadIdx.x == 0 && next_tile >= 0) next_mail = p.mail_in[next_tile];

    // Local iterations: relabel + list, then push.  The pass over the owned
    // nodes relabels every node with excess and no downhill residual arc
    // (residuals frozen: the push phase before it ended at a barrier; labels
    // only grow, so concurrent reads of a neighbour's label are safe) and lists
    // the nodes that are then active; the push phase walks the list with
    // labels frozen.  Two barriers per iteration.  Nodes left with excess after
    // the last push phase are relabelled by the first pass of their tile's next
    // round (or by the next global relabel).
    for (int it = 0; it < p.iters; ++it) {
      int* cnt = &s_cnt[it & 1];
#pragma unroll
      for (int j = 0; j < G::NPT; ++j) {
        const int t = threadIdx.x + j * NTHREADS;
        const T e = s_ex[t];
        int z, y, x;
        tile_coords<G>(t, z, y, x);
        const int hi = ((z + G::RZ) * G::HH + (y + G::RY)) * G::HW + (x + G::RX);
        int hv = s_h[hi];
        if (e > 0 && hv < INF_H) {
          int m = INF_H;
#pragma unroll
          for (int c = 0; c < K; c += HOIST) {
            int hn[HOIST];
            T rr[HOIST];
#pragma unroll
            for (int u = 0; u < HOIST && c + u < K; ++u) {
              const int i = c + u;
              hn[u] = s_h[hi + (DIR::z(p.dirs, i) * G::HH + DIR::y(p.dirs, i)) * G::HW + DIR::x(p.dirs, i)];
              rr[u] = s_res[i * NT + t];
            }
#pragma unroll
            for (int u = 0; u < HOIST && c + u < K; ++u) m = (rr[u] > 0 && hn[u] < m) ? hn[u] : m;
          }
          if (m >= hv) {
            hv = (m >= INF_H - 1) ? INF_H : m + 1;
            s_h[hi] = hv;
          }
        }
        // list the active nodes (ex > 0, label < INF) with one LDS atomic per wave
        const bool act = e > 0 && hv < INF_H;
        const unsigned long long b = __ballot(act);
        const int lane = threadIdx.x & 63;
        int base = 0;
        if (lane == 0 && b) base = atomicAdd(cnt, __popcll(b));
        base = __shfl(base, 0, 64);
        if (act) s_list[base + __popcll(b & ((1ull << lane) - 1))] = (unsigned short)t;
      }
      __syncthreads();
      const int n = *cnt;
      if (n == 0) break;
      if (threadIdx.x == 0) s_cnt[(it + 1) & 1] = 0;

      // push along downhill residual arcs, labels frozen
      for (int q = threadIdx.x; q < n; q += NTHREADS) {
        const int t = s_list[q];
        int z, y, x;
        tile_coords<G>(t, z, y, x);
        const int hi = ((z + G::RZ) * G::HH + (y + G::RY)) * G::HW + (x + G::RX);
        const int hv = s_h[hi];
        // excess moves only through LDS atomics in this phase: in-tile receivers
        // add, the owner subtracts what it pushed
        const T e0 = s_ex[t];
        // the labels and residuals the node looks at, read in batches of HOIST
        // directions (labels are frozen in this phase; only this thread lowers
        // node t's residuals, and a receiver's increase lands on an arc that t
        // does not push along, since both ends cannot be lower than the other)
        T e = e0;
#pragma unroll
        for (int c = 0; c < K; c += HOIST) {
          int hn[HOIST];
          T rr[HOIST];
#pragma unroll
          for (int u = 0; u < HOIST && c + u < K; ++u) {
            const int i = c + u;
            hn[u] = s_h[hi + (DIR::z(p.dirs, i) * G::HH + DIR::y(p.dirs, i)) * G::HW + DIR::x(p.dirs, i)];
            rr[u] = s_res[i * NT + t];
          }
#pragma unroll
          for (int u = 0; u < HOIST && c + u < K; ++u) {
            const int i = c + u;
            const int dz = DIR::z(p.dirs, i), dy = DIR::y(p.dirs, i), dx = DIR::x(p.dirs, i);
            if (e > 0 && hn[u] < hv) {
              const T r = rr[u];
              if (r > 0) {
                const T d = r < e ? r : e;
                s_res[i * NT + t] = r - d;
                e -= d;
                const int lz = z + dz, ly = y + dy, lx = x + dx;
                if (lz >= 0 && lz < G::TD && ly >= 0 && ly < G::TH && lx >= 0 && lx < G::TW) {
                  const int tn = t + (dz * G::TH + dy) * G::TW + dx;
                  lds_add(&s_res[((i + K / 2) % K) * NT + tn], d);
                  lds_add(&s_ex[tn], d);
                } else {
                  const long long g = ((long long)(z0 + z) * p.H + (y0 + y)) * p.W + (x0 + x);
                  gadd(&p.mbox[(long long)i * N + g], d);
                  const int sz = lz < 0 ? 0 : (lz >= G::TD ? 2 : 1), sy = ly < 0 ? 0 : (ly >= G::TH ? 2 : 1),
                            sx = lx < 0 ? 0 : (lx >= G::TW ? 2 : 1);
                  atomicOr(&s_nbr, 1u << (sz * 9 + sy * 3 + sx));
                }
              }
            }
          }
        }
        if (e != e0) lds_add(&s_ex[t], e - e0);
      }
      __syncthreads();
    }

    // the tile stays listed while it holds active nodes (nodes past the
    // lattice edge hold no excess)
    bool act = false;
#pragma unroll
    for (int j = 0; j < G::NPT; ++j) {
      const int t = threadIdx.x + j * NTHREADS;
      int z, y, x;
      tile_coords<G>(t, z, y, x);
      act |= s_ex[t] > 0 && s_h[((z + G::RZ) * G::HH + (y + G::RY)) * G::HW + (x + G::RX)] < INF_H;
    }
    const unsigned nbr_sides = s_nbr;  // final: the last push phase ended at a barrier
    const bool self_active = __syncthreads_or(act);
    // next round: this tile while it holds active nodes, and every tile whose
    // mailbox it filled.  Lane b of the last wave handles side b, so the
    // appends of all sides are in flight together (one after the other they
    // were a chain of dependent atomics at the end of every tile), and the
    // list counter takes one atomic per tile (a single address every tile of
    // the round adds to)
    bool claim = false;
    int claimed_tile = 0;
    if (threadIdx.x >= NTHREADS - 64 && threadIdx.x < NTHREADS - 64 + 27) {
      const int b = threadIdx.x - (NTHREADS - 64);
      const unsigned m = nbr_sides | (self_active ? 1u << 13 : 0u);
      const int nz = tz + b / 9 - 1, ny = ty + (b / 3) % 3 - 1, nx = tx + b % 3 - 1;
      if (((m >> b) & 1u) && nz >= 0 && nz < p.ntz && ny >= 0 && ny < p.nty && nx >= 0 && nx < p.ntx) {
        const int nt = (nz * p.nty + ny) * p.ntx + nx;
        if (b != 13) {  // seen from nt, this tile is at the mirrored offset
          atomicOr(&p.mail[nt], 1u << (26 - b));
          atomicOr(&p.dmail[nt], 1u << (26 - b));
        }
        claim = atomicExch(&p.flag_out[nt], 1) == 0;
        claimed_tile = nt;
      }
    }
    if (threadIdx.x >= NTHREADS - 64) {
      // one counter atomic for all the tiles this one appends
      const unsigned long long won = __ballot(claim);
      const int lane = threadIdx.x & 63;
      unsigned base = 0;
      if (won) {
        const int first = __builtin_ctzll(won);
        if (lane == first) base = atomicAdd(p.wl_cnt_out, (unsigned)__popcll(won));
        base = __shfl(base, first, 64);
        if (claim) p.wl_out[base + __popcll(won & ((1ull << lane) - 1))] = claimed_tile;
      }
    }
    // write back the tile's residual planes, excess, labels and arc masks
    // (whole: tracking which words changed measured slower, nearly every node
    // of a processed tile changes in a round); the appends above are in
    // flight meanwhile
#pragma unroll
    for (int j = 0; j < G::NPT; ++j) {
      const int t = threadIdx.x + j * NTHREADS;
      int z, y, x;
      tile_coords<G>(t, z, y, x);
      if (z >= ld || y >= lh || x >= lw) continue;
      const T e = s_ex[t];
      const int hv = s_h[((z + G::RZ) * G::HH + (y + G::RY)) * G::HW + (x + G::RX)];
      const unsigned g = (unsigned)(((long long)(z0 + z) * p.H + (y0 + y)) * p.W + (x0 + x));
      const unsigned o = g * (unsigned)sizeof(T);
      MaskT<K> m = 0;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const T r = s_res[i * NT + t];
        bst<T>(p.res + (long long)i * N, o, r);
        if (r > 0) m |= (MaskT<K>)1 << i;
      }
      bst<T>(p.ex, o, e);
      bst<int>(p.h, g * 4u, hv);
      bst<MaskT<K>>(p.mask, g * (unsigned)sizeof(MaskT<K>), m);
    }
    // no barrier here: the next tile's first barrier follows thread 0's
    // prologue; the appends above use only registers (nbr_sides was read
    // before the last barrier), while thread 0 may already reset s_nbr
  }
  // processed tiles: one atomic per block, not per tile
  if (threadIdx.x == 0 && ntiles_done) atomicAdd(p.tiles, ntiles_done);
}

// Work list of the tiles whose flag is set (the first round after a relabel).
__global__ __launch_bounds__(NTHREADS) void k_flags_to_list(const int* __restrict__ flags, int ntiles, int* list,
                                                            unsigned* cnt) {
  const int t = blockIdx.x * NTHREADS + threadIdx.x;
  if (t < ntiles && flags[t]) list[atomicAdd(cnt, 1u)] = t;
}

// Lattice coordinates of node v in 32-bit unsigned arithmetic (a short
// sequence where 64-bit division by the run-time extents is a long software
// one): the tiled engines keep N * sizeof(element) below 2^32, so v < 2^30.
__device__ inline void node_coords(long long v, int H, int W, int& z, int& y, int& x) {
  const unsigned u = (unsigned)v;
  const unsigned r = u / (unsigned)W;
  x = (int)(u - r * (unsigned)W);
  const unsigned zz = r / (unsigned)H;
  y = (int)(r - zz * (unsigned)H);
  z = (int)zz;
}

// ---- mailbox delivery (before any exact pass) --------------------------------

template <int K, class G, typename T, class DIR>
__device__ inline void deliver_node(const KP<T>& p, long long v) {
  int z, y, x;
  node_coords(v, p.H, p.W, z, y, x);
  const int lx = x % G::TW, ly = y % G::TH, lz = z % G::TD;
  if (lx >= G::RX && lx < G::TW - G::RX && ly >= G::RY && ly < G::TH - G::RY && lz >= G::RZ &&
      lz < G::TD - G::RZ)
    return;
  // only sides that sent this tile flow since the last delivery hold slots
  // with flow (every sender sets the bit after its rounds)
  const unsigned dm = p.dmail[((z / G::TD) * p.nty + y / G::TH) * p.ntx + x / G::TW];
  if (!dm) return;
  // every slot of the node read in one batch (no round runs concurrently here)
  long long slot[K];
  T a[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int dz = DIR::z(p.dirs, i), dy = DIR::y(p.dirs, i), dx = DIR::x(p.dirs, i);
    const int nz = z + dz, ny = y + dy, nx = x + dx;
    const int qz = lz + dz, qy = ly + dy, qx = lx + dx;
    const int side = (qz < 0 ? 0 : (qz >= G::TD ? 2 : 1)) * 9 + (qy < 0 ? 0 : (qy >= G::TH ? 2 : 1)) * 3 +
                     (qx < 0 ? 0 : (qx >= G::TW ? 2 : 1));
    const bool cross = nz >= 0 && nz < p.D && ny >= 0 && ny < p.H && nx >= 0 && nx < p.W && side != 13 &&
                       ((dm >> side) & 1u);
    slot[i] = cross ? (long long)((i + K / 2) % K) * p.N + ((long long)nz * p.H + ny) * p.W + nx : -1;
    const T m = p.mbox[slot[i] < 0 ? 0 : slot[i]];
    a[i] = slot[i] < 0 ? (T)0 : m;
  }
  T add_ex = 0;
  MaskT<K> got = 0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    if (a[i] == 0) continue;
    p.mbox[slot[i]] = 0;
    add_ex += a[i];
    got |= (MaskT<K>)1 << i;  // flow arrived along i: the arc back is residual
    p.res[(long long)i * p.N + v] += a[i];
  }
  if (add_ex != 0) {
    p.ex[v] += add_ex;
    reinterpret_cast<MaskT<K>*>(p.mask)[v] |= got;
  }
}

// Delivery pass; with init_labels it also sets the BFS start labels (h = 0 on
// sink-capable nodes, INF_H elsewhere) from the delivered excess, which saves
// the separate label pass when the arc masks are current.
template <int K, class G, typename T, class DIR>
__global__ __launch_bounds__(NTHREADS) void k_deliver(KP<T> p, int init_labels) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= p.N) return;
  deliver_node<K, G, T, DIR>(p, v);
  if (init_labels) p.h[v] = p.ex[v] < 0 ? 0 : INF_H;
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

// BFS start labels only (the arc masks are current): h[v] = 0 on sink-capable
// nodes, INF_H elsewhere.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_label_init(KP<T> p) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= p.N) return;
  p.h[v] = p.ex[v] < 0 ? 0 : INF_H;
}

// Tiles of the exact passes (BFS labels, S_min flood) are larger than the push
// tiles: only labels / reach bits and 1-bit arc masks are staged in LDS.
constexpr int BTHREADS = 512;

template <int TD_, int TH_, int TW_, int RZ_, int RY_, int RX_>
struct BGeo {
  static constexpr int TD = TD_, TH = TH_, TW = TW_, RZ = RZ_, RY = RY_, RX = RX_;
  static constexpr int NT = TD * TH * TW;
  static constexpr int HD = TD + 2 * RZ, HH = TH + 2 * RY, HW = TW + 2 * RX;
  static constexpr int NH = HD * HH * HW;
  static constexpr int NPT = NT / BTHREADS;
  static_assert(NT % BTHREADS == 0, "tile must be a multiple of the block");
};

struct BP {
  int D, H, W;
  long long N;
  int ntx, nty, ntz, ntiles, nblocks;
  int* flag_in;
  int* flag_out;
  unsigned long long* counter;
  // node visits of the exact BFS (the pull pass over a tile's nodes plus every
  // listed node of its relaxation steps), one slot per XCD; null: not counted
  unsigned long long* visits;
  DirTab dirs;
};

template <class B>
__device__ inline int bhalo(int t) {
  const int z = t / (B::TH * B::TW), y = (t / B::TW) % B::TH, x = t % B::TW;
  return ((z + B::RZ) * B::HH + (y + B::RY)) * B::HW + (x + B::RX);
}

template <class B>
__device__ inline bool bborder(int t) {
  const int z = t / (B::TH * B::TW), y = (t / B::TW) % B::TH, x = t % B::TW;
  return z < B::RZ || z >= B::TD - B::RZ || y < B::RY || y >= B::TH - B::RY || x < B::RX || x >= B::TW - B::RX;
}

// Flag the 3^d - 1 neighbouring tiles and record a border change (the host
// only tests the counter for zero, so a plain store serves every tile of the
// launch without an atomic on one address).
__device__ inline void flag_around(const BP& p, int tile) {
  *p.counter = 1ull;
  const int tx = tile % p.ntx, ty = (tile / p.ntx) % p.nty, tz = tile / (p.ntx * p.nty);
  for (int a = -1; a <= 1; ++a)
    for (int b = -1; b <= 1; ++b)
      for (int e = -1; e <= 1; ++e) {
        const int z = tz + a, y = ty + b, x = tx + e;
        if ((a | b | e) == 0 || z < 0 || z >= p.ntz || y < 0 || y >= p.nty || x < 0 || x >= p.ntx) continue;
        p.flag_out[(z * p.nty + y) * p.ntx + x] = 1;
      }
}

// One tile of the exact distance labelling d(v) = 1 + min{d(w) : arc v->w
// residual}, d = 0 on sink-capable nodes.  The tile first pulls every node once
// from its neighbours (halo included), then only nodes whose label dropped
// relax their in-neighbours inside the tile, until no label changes.  Labels
// only decrease and every value is the length of a real residual path, so the
// fixpoint over all tiles is the exact BFS distance.
template <int K, class B, typename MT, class DIR>
__global__ __launch_bounds__(BTHREADS) void k_bfs_tile(BP p, const MT* __restrict__ mask, int* __restrict__ hgl) {
  // a list holds up to half the tile; a step whose list overflowed walks the
  // listed nodes instead, so the lists cost half the LDS and more tiles are
  // resident per CU.
  // s_d[q] = 2 label + 1 while node q is not listed, 2 label while it is: one
  // atomicOr releases a node and returns its label, one atomicMin lowers a
  // neighbour's label and says whether it must be listed (it was not, and the
  // word changed), where a separate listed-bit array took a release, a fence
  // and a read, and a lowering plus a claim
  constexpr int LCAP = B::NT / 2;
  __shared__ unsigned s_d[B::NH];
  __shared__ MT s_m[B::NT];
  // 2-D tiles of byte masks (at most 8 directions): bit j of s_in[v] says the
  // in-tile neighbour v + dir j has an arc to v, so a relaxation reads one
  // byte instead of every neighbour's mask (2048^2 segmentation 1.257 -> 1.207
  // ms, 1024^2 TV path 3.70 -> 3.38 s; on 16^3 tiles of 32-bit masks its 16 KB
  // cost more residency than the reads it saves)
  constexpr bool INM = B::TD == 1 && sizeof(MT) == 1;
  __shared__ MT s_in[INM ? B::NT : 1];
  __shared__ unsigned short s_f[2][LCAP];
  __shared__ int s_n[3];
  __shared__ int s_flag;
  const int tile = tile_of_block(p.nblocks);
  if (tile >= p.ntiles) return;
  if (threadIdx.x == 0) {
    s_flag = p.flag_in[tile];
    p.flag_in[tile] = 0;
    s_n[0] = 0;
    s_n[1] = 0;
    s_n[2] = 0;
  }
  __syncthreads();
  if (!s_flag) return;
  const int tx = tile % p.ntx, ty = (tile / p.ntx) % p.nty, tz = tile / (p.ntx * p.nty);
  const int x0 = tx * B::TW, y0 = ty * B::TH, z0 = tz * B::TD;
  const int lw = min(B::TW, p.W - x0), lh = min(B::TH, p.H - y0), ld = min(B::TD, p.D - z0);
  // labels of the tile and its halo, and the arc masks.  Slots past the
  // lattice edge load node 0 and discard it, so that a thread's loads are all
  // in flight at once instead of waiting one by one behind branches.
  {
    constexpr int NHI = (B::NH + BTHREADS - 1) / BTHREADS;
    int hv[NHI];
    bool hin[NHI];
#pragma unroll
    for (int k = 0; k < NHI; ++k) {
      const int q = threadIdx.x + k * BTHREADS;
      const int hz = q / (B::HH * B::HW), hy = (q / B::HW) % B::HH, hx = q % B::HW;
      const int gz = z0 + hz - B::RZ, gy = y0 + hy - B::RY, gx = x0 + hx - B::RX;
      hin[k] = q < B::NH && gz >= 0 && gz < p.D && gy >= 0 && gy < p.H && gx >= 0 && gx < p.W;
      hv[k] = hgl[hin[k] ? ((long long)gz * p.H + gy) * p.W + gx : 0];
    }
    MT mv[B::NPT];
    bool in[B::NPT];
#pragma unroll
    for (int j = 0; j < B::NPT; ++j) {
      const int t = threadIdx.x + j * BTHREADS;
      const int z = t / (B::TH * B::TW), y = (t / B::TW) % B::TH, x = t % B::TW;
      in[j] = z < ld && y < lh && x < lw;
      mv[j] = mask[in[j] ? ((long long)(z0 + z) * p.H + (y0 + y)) * p.W + (x0 + x) : 0];
    }
#pragma unroll
    for (int k = 0; k < NHI; ++k) {
      const int q = threadIdx.x + k * BTHREADS;
      if (q < B::NH) s_d[q] = 2u * (unsigned)(hin[k] ? hv[k] : INF_H) + 1u;
    }
#pragma unroll
    for (int j = 0; j < B::NPT; ++j) {
      const int t = threadIdx.x + j * BTHREADS;
      s_m[t] = in[j] ? mv[j] : (MT)0;
    }
  }
  int d0[B::NPT];
  __syncthreads();
  // pull pass
#pragma unroll
  for (int j = 0; j < B::NPT; ++j) {
    const int t = threadIdx.x + j * BTHREADS;
    const int hi = bhalo<B>(t);
    const int dv = (int)(s_d[hi] >> 1);
    d0[j] = dv;
    const MT m = s_m[t];
    if constexpr (INM) {
      const int z = t / (B::TH * B::TW), y = (t / B::TW) % B::TH, x = t % B::TW;
      MT im = 0;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int dz = DIR::z(p.dirs, i), dy = DIR::y(p.dirs, i), dx = DIR::x(p.dirs, i);
        const int uz = z + dz, uy = y + dy, ux = x + dx;
        const bool intile = uz >= 0 && uz < B::TD && uy >= 0 && uy < B::TH && ux >= 0 && ux < B::TW;
        const MT mu = s_m[intile ? t + (dz * B::TH + dy) * B::TW + dx : t];
        if (intile && (mu & ((MT)1 << ((i + K / 2) % K)))) im |= (MT)1 << i;
      }
      s_in[t] = im;
    }
    if (dv == 0 || m == 0) continue;
    int best = dv;
#pragma unroll
    for (int i = 0; i < K; ++i)
      if (m & ((MT)1 << i)) {
        const int dn =
            (int)(s_d[hi + (DIR::z(p.dirs, i) * B::HH + DIR::y(p.dirs, i)) * B::HW + DIR::x(p.dirs, i)] >> 1) + 1;
        best = dn < best ? dn : best;
      }
    if (best < dv) {
      // only the owner writes its node in this pass: lowered and listed
      s_d[hi] = 2u * (unsigned)best;
      const int pos = atomicAdd(&s_n[0], 1);
      if (pos < LCAP) s_f[0][pos] = (unsigned short)t;
    }
  }
  __syncthreads();
  // one barrier per step: step k reads list k&1 with count s_n[k%3], appends
  // to list (k+1)&1 with count s_n[(k+1)%3], and clears s_n[(k+2)%3], which
  // every thread last read before the previous step's barrier
  unsigned long long nvis = (unsigned long long)ld * lh * lw;  // the pull pass
  for (int k = 0;; ++k) {
    const int cur = k & 1, c_in = k % 3, c_out = (k + 1) % 3;
    const int n = s_n[c_in];
    if (n == 0) break;
    nvis += (unsigned long long)n;
    if (threadIdx.x == 0) s_n[(k + 2) % 3] = 0;
    // n counts every claim of the previous step, also those past LCAP
    const bool scan = n > LCAP;
    const int lim = scan ? B::NT : n;
    for (int q = threadIdx.x; q < lim; q += BTHREADS) {
      int v;
      if (scan) {
        // a node listed during this step is a claim for the next one: taking
        // it now as well only repeats an idempotent relaxation
        v = q;
        if (s_d[bhalo<B>(v)] & 1u) continue;
      } else {
        v = s_f[cur][q];
      }
      // release v and read its label in one atomic: a later drop of the label
      // then lists v again instead of being lost
      const int hv = bhalo<B>(v);
      const int nd = (int)(atomicOr(&s_d[hv], 1u) >> 1) + 1;
      const unsigned e_nd = 2u * (unsigned)nd;  // label nd, listed
      const MT im = INM ? s_in[v] : (MT)0;
      const int z = v / (B::TH * B::TW), y = (v / B::TW) % B::TH, x = v % B::TW;
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int dz = DIR::z(p.dirs, j), dy = DIR::y(p.dirs, j), dx = DIR::x(p.dirs, j);
        const int u = v + (dz * B::TH + dy) * B::TW + dx;
        // the in-tile neighbours u = v + dir j with an arc u -> v (u's
        // direction rev(j))
        if constexpr (INM) {
          if (!((im >> j) & 1)) continue;
        } else {
          const int uz = z + dz, uy = y + dy, ux = x + dx;
          if (uz < 0 || uz >= B::TD || uy < 0 || uy >= B::TH || ux < 0 || ux >= B::TW) continue;
          if (!(s_m[u] & ((MT)1 << ((j + K / 2) % K)))) continue;
        }
        const int hu = hv + (dz * B::HH + dy) * B::HW + dx;
        if ((int)(s_d[hu] >> 1) > nd) {
          const unsigned o = atomicMin(&s_d[hu], e_nd);
          // this atomic listed u: it changed the word of a node not listed
          if (o > e_nd && (o & 1u)) {
            const int pos = atomicAdd(&s_n[c_out], 1);
            if (pos < LCAP) s_f[cur ^ 1][pos] = (unsigned short)u;
          }
        }
      }
    }
    __syncthreads();
  }
  bool border = false;
#pragma unroll
  for (int j = 0; j < B::NPT; ++j) {
    const int t = threadIdx.x + j * BTHREADS;
    const int z = t / (B::TH * B::TW), y = (t / B::TW) % B::TH, x = t % B::TW;
    if (z >= ld || y >= lh || x >= lw) continue;
    const int dv = (int)(s_d[bhalo<B>(t)] >> 1);
    if (dv != d0[j]) {
      hgl[((long long)(z0 + z) * p.H + (y0 + y)) * p.W + (x0 + x)] = dv;
      if (bborder<B>(t)) border = true;
    }
  }
  if (threadIdx.x == 0 && p.visits) atomicAdd(&p.visits[blockIdx.x & 7], nvis);
  if (__syncthreads_or(border) && threadIdx.x == 0) flag_around(p, tile);
}

// count active nodes (ex > 0, h < INF) and flag their tiles for the push rounds
template <class G, typename T>
__global__ __launch_bounds__(NTHREADS) void k_count_active(KP<T> p, long long* __restrict__ part) {
  long long cnt = 0;
  for (long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x; v < p.N; v += (long long)gridDim.x * NTHREADS) {
    if (p.ex[v] > 0 && p.h[v] < INF_H) {
      ++cnt;
      int z, y, x;
      node_coords(v, p.H, p.W, z, y, x);
      flag_neighbor_tile<G, T>(p, z, y, x, p.flag_out);
    }
  }
  cnt = block_reduce<false>(cnt);
  if (threadIdx.x == 0) part[blockIdx.x] = cnt;
}

// ---- S_min: flood fill from excess nodes over residual arcs --------------------

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_flood_init(KP<T> p, int* __restrict__ reach) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= p.N) return;
  reach[v] = p.ex[v] > 0 ? 1 : 0;
}

template <int K, class B, typename MT, class DIR>
__global__ __launch_bounds__(BTHREADS) void k_flood_tile(BP p, const MT* __restrict__ mask, int* __restrict__ reach) {
  __shared__ __align__(4) unsigned char s_r[(B::NH + 3) & ~3];
  __shared__ MT s_m[B::NH];
  __shared__ unsigned short s_f[2][B::NT];
  __shared__ int s_n[3];
  __shared__ int s_flag;
  const int tile = tile_of_block(p.nblocks);
  if (tile >= p.ntiles) return;
  if (threadIdx.x == 0) {
    s_flag = p.flag_in[tile];
    p.flag_in[tile] = 0;
    s_n[0] = 0;
    s_n[1] = 0;
    s_n[2] = 0;
  }
  __syncthreads();
  if (!s_flag) return;
  const int tx = tile % p.ntx, ty = (tile / p.ntx) % p.nty, tz = tile / (p.ntx * p.nty);
  const int x0 = tx * B::TW, y0 = ty * B::TH, z0 = tz * B::TD;
  const int lw = min(B::TW, p.W - x0), lh = min(B::TH, p.H - y0), ld = min(B::TD, p.D - z0);
  // reach bits and arc masks of the tile and its halo, all loads of a thread
  // in flight at once (slots past the lattice edge read node 0, discarded)
  {
    constexpr int NHI = (B::NH + BTHREADS - 1) / BTHREADS;
    int rv[NHI];
    MT mv[NHI];
    bool hin[NHI];
#pragma unroll
    for (int k = 0; k < NHI; ++k) {
      const int q = threadIdx.x + k * BTHREADS;
      const int hz = q / (B::HH * B::HW), hy = (q / B::HW) % B::HH, hx = q % B::HW;
      const int gz = z0 + hz - B::RZ, gy = y0 + hy - B::RY, gx = x0 + hx - B::RX;
      hin[k] = q < B::NH && gz >= 0 && gz < p.D && gy >= 0 && gy < p.H && gx >= 0 && gx < p.W;
      const long long g = hin[k] ? ((long long)gz * p.H + gy) * p.W + gx : 0;
      rv[k] = reach[g];
      mv[k] = mask[g];
    }
#pragma unroll
    for (int k = 0; k < NHI; ++k) {
      const int q = threadIdx.x + k * BTHREADS;
      if (q >= B::NH) continue;
      s_r[q] = hin[k] ? (unsigned char)rv[k] : (unsigned char)0;
      s_m[q] = hin[k] ? mv[k] : (MT)0;
    }
  }
  __syncthreads();
  int r0[B::NPT];
  // pull pass: w becomes reached when an in-neighbour v = w + dir[i] is reached
  // and v's arc rev(i) (pointing at w) is residual
#pragma unroll
  for (int j = 0; j < B::NPT; ++j) {
    const int t = threadIdx.x + j * BTHREADS;
    const int hi = bhalo<B>(t);
    r0[j] = s_r[hi];
    const int z = t / (B::TH * B::TW), y = (t / B::TW) % B::TH, x = t % B::TW;
    if (r0[j] || z >= ld || y >= lh || x >= lw) continue;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int q = hi + (DIR::z(p.dirs, i) * B::HH + DIR::y(p.dirs, i)) * B::HW + DIR::x(p.dirs, i);
      if (s_r[q] && (s_m[q] & ((MT)1 << ((i + K / 2) % K)))) {
        s_r[hi] = 1;
        s_f[0][atomicAdd(&s_n[0], 1)] = (unsigned short)t;
        break;
      }
    }
  }
  __syncthreads();
  // every owned node that was reached at the start also spreads inside the tile
#pragma unroll
  for (int j = 0; j < B::NPT; ++j) {
    const int t = threadIdx.x + j * BTHREADS;
    if (r0[j]) s_f[0][atomicAdd(&s_n[0], 1)] = (unsigned short)t;
  }
  __syncthreads();
  // one barrier per step: step k reads list k&1 with count s_n[k%3], appends
  // to list (k+1)&1 with count s_n[(k+1)%3], and clears s_n[(k+2)%3], which
  // every thread last read before the previous step's barrier
  for (int k = 0;; ++k) {
    const int cur = k & 1, c_in = k % 3, c_out = (k + 1) % 3;
    const int n = s_n[c_in];
    if (n == 0) break;
    if (threadIdx.x == 0) s_n[(k + 2) % 3] = 0;
    for (int q = threadIdx.x; q < n; q += BTHREADS) {
      const int v = s_f[cur][q];
      const int hv = bhalo<B>(v);
      const MT m = s_m[hv];
      const int z = v / (B::TH * B::TW), y = (v / B::TW) % B::TH, x = v % B::TW;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        if (!(m & ((MT)1 << i))) continue;
        const int dz = DIR::z(p.dirs, i), dy = DIR::y(p.dirs, i), dx = DIR::x(p.dirs, i);
        const int wz = z + dz, wy = y + dy, wx = x + dx;
        if (wz < 0 || wz >= B::TD || wy < 0 || wy >= B::TH || wx < 0 || wx >= B::TW) continue;
        const int hw = hv + (dz * B::HH + dy) * B::HW + dx;
        if (s_r[hw]) continue;
        // claim w through the byte's containing word
        unsigned int* wd = (unsigned int*)&s_r[hw & ~3];
        const int sh = (hw & 3) * 8;
        const unsigned int old = atomicOr(wd, 1u << sh);
        if (!((old >> sh) & 0xffu))
          s_f[cur ^ 1][atomicAdd(&s_n[c_out], 1)] = (unsigned short)(v + (dz * B::TH + dy) * B::TW + dx);
      }
    }
    __syncthreads();
  }
  bool border = false;
#pragma unroll
  for (int j = 0; j < B::NPT; ++j) {
    const int t = threadIdx.x + j * BTHREADS;
    const int z = t / (B::TH * B::TW), y = (t / B::TW) % B::TH, x = t % B::TW;
    if (z >= ld || y >= lh || x >= lw) continue;
    if (s_r[bhalo<B>(t)] && !r0[j]) {
      reach[((long long)(z0 + z) * p.H + (y0 + y)) * p.W + (x0 + x)] = 1;
      if (bborder<B>(t)) border = true;
    }
  }
  if (__syncthreads_or(border) && threadIdx.x == 0) flag_around(p, tile);
}

// ---- wave-tile exact passes (2D lattices, 3x3 neighbourhoods) ------------------
//
// A 64 x 64 tile is one wavefront: lane y holds row y of the tile as 64-bit
// words, bit x = column x.  The residual arcs are bit planes, one per
// neighbour slot s = 3 (dy + 1) + (dx + 1) (the node itself, s = 4, has none),
// stored tile-major so that the 64 words of one tile and plane are contiguous;
// plane 8 holds the sink-capable nodes.  A BFS level is then a handful of
// word operations per lane (shift the frontier of the row and of the rows
// above and below, AND with the arc planes), with no block barrier at all,
// and a flood closes each row with one carry-propagating add.

constexpr int WT = 64;      // side of a wave tile
constexpr int WPLANES = 9;  // 8 neighbour slots + sinks

__host__ __device__ constexpr int slot_plane(int dy, int dx) {
  return (dy + 1) * 3 + (dx + 1) < 4 ? (dy + 1) * 3 + (dx + 1) : (dy + 1) * 3 + (dx + 1) - 1;
}

struct WP {
  int D, H, W, ntx, nty;
  long long ps;  // words per plane = D * nty * ntx * 64
};

__device__ inline long long widx(const WP& w, int pl, int z, int ty, int tx, int lane) {
  return (long long)pl * w.ps + ((((long long)z * w.nty + ty) * w.ntx + tx) << 6) + lane;
}

// Bit planes of the residual arcs and sinks from the working state, and the
// BFS start labels (0 on sink-capable nodes, INF_H elsewhere).  One wave per
// 64-node word of a row; the word is the wave's ballot.  Plane words of absent
// directions and of rows past the lattice stay zero from allocation.  With
// use_mask the arcs come from the per-node arc masks (one byte per node
// instead of k residuals).
template <int K, typename T>
__global__ __launch_bounds__(NTHREADS) void k_bits_init(KP<T> p, WP w, unsigned long long* __restrict__ planes,
                                                        int use_mask) {
  const long long wave = ((long long)blockIdx.x * NTHREADS + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= (long long)p.D * p.H * w.ntx) return;  // uniform over the wave
  const int tx = (int)(wave % w.ntx);
  const long long r = wave / w.ntx;  // row over all layers
  const int y = (int)(r % p.H), z = (int)(r / p.H);
  const int x = tx * WT + lane;
  const bool in = x < p.W;
  const long long v = r * p.W + x;
  const long long base = widx(w, 0, z, y >> 6, tx, y & 63);
  const MaskT<K> m = (use_mask && in) ? reinterpret_cast<const MaskT<K>*>(p.mask)[v] : (MaskT<K>)0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const bool arc = use_mask ? ((m >> i) & 1) != 0 : (in && p.res[(long long)i * p.N + v] > 0);
    const unsigned long long b = __ballot(arc);
    if (lane == 0) planes[base + (long long)slot_plane(p.dirs.dy[i], p.dirs.dx[i]) * w.ps] = b;
  }
  const T e = in ? p.ex[v] : (T)0;
  const unsigned long long s = __ballot(e < 0);
  if (lane == 0) planes[base + 8LL * w.ps] = s;
  if (in) p.h[v] = e < 0 ? 0 : INF_H;
}

__device__ inline int wave_min_int(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int u = __shfl_xor(v, o, 64);
    v = u < v ? u : v;
  }
  return v;
}

// Flag the 8 neighbouring tiles of the same layer and count a border change
// (2D neighbourhoods never cross layers).
__device__ inline void flag_around2d(const BP& p, int tile) {
  atomicAdd(p.counter, 1ull);
  const int tx = tile % p.ntx, ty = (tile / p.ntx) % p.nty, tz = tile / (p.ntx * p.nty);
  for (int b = -1; b <= 1; ++b)
    for (int e = -1; e <= 1; ++e) {
      const int y = ty + b, x = tx + e;
      if ((b | e) == 0 || y < 0 || y >= p.nty || x < 0 || x >= p.ntx) continue;
      p.flag_out[(tz * p.nty + y) * p.ntx + x] = 1;
    }
}

// Whole-wave lane shifts by one (DPP wave_shr:1 / wave_shl:1, VALU moves
// instead of LDS permutes): up(v) gives lane i the value of lane i - 1 (lane 0
// gets 0), down(v) the value of lane i + 1 (lane 63 gets 0).
__device__ inline unsigned wave_up1(unsigned v) { return __builtin_amdgcn_update_dpp(0u, v, 0x138, 0xf, 0xf, false); }
__device__ inline unsigned wave_down1(unsigned v) { return __builtin_amdgcn_update_dpp(0u, v, 0x130, 0xf, 0xf, false); }
__device__ inline unsigned long long wave_up1(unsigned long long v) {
  return ((unsigned long long)wave_up1((unsigned)(v >> 32)) << 32) | wave_up1((unsigned)v);
}
__device__ inline unsigned long long wave_down1(unsigned long long v) {
  return ((unsigned long long)wave_down1((unsigned)(v >> 32)) << 32) | wave_down1((unsigned)v);
}
// value of a 64-bit register in lane r (r uniform over the wave)
__device__ inline unsigned long long lane_u64(unsigned long long v, int r) {
  return ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(v >> 32), r) << 32) |
         (unsigned)__builtin_amdgcn_readlane((int)v, r);
}

// One wave tile of the exact distance labelling (same fixpoint as k_bfs_tile):
// the tile's labels are recomputed level by level from its sinks and from the
// current labels of its halo, d(v) = 1 + min d(w) over residual arcs v -> w.
// Halo labels only decrease from launch to launch, so the recomputed labels
// never exceed the previous ones (a node the recomputation does not reach had
// no finite label); a changed border label flags the neighbours.
//
// A level costs a few word operations per lane; the labels of the nodes it
// reaches are stored into LDS without reading anything back, either bit by
// bit in each row or row by row across the lanes, whichever takes fewer
// stores, and the whole tile is written back at the end with row-coalesced
// stores.
__global__ __launch_bounds__(WT) void k_bfs_wave(BP p, WP w, const unsigned long long* __restrict__ planes,
                                                 int* __restrict__ hgl) {
  __shared__ int s_lab[WT * (WT + 1)];
  const int tile = tile_of_block(p.nblocks);
  if (tile >= p.ntiles) return;
  const int lane = threadIdx.x;
  int fl = 0;
  if (lane == 0) {
    fl = p.flag_in[tile];
    p.flag_in[tile] = 0;
  }
  if (!__shfl(fl, 0, 64)) return;
  const int tx = tile % p.ntx, ty = (tile / p.ntx) % p.nty, tz = tile / (p.ntx * p.nty);
  const int x0 = tx * WT, y0 = ty * WT, H = p.H, W = p.W;
  const long long lay = (long long)tz * H;
  const int y = y0 + lane, xc = x0 + lane;
  const bool rowok = y < H, colok = xc < W;
  const bool up = y0 > 0, dn = y0 + WT < H, lf = x0 > 0, rt = x0 + WT < W;
  const bool full_r = y0 + WT <= H, full_c = x0 + WT <= W;  // the tile has its last row / column
  unsigned long long P[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) P[s] = rowok ? planes[widx(w, s, tz, ty, tx, lane)] : 0ull;
  const unsigned long long sink = rowok ? planes[widx(w, 8, tz, ty, tx, lane)] : 0ull;
  const long long ry = (lay + y) * W;
  const int hl = (rowok && lf) ? hgl[ry + x0 - 1] : INF_H;
  const int hr = (rowok && rt) ? hgl[ry + x0 + WT] : INF_H;
  const long long rtop = (lay + y0 - 1) * W, rbot = (lay + y0 + WT) * W;
  const int ht = (colok && up) ? hgl[rtop + xc] : INF_H;
  const int hb = (colok && dn) ? hgl[rbot + xc] : INF_H;
  const int ctl = (up && lf) ? hgl[rtop + x0 - 1] : INF_H;
  const int ctr = (up && rt) ? hgl[rtop + x0 + WT] : INF_H;
  const int cbl = (dn && lf) ? hgl[rbot + x0 - 1] : INF_H;
  const int cbr = (dn && rt) ? hgl[rbot + x0 + WT] : INF_H;
  // the tile's own border labels before this launch (row 0 / 63 by column,
  // column 0 / 63 by row), to tell whether a border label changed
  const int o_top = colok ? hgl[(lay + y0) * W + xc] : INF_H;
  const int o_bot = (colok && full_r) ? hgl[(lay + y0 + WT - 1) * W + xc] : INF_H;
  const int o_lft = rowok ? hgl[ry + x0] : INF_H;
  const int o_rgt = (rowok && full_c) ? hgl[ry + x0 + WT - 1] : INF_H;

  // lowest finite halo label >= a (uniform over the wave), INF_H if none
  auto next_level = [&](int a) {
    int c = INF_H;
    c = (hl >= a && hl < c) ? hl : c;
    c = (hr >= a && hr < c) ? hr : c;
    c = (ht >= a && ht < c) ? ht : c;
    c = (hb >= a && hb < c) ? hb : c;
    c = (ctl >= a && ctl < c) ? ctl : c;
    c = (ctr >= a && ctr < c) ? ctr : c;
    c = (cbl >= a && cbl < c) ? cbl : c;
    c = (cbr >= a && cbr < c) ? cbr : c;
    return wave_min_int(c);
  };
  unsigned long long vis = sink, F = sink;
  int L = __ballot(sink != 0ull) ? 0 : next_level(0);
  while (L < INF_H) {
    // frontier of level L: tile nodes F, halo nodes labelled L
    const unsigned long long topw = __ballot(ht == L), botw = __ballot(hb == L);
    const unsigned e = (hl == L ? 1u : 0u) | (hr == L ? 2u : 0u);
    unsigned long long Fu = wave_up1(F), Fd = wave_down1(F);
    unsigned eu = wave_up1(e), ed = wave_down1(e);
    if (lane == 0) {
      Fu = topw;
      eu = (ctl == L ? 1u : 0u) | (ctr == L ? 2u : 0u);
    }
    if (lane == WT - 1) {
      Fd = botw;
      ed = (cbl == L ? 1u : 0u) | (cbr == L ? 2u : 0u);
    }
    // node (y, x) joins level L+1 when its arc towards (y+dy, x+dx) is
    // residual and that node is in the frontier
    unsigned long long c = 0;
    c |= P[0] & ((Fu << 1) | (unsigned long long)(eu & 1));
    c |= P[1] & Fu;
    c |= P[2] & ((Fu >> 1) | ((unsigned long long)(eu >> 1) << 63));
    c |= P[3] & ((F << 1) | (unsigned long long)(e & 1));
    c |= P[4] & ((F >> 1) | ((unsigned long long)(e >> 1) << 63));
    c |= P[5] & ((Fd << 1) | (unsigned long long)(ed & 1));
    c |= P[6] & Fd;
    c |= P[7] & ((Fd >> 1) | ((unsigned long long)(ed >> 1) << 63));
    const unsigned long long nw = c & ~vis;
    vis |= nw;
    ++L;
    const unsigned long long rows = __ballot(nw != 0ull);
    if (rows) {
      // few rows reached (a front along a row): the lanes store one row at a
      // time; otherwise each lane stores its own row's new bits
      if (__popcll(rows) <= 4) {
        for (unsigned long long rr = rows; rr; rr &= rr - 1) {
          const int r = __builtin_ctzll(rr);
          if ((lane_u64(nw, r) >> lane) & 1ull) s_lab[r * (WT + 1) + lane] = L;
        }
      } else {
        for (unsigned long long b = nw; b; b &= b - 1) s_lab[lane * (WT + 1) + __builtin_ctzll(b)] = L;
      }
    }
    F = nw;
    if (!rows) L = next_level(L);
  }
  __syncthreads();
  // write back every label of the tile (row r, column lane): sinks 0, reached
  // nodes their level, the rest INF_H
  bool border = false;
  for (int r = 0; r < WT; ++r) {
    const int yy = y0 + r;
    if (yy >= H) break;
    const unsigned long long vr = lane_u64(vis, r), sr = lane_u64(sink, r);
    const int lab = ((sr >> lane) & 1ull) ? 0 : (((vr >> lane) & 1ull) ? s_lab[r * (WT + 1) + lane] : INF_H);
    if (colok) {
      hgl[(lay + yy) * W + xc] = lab;
      if (r == 0 && lab != o_top) border = true;
      if (r == WT - 1 && lab != o_bot) border = true;
    }
  }
  if (rowok) {
    const int l0 = ((sink & 1ull) ? 0 : ((vis & 1ull) ? s_lab[lane * (WT + 1)] : INF_H));
    if (l0 != o_lft) border = true;
    if (full_c) {
      const int l1 = ((sink >> 63) ? 0 : ((vis >> 63) ? s_lab[lane * (WT + 1) + WT - 1] : INF_H));
      if (l1 != o_rgt) border = true;
    }
  }
  if (__any(border) && lane == 0) flag_around2d(p, tile);
}

// One wave tile of the S_min flood (same fixpoint as k_flood_tile): the
// tile's reached set grows by whole-row closures along the horizontal arcs
// (one add each way) and one-row steps along the other arcs, fed by the
// reached halo nodes' arcs into the tile, until it stops changing.
__global__ __launch_bounds__(WT) void k_flood_wave(BP p, WP w, const unsigned long long* __restrict__ planes,
                                                   int* __restrict__ reach) {
  const int tile = tile_of_block(p.nblocks);
  if (tile >= p.ntiles) return;
  const int lane = threadIdx.x;
  int fl = 0;
  if (lane == 0) {
    fl = p.flag_in[tile];
    p.flag_in[tile] = 0;
  }
  if (!__shfl(fl, 0, 64)) return;
  const int tx = tile % p.ntx, ty = (tile / p.ntx) % p.nty, tz = tile / (p.ntx * p.nty);
  const int x0 = tx * WT, y0 = ty * WT, H = p.H, W = p.W;
  const long long lay = (long long)tz * H;
  const int y = y0 + lane, xc = x0 + lane;
  const bool rowok = y < H, colok = xc < W;
  const bool up = y0 > 0, dn = y0 + WT < H, lf = x0 > 0, rt = x0 + WT < W;
  unsigned long long P[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) P[s] = rowok ? planes[widx(w, s, tz, ty, tx, lane)] : 0ull;
  // reached nodes of the tile, one ballot per row
  unsigned long long R = 0;
#pragma unroll 16
  for (int r = 0; r < WT; ++r) {
    const int yy = y0 + r;
    const unsigned long long b = __ballot(yy < H && colok && reach[(lay + yy) * W + xc] != 0);
    if (lane == r) R = b;
  }
  const unsigned long long R0 = R;

  // inflow from reached halo nodes (fixed during the launch)
  const long long ry = (lay + y) * W;
  // left column node (y, x0-1): arcs (dy, +1) enter (y+dy, 0); bit 63 of the left tile's words
  int cl = 0, cr = 0;
  if (rowok && lf && reach[ry + x0 - 1]) {
    cl = (int)(planes[widx(w, slot_plane(-1, 1), tz, ty, tx - 1, lane)] >> 63) |
         (int)(planes[widx(w, slot_plane(0, 1), tz, ty, tx - 1, lane)] >> 63) << 1 |
         (int)(planes[widx(w, slot_plane(1, 1), tz, ty, tx - 1, lane)] >> 63) << 2;
  }
  // right column node (y, x0+64): arcs (dy, -1) enter (y+dy, 63); bit 0 of the right tile's words
  if (rowok && rt && reach[ry + x0 + WT]) {
    cr = (int)(planes[widx(w, slot_plane(-1, -1), tz, ty, tx + 1, lane)] & 1ull) |
         (int)(planes[widx(w, slot_plane(0, -1), tz, ty, tx + 1, lane)] & 1ull) << 1 |
         (int)(planes[widx(w, slot_plane(1, -1), tz, ty, tx + 1, lane)] & 1ull) << 2;
  }
  const int clb = __shfl_down(cl, 1, 64), cla = __shfl_up(cl, 1, 64);
  const int crb = __shfl_down(cr, 1, 64), cra = __shfl_up(cr, 1, 64);
  unsigned long long hin = 0;
  // (y+1, x0-1) --(-1,+1)--> (y, 0);  (y, x0-1) --(0,+1)--> (y, 0);  (y-1, x0-1) --(+1,+1)--> (y, 0)
  if (((lane < WT - 1 ? clb : 0) & 1) | ((cl >> 1) & 1) | ((lane > 0 ? cla : 0) >> 2 & 1)) hin |= 1ull;
  if (((lane < WT - 1 ? crb : 0) & 1) | ((cr >> 1) & 1) | ((lane > 0 ? cra : 0) >> 2 & 1)) hin |= 1ull << 63;
  // top halo row y0-1 (last row of the tile above): arcs (+1, dx) enter row 0
  const unsigned long long rtw = __ballot(up && colok && reach[(lay + y0 - 1) * W + xc] != 0);
  const unsigned long long rbw = __ballot(dn && colok && reach[(lay + y0 + WT) * W + xc] != 0);
  if (lane == 0 && up) {
    const unsigned long long a5 = planes[widx(w, slot_plane(1, -1), tz, ty - 1, tx, WT - 1)];
    const unsigned long long a6 = planes[widx(w, slot_plane(1, 0), tz, ty - 1, tx, WT - 1)];
    const unsigned long long a7 = planes[widx(w, slot_plane(1, 1), tz, ty - 1, tx, WT - 1)];
    hin |= ((rtw & a5) >> 1) | (rtw & a6) | ((rtw & a7) << 1);
    const long long rtop = (lay + y0 - 1) * W;
    if (lf && reach[rtop + x0 - 1] && (planes[widx(w, slot_plane(1, 1), tz, ty - 1, tx - 1, WT - 1)] >> 63))
      hin |= 1ull;
    if (rt && reach[rtop + x0 + WT] && (planes[widx(w, slot_plane(1, -1), tz, ty - 1, tx + 1, WT - 1)] & 1ull))
      hin |= 1ull << 63;
  }
  // bottom halo row y0+64 (first row of the tile below): arcs (-1, dx) enter row 63
  if (lane == WT - 1 && dn) {
    const unsigned long long a0 = planes[widx(w, slot_plane(-1, -1), tz, ty + 1, tx, 0)];
    const unsigned long long a1 = planes[widx(w, slot_plane(-1, 0), tz, ty + 1, tx, 0)];
    const unsigned long long a2 = planes[widx(w, slot_plane(-1, 1), tz, ty + 1, tx, 0)];
    hin |= ((rbw & a0) >> 1) | (rbw & a1) | ((rbw & a2) << 1);
    const long long rbot = (lay + y0 + WT) * W;
    if (lf && reach[rbot + x0 - 1] && (planes[widx(w, slot_plane(-1, 1), tz, ty + 1, tx - 1, 0)] >> 63))
      hin |= 1ull;
    if (rt && reach[rbot + x0 + WT] && (planes[widx(w, slot_plane(-1, -1), tz, ty + 1, tx + 1, 0)] & 1ull))
      hin |= 1ull << 63;
  }

  R |= hin;
  const unsigned long long gr = P[slot_plane(0, 1)];                         // arcs x -> x+1
  const unsigned long long gl = __builtin_bitreverse64(P[slot_plane(0, -1)]);  // arcs x -> x-1, mirrored
  for (;;) {
    // closure along the row: a run of open arcs carries every seed in it to
    // the run's far end in one add
    R |= (gr + (R & gr)) ^ gr;
    unsigned long long m = __builtin_bitreverse64(R);
    m |= (gl + (m & gl)) ^ gl;
    R = __builtin_bitreverse64(m);
    // one step along the arcs between rows
    const unsigned long long to_dn = ((R & P[slot_plane(1, -1)]) >> 1) | (R & P[slot_plane(1, 0)]) |
                                     ((R & P[slot_plane(1, 1)]) << 1);
    const unsigned long long to_up = ((R & P[slot_plane(-1, -1)]) >> 1) | (R & P[slot_plane(-1, 0)]) |
                                     ((R & P[slot_plane(-1, 1)]) << 1);
    const unsigned long long from_up = wave_up1(to_dn), from_dn = wave_down1(to_up);
    const unsigned long long Rn = R | from_up | from_dn;
    if (!__any(Rn != R)) break;
    R = Rn;
  }
  const unsigned long long add = R & ~R0;
  const unsigned long long bord = (lane == 0 || lane == WT - 1) ? add : (add & 0x8000000000000001ull);
  const bool border = __any(bord != 0ull);
  for (int r = 0; r < WT; ++r) {
    const unsigned long long a = lane_u64(add, r);
    if ((a >> lane) & 1ull) reach[(lay + y0 + r) * W + xc] = 1;
  }
  if (border && lane == 0) flag_around2d(p, tile);
}

// ---- outputs -------------------------------------------------------------------

// flow = sum(max(red0,0)) - sum(max(-ex,0)), exact in int64
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_flow(KP<T> p, const long long* __restrict__ red0,
                                                   long long* __restrict__ part) {
  long long f = 0;
  for (long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x; v < p.N; v += (long long)gridDim.x * NTHREADS) {
    const long long r = red0[v];
    const long long e = (long long)p.ex[v];
    f += (r > 0 ? r : 0) - (e < 0 ? -e : 0);
  }
  f = block_reduce<false>(f);
  if (threadIdx.x == 0) part[blockIdx.x] = f;
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_cut(KP<T> p, const int* __restrict__ reach, int which,
                                                  int* __restrict__ out) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= p.N) return;
  out[v] = which == LF_CUT_SMAX ? (p.h[v] >= INF_H ? 1 : 0) : reach[v];
}

// ---- batched lattices on one canvas ------------------------------------------
//
// Lattice i of a batch owns the canvas rectangle rects[i]; a flat list of
// blocks covers every lattice in chunks of BCHUNK nodes (blk[i] = first block
// of lattice i), so the per-lattice loops touch no filler node.

constexpr int BCHUNK = NTHREADS * 8;

__device__ inline int batch_block_owner(const long long* __restrict__ blk, int n) {
  __shared__ int s_owner;
  if (threadIdx.x == 0) {
    int lo = 0, hi = n - 1;  // largest i with blk[i] <= blockIdx.x
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (blk[mid] <= (long long)blockIdx.x) lo = mid;
      else hi = mid - 1;
    }
    s_owner = lo;
  }
  __syncthreads();
  return s_owner;
}

// flow of every lattice: sum(max(red0,0)) - sum(max(-ex,0)) over its nodes
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_batch_flow(KP<T> p, const long long* __restrict__ red0,
                                                         const BatchRect* __restrict__ rects,
                                                         const long long* __restrict__ blk, int n,
                                                         unsigned long long* __restrict__ flows) {
  const int i = batch_block_owner(blk, n);
  const BatchRect R = rects[i];
  const long long m = R.h * R.w;
  const long long j0 = ((long long)blockIdx.x - blk[i]) * BCHUNK;
  const long long j1 = j0 + BCHUNK < m ? j0 + BCHUNK : m;
  long long f = 0;
  for (long long j = j0 + threadIdx.x; j < j1; j += NTHREADS) {
    const long long v = (R.y0 + j / R.w) * p.W + R.x0 + j % R.w;
    const long long r = red0[v];
    const long long e = (long long)p.ex[v];
    f += (r > 0 ? r : 0) - (e < 0 ? -e : 0);
  }
  f = block_reduce<false>(f);
  if (threadIdx.x == 0) atomicAdd(&flows[i], (unsigned long long)f);
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_batch_cut(KP<T> p, const int* __restrict__ reach, int which,
                                                        const BatchRect* __restrict__ rects,
                                                        const long long* __restrict__ blk, int n,
                                                        int* __restrict__ out) {
  const int i = batch_block_owner(blk, n);
  const BatchRect R = rects[i];
  const long long m = R.h * R.w;
  const long long j0 = ((long long)blockIdx.x - blk[i]) * BCHUNK;
  const long long j1 = j0 + BCHUNK < m ? j0 + BCHUNK : m;
  for (long long j = j0 + threadIdx.x; j < j1; j += NTHREADS) {
    const long long v = (R.y0 + j / R.w) * p.W + R.x0 + j % R.w;
    out[R.off + j] = which == LF_CUT_SMAX ? (p.h[v] >= INF_H ? 1 : 0) : reach[v];
  }
}

// Energies of one lattice of the batch (its own layout, rows of R.w nodes)
// into the canvas network; the gather form of k_reparam over local
// coordinates, so arcs that would leave the lattice are never written.
template <int K>
__global__ __launch_bounds__(NTHREADS) void k_batch_reparam(BatchRect R, int Wc, long long N, DirTab dirs,
                                                            const long long* __restrict__ unary,
                                                            const long long* __restrict__ binary,
                                                            long long* __restrict__ red0,
                                                            long long* __restrict__ alpha0, int* err) {
  const long long j = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (j >= R.h * R.w) return;
  constexpr int KP2 = K / 2;
  const int ly = (int)(j / R.w), lx = (int)(j % R.w);
  const long long v = (R.y0 + ly) * (long long)Wc + R.x0 + lx;
  long long r = 2 * (unary[2 * j + 1] - unary[2 * j + 0]);
  bool bad = false;
#pragma unroll
  for (int e = 0; e < KP2; ++e) {
    const int dy = dirs.dy[e], dx = dirs.dx[e], pe = R.perm[e];
    {
      const int ny = ly + dy, nx = lx + dx;
      if (ny >= 0 && ny < R.h && nx >= 0 && nx < R.w) {
        const long long* E = binary + (j * KP2 + pe) * 4;
        const long long e00 = E[0], e01 = E[1], e10 = E[2], e11 = E[3];
        const long long q1 = e10 - e00, q12 = e00 + e11 - e01 - e10;
        bad |= q12 > 0;
        r += 2 * q1 + q12;
        alpha0[(long long)e * N + v] += -q12;
      }
    }
    {
      const int ny = ly - dy, nx = lx - dx;
      if (ny >= 0 && ny < R.h && nx >= 0 && nx < R.w) {
        const long long w = (long long)ny * R.w + nx;
        const long long* E = binary + (w * KP2 + pe) * 4;
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
// capacity; part[2b], part[2b+1] receive block maxima of |red0| and alpha0,
// which bound every excess (|red0| + k max alpha) and residual (2 max alpha).
template <int K, typename T>
__global__ __launch_bounds__(NTHREADS) void k_load(KP<T> p, const long long* __restrict__ red0,
                                                   const long long* __restrict__ alpha0, int* err,
                                                   long long* __restrict__ part, int zero_mbox) {
  long long mr = 0, ma = 0;
  bool bad = false;
  const long long tid = (long long)blockIdx.x * NTHREADS + threadIdx.x, step = (long long)gridDim.x * NTHREADS;
  // two nodes per thread with 16-byte loads when every plane starts 16-byte
  // aligned (N even; the planes are separate allocations)
  using T2 = typename std::conditional<sizeof(T) == 8, longlong2, int2>::type;
  const long long npair = (p.N & 1) ? 0 : p.N / 2;
  for (long long w = tid; w < npair; w += step) {
    const long long v = 2 * w;
    const longlong2 r = *reinterpret_cast<const longlong2*>(red0 + v);
    const long long r0 = r.x < 0 ? -r.x : r.x, r1 = r.y < 0 ? -r.y : r.y;
    mr = r0 > mr ? r0 : mr;
    mr = r1 > mr ? r1 : mr;
    MaskT<K> m0 = 0, m1 = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const longlong2 a = *reinterpret_cast<const longlong2*>(alpha0 + (long long)i * p.N + v);
      bad |= (a.x < 0) | (a.y < 0);
      ma = a.x > ma ? a.x : ma;
      ma = a.y > ma ? a.y : ma;
      if (a.x > 0) m0 |= (MaskT<K>)1 << i;
      if (a.y > 0) m1 |= (MaskT<K>)1 << i;
      T2 o;
      o.x = (T)a.x;
      o.y = (T)a.y;
      *reinterpret_cast<T2*>(p.res + (long long)i * p.N + v) = o;
      if (zero_mbox) {
        T2 z;
        z.x = 0;
        z.y = 0;
        *reinterpret_cast<T2*>(p.mbox + (long long)i * p.N + v) = z;
        *reinterpret_cast<T2*>(p.mbox_in + (long long)i * p.N + v) = z;
      }
    }
    T2 e;
    e.x = (T)(-r.x);
    e.y = (T)(-r.y);
    *reinterpret_cast<T2*>(p.ex + v) = e;
    reinterpret_cast<MaskT<K>*>(p.mask)[v] = m0;
    reinterpret_cast<MaskT<K>*>(p.mask)[v + 1] = m1;
  }
  for (long long v = npair ? p.N : tid; v < p.N; v += step) {
    const long long r = red0[v];
    mr = (r < 0 ? -r : r) > mr ? (r < 0 ? -r : r) : mr;
    MaskT<K> m = 0;
    // (odd N only; a fully unrolled loop of more than 16 planes held every
    // plane's load and address in registers and set the register count of
    // the whole kernel: 200 VGPRs for 26 int64 planes, 2 waves per SIMD; the
    // 256^3 26-direction load measured 6 % of a solve faster without it)
#pragma unroll(K > 16 ? 1 : K)
    for (int i = 0; i < K; ++i) {
      const long long a = alpha0[(long long)i * p.N + v];
      bad |= a < 0;
      ma = a > ma ? a : ma;
      if (a > 0) m |= (MaskT<K>)1 << i;
      p.res[(long long)i * p.N + v] = (T)a;
      if (zero_mbox) {
        p.mbox[(long long)i * p.N + v] = 0;
        p.mbox_in[(long long)i * p.N + v] = 0;
      }
    }
    p.ex[v] = (T)(-r);
    reinterpret_cast<MaskT<K>*>(p.mask)[v] = m;
  }
  mr = block_reduce<true>(mr);
  ma = block_reduce<true>(ma);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = mr;
    part[2 * blockIdx.x + 1] = ma;
  }
  if (bad) atomicOr(err, 2);
}

template <typename V>
__global__ void k_fill(V* a, V val, long long n) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v < n) a[v] = val;
}

// Up to six 32-bit fills in one launch (the per-relabel resets of tile flags,
// mail bits and list counters are each a few thousand words).
struct FillSeg {
  unsigned* p;
  unsigned v;
  long long n;
};
struct FillSet {
  FillSeg s[6];
  int ns;
};
__global__ __launch_bounds__(NTHREADS) void k_fill_multi(FillSet f) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  for (int i = 0; i < f.ns; ++i)
    if (v < f.s[i].n) f.s[i].p[v] = f.s[i].v;
}

// ---- parametric divide and conquer (ParametricFlowSolver, pf_solver.hpp:33-208) ----
//
// All regions of one depth of the divide-and-conquer tree are cut in one
// lattice min-cut: each node carries a region id, arcs between regions are
// zeroed when a region splits (the reference gives the two parts different
// keys, push_relabel.hpp:285-303, so those arcs are never used again), and the
// excess of a node is t(region) - r with t = ceil(mean r) (minus one when it
// equals max r) (setToMeanReductionPoint, pf_flow_node.hpp:331-396).  A region
// is final when max r - min r <= 1.  Region sums are exact (128-bit).

struct RegionStat {
  unsigned long long sum_lo, sum_hi;  // 128-bit two's complement sum of r
  unsigned long long mn, mx;          // order-preserving encodings of min / max r
  unsigned long long cnt, cnt1;       // nodes; nodes on the source side after the cut
  long long thr;                      // threshold of this level
  int state;                          // 1 active, 2 final, 3 waiting (no cut at this depth)
  int newid;                          // id of the split-off source-side part, -1 when none
  unsigned long long pmn, pmx;        // order-preserving min / max of the predicted values
  int mode;                           // 0: predicted thresholds allowed, 1: mean thresholds only
  int hslot;                          // histogram of this depth, -1 when none
  int kind;                           // this depth's cut: 0 at the mean, 1 at a predicted gap
  int fresh;                          // threshold set since the last setup of excesses
  int busy;                           // active nodes seen at the last relabel
  int pad_;
};

struct PFP {
  long long* r;
  int* reg;
  RegionStat* tab;
  int* nreg;
  int* err;       // bit 0: the region table is full
  long long N;
  long long cap;  // entries of tab
  int D, H, W;
  int collapse;   // a region that a cut at its mean does not split is final
  // thresholds from predictions: predicted value of node v in r units is
  // pred[v] * pscale + poff; hist holds nbins counters per histogram slot
  const float* pred;
  double pscale, poff;
  unsigned* hist;
  int nbins;
};

__device__ inline long long pred_r(const PFP& f, long long v) {
  return (long long)floor((double)f.pred[v] * f.pscale + f.poff);
}

__device__ inline unsigned long long ord64(long long x) { return (unsigned long long)x ^ 0x8000000000000000ull; }
__device__ inline long long unord64(unsigned long long u) { return (long long)(u ^ 0x8000000000000000ull); }

__device__ inline void add128(RegionStat* t, long long x) {
  const unsigned long long ux = (unsigned long long)x;
  const unsigned long long old = atomicAdd(&t->sum_lo, ux);
  const unsigned long long carry = (old + ux) < old ? 1ull : 0ull;
  const unsigned long long hi = (x < 0 ? ~0ull : 0ull) + carry;
  if (hi) atomicAdd(&t->sum_hi, hi);
}

// correctly rounded conversion of a 128-bit two's complement integer
__device__ inline double i128_to_double(unsigned long long hi, unsigned long long lo) {
  const bool neg = (long long)hi < 0;
  if (neg) {
    lo = ~lo + 1ull;
    hi = ~hi + (lo == 0 ? 1ull : 0ull);
  }
  double d;
  if (hi == 0) {
    d = (double)lo;
  } else {
    // keep the top 63 significant bits, fold the rest into a sticky bit, so the
    // single rounding of the 63-bit conversion is the correct one
    const int bl = 128 - __clzll((long long)hi);  // 65..128
    const int sh = bl - 63;                        // 2..65
    unsigned long long top, sticky;
    if (sh < 64) {
      top = (lo >> sh) | (hi << (64 - sh));
      sticky = (lo & ((1ull << sh) - 1ull)) != 0 ? 1ull : 0ull;
    } else {
      top = hi >> (sh - 64);
      sticky = (lo != 0 || (hi & ((1ull << (sh - 64)) - 1ull)) != 0) ? 1ull : 0ull;
    }
    d = ldexp((double)(top | sticky), sh);
  }
  return neg ? -d : d;
}

__global__ __launch_bounds__(NTHREADS) void k_pf_clear(PFP f, int nreg, int st) {
  const int k = blockIdx.x * NTHREADS + threadIdx.x;
  if (k >= nreg) return;
  RegionStat& t = f.tab[k];
  if (t.state == 3) t.state = 1;
  if (t.state != st) return;
  t.sum_lo = 0;
  t.sum_hi = 0;
  t.mn = ~0ull;
  t.mx = 0;
  t.cnt = 0;
  t.cnt1 = 0;
  t.newid = -1;
  t.pmn = ~0ull;
  t.pmx = 0;
  t.hslot = -1;
  t.kind = 0;
}

// Per-region aggregation of the active nodes.  MODE 0: sum / min / max / count
// of r.  MODE 1: count of the nodes with side[v] != 0.  MODE 2: min / max of
// the predicted values.  Each block folds its
// nodes into an LDS hash table keyed by region (spatially coherent, so a block
// sees few regions) and flushes it with one set of atomics per region.
template <int MODE>
__global__ __launch_bounds__(NTHREADS) void k_pf_stats_impl(PFP f, const int* __restrict__ side, int st) {
  constexpr int HSZ = 256;
  __shared__ int hk[HSZ];
  __shared__ long long hs[HSZ];
  __shared__ unsigned long long hmn[HSZ], hmx[HSZ];
  __shared__ unsigned int hc[HSZ];
  for (int i = threadIdx.x; i < HSZ; i += NTHREADS) {
    hk[i] = -1;
    hs[i] = 0;
    hmn[i] = ~0ull;
    hmx[i] = 0;
    hc[i] = 0;
  }
  __syncthreads();
  // Each thread folds runs of RUN consecutive nodes in registers: neighbouring
  // nodes mostly share a region, so the table takes one update per region
  // change instead of one per node (per-node updates serialise the wave on
  // one LDS slot).  Sums, extremes and counts are integers: the result does
  // not depend on the order.
  constexpr int RUN = 8;
  auto put = [&](int k, long long sum, unsigned long long mn, unsigned long long mx, unsigned c) {
    int h = (int)(((unsigned)k * 2654435761u) >> 24) & (HSZ - 1);
    int slot = -1;
    for (int probe = 0; probe < 16; ++probe) {
      const int cur = hk[h];
      if (cur == k) { slot = h; break; }
      if (cur == -1) {
        const int old = atomicCAS(&hk[h], -1, k);
        if (old == -1 || old == k) { slot = h; break; }
      }
      h = (h + 1) & (HSZ - 1);
    }
    if (slot >= 0) {
      if (MODE == 0) atomicAdd((unsigned long long*)&hs[slot], (unsigned long long)sum);
      if (MODE != 1) {
        atomicMin(&hmn[slot], mn);
        atomicMax(&hmx[slot], mx);
      }
      atomicAdd(&hc[slot], c);
    } else {
      RegionStat* t = &f.tab[k];
      if (MODE == 0) {
        add128(t, sum);
        atomicMin(&t->mn, mn);
        atomicMax(&t->mx, mx);
        atomicAdd(&t->cnt, (unsigned long long)c);
      } else if (MODE == 2) {
        atomicMin(&t->pmn, mn);
        atomicMax(&t->pmx, mx);
      } else {
        atomicAdd(&t->cnt1, (unsigned long long)c);
      }
    }
  };
  const long long nrun = (f.N + RUN - 1) / RUN;
  for (long long c = (long long)blockIdx.x * NTHREADS + threadIdx.x; c < nrun; c += (long long)gridDim.x * NTHREADS) {
    const long long v0 = c * RUN, v1 = v0 + RUN < f.N ? v0 + RUN : f.N;
    int ck = -1;
    bool live = false;
    long long sum = 0;
    unsigned long long mn = ~0ull, mx = 0;
    unsigned cnt = 0;
    for (long long v = v0; v < v1; ++v) {
      const int k = f.reg[v];
      if (k != ck) {
        if (cnt) put(ck, sum, mn, mx, cnt);
        ck = k;
        live = f.tab[k].state == st;
        sum = 0;
        mn = ~0ull;
        mx = 0;
        cnt = 0;
      }
      if (!live) continue;
      const long long x = MODE == 0 ? f.r[v] : MODE == 2 ? pred_r(f, v) : (long long)(side[v] != 0);
      if (MODE == 1 && x == 0) continue;
      sum += x;
      mn = ord64(x) < mn ? ord64(x) : mn;
      mx = ord64(x) > mx ? ord64(x) : mx;
      ++cnt;
    }
    if (cnt) put(ck, sum, mn, mx, cnt);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < HSZ; i += NTHREADS) {
    const int k = hk[i];
    if (k < 0 || hc[i] == 0) continue;
    RegionStat* t = &f.tab[k];
    if (MODE == 0) {
      add128(t, hs[i]);
      atomicMin(&t->mn, hmn[i]);
      atomicMax(&t->mx, hmx[i]);
      atomicAdd(&t->cnt, (unsigned long long)hc[i]);
    } else if (MODE == 2) {
      atomicMin(&t->pmn, hmn[i]);
      atomicMax(&t->pmx, hmx[i]);
    } else {
      atomicAdd(&t->cnt1, (unsigned long long)hc[i]);
    }
  }
}

// n[0]: active regions; n[1]: regions that get a histogram of predicted
// values (guided, mode 0, predicted range >= min_range)
__global__ __launch_bounds__(NTHREADS) void k_pf_thresh(PFP f, int nreg, unsigned long long* n, long long min_range,
                                                        int st) {
  const int k = blockIdx.x * NTHREADS + threadIdx.x;
  if (k >= nreg) return;
  RegionStat& t = f.tab[k];
  if (t.state != st) return;
  const long long mn = unord64(t.mn), mx = unord64(t.mx);
  if (mx - mn <= 1) {
    t.state = 2;
    return;
  }
  const double sum = i128_to_double(t.sum_hi, t.sum_lo);
  long long thr = (long long)ceil(sum / (double)t.cnt);
  if (thr == mx) --thr;
  t.thr = thr;
  t.kind = 0;
  t.state = 1;
  t.fresh = 1;
  atomicAdd(&n[0], 1ull);
  if (f.pred && t.mode == 0 && unord64(t.pmx) - unord64(t.pmn) >= min_range)
    t.hslot = (int)atomicAdd(&n[1], 1ull);
}

// histogram of the predicted values of every region with a slot (nbins bins
// over [pmn, pmx]); an LDS table keyed by (slot, bin) folds each block's nodes
__global__ __launch_bounds__(NTHREADS) void k_pf_hist(PFP f) {
  constexpr int HSZ = 512;
  __shared__ int hk[HSZ];
  __shared__ unsigned hc[HSZ];
  for (int i = threadIdx.x; i < HSZ; i += NTHREADS) {
    hk[i] = -1;
    hc[i] = 0;
  }
  __syncthreads();
  // runs of RUN consecutive nodes counted in registers while their bin stays
  // the same (smooth predictions: neighbours mostly share one), one table
  // update per run instead of one per node
  constexpr int RUN = 8;
  auto put = [&](int key, unsigned c) {
    int h = (int)(((unsigned)key * 2654435761u) >> 23) & (HSZ - 1);
    int slot = -1;
    for (int probe = 0; probe < 16; ++probe) {
      const int cur = hk[h];
      if (cur == key) { slot = h; break; }
      if (cur == -1) {
        const int old = atomicCAS(&hk[h], -1, key);
        if (old == -1 || old == key) { slot = h; break; }
      }
      h = (h + 1) & (HSZ - 1);
    }
    if (slot >= 0) atomicAdd(&hc[slot], c);
    else atomicAdd(&f.hist[key], c);
  };
  const long long nrun = (f.N + RUN - 1) / RUN;
  for (long long c = (long long)blockIdx.x * NTHREADS + threadIdx.x; c < nrun; c += (long long)gridDim.x * NTHREADS) {
    const long long v0 = c * RUN, v1 = v0 + RUN < f.N ? v0 + RUN : f.N;
    int ckey = -1;
    unsigned cnt = 0;
    for (long long v = v0; v < v1; ++v) {
      const RegionStat& t = f.tab[f.reg[v]];
      if (t.state != 1 || !t.fresh || t.hslot < 0) continue;
      const long long lo = unord64(t.pmn), range = unord64(t.pmx) - lo + 1;
      long long b = (long long)((double)(pred_r(f, v) - lo) * f.nbins / (double)range);
      b = b < 0 ? 0 : (b >= f.nbins ? f.nbins - 1 : b);
      const int key = t.hslot * f.nbins + (int)b;
      if (key != ckey) {
        if (cnt) put(ckey, cnt);
        ckey = key;
        cnt = 0;
      }
      ++cnt;
    }
    if (cnt) put(ckey, cnt);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < HSZ; i += NTHREADS)
    if (hk[i] >= 0 && hc[i]) atomicAdd(&f.hist[hk[i]], hc[i]);
}

// Threshold at the emptiest bin boundary of the predicted values (each side
// holding at least min_side of the region): a cut there separates predicted
// clusters without a large level set near the threshold.  Accepted when the two
// bins around it hold at most gap_frac of the region.  n_pred counts them.
__global__ __launch_bounds__(NTHREADS) void k_pf_pick(PFP f, int nreg, double gap_frac, double min_side,
                                                      unsigned long long* n_pred) {
  const int k = blockIdx.x * NTHREADS + threadIdx.x;
  if (k >= nreg) return;
  RegionStat& t = f.tab[k];
  if (t.state != 1 || !t.fresh || t.hslot < 0) return;
  const unsigned* h = f.hist + (long long)t.hslot * f.nbins;
  const double n = (double)t.cnt;
  unsigned long long left = h[0];
  int best = -1;
  double best_mass = 1e300, best_bal = 1e300;
  for (int j = 1; j < f.nbins; ++j) {
    const double l = (double)left, r = n - l;
    if (l >= min_side * n && r >= min_side * n && l >= 1 && r >= 1) {
      const double mass = (double)h[j - 1] + (double)h[j];
      const double bal = fabs(l - r);
      if (mass < best_mass || (mass == best_mass && bal < best_bal)) {
        best = j;
        best_mass = mass;
        best_bal = bal;
      }
    }
    left += h[j];
  }
  if (best < 0 || best_mass > gap_frac * n) return;
  const long long lo = unord64(t.pmn), range = unord64(t.pmx) - lo + 1;
  // first predicted value of bin `best`
  const long long thr = lo + (long long)ceil((double)best * (double)range / (double)f.nbins);
  const long long mn = unord64(t.mn), mx = unord64(t.mx);
  if (thr <= mn || thr > mx) return;  // no split possible at thr
  t.thr = thr;
  t.kind = 1;
  atomicAdd(n_pred, 1ull);
}

// while some region cuts at a predicted gap, the mean cuts wait
__global__ __launch_bounds__(NTHREADS) void k_pf_defer(PFP f, int nreg) {
  const int k = blockIdx.x * NTHREADS + threadIdx.x;
  if (k >= nreg) return;
  RegionStat& t = f.tab[k];
  if (t.state == 1 && t.kind == 0) t.state = 3;
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_pf_setup(PFP f, KP<T> p, int* err, int only_fresh) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  const RegionStat& t = f.tab[f.reg[v]];
  // (only_fresh: running regions keep their excess; final ones drop theirs)
  if (only_fresh && !(t.state == 1 && t.fresh) && t.state != 2) return;
  long long e = 0;
  if (t.state == 1) e = t.thr - f.r[v];
  if (sizeof(T) == 4 && (e > 0x3fffffffLL || e < -0x3fffffffLL)) atomicOr(err, 1);
  p.ex[v] = (T)e;
}

template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_pf_after(PFP f, KP<T> p, int st) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  const RegionStat& t = f.tab[f.reg[v]];
  if (t.state == st) f.r[v] = t.thr - (long long)p.ex[v];
}

// regions in state `st` whose cut is solved; both parts continue in state `next`
__global__ __launch_bounds__(NTHREADS) void k_pf_split(PFP f, int nreg, int st, int next) {
  const int k = blockIdx.x * NTHREADS + threadIdx.x;
  if (k >= nreg) return;
  RegionStat& t = f.tab[k];
  if (t.state != st) return;
  if (t.cnt1 > 0 && t.cnt1 < t.cnt) {
    const int nid = atomicAdd(f.nreg, 1);
    if (nid >= f.cap) {
      atomicOr(f.err, 1);
      t.newid = -1;
      t.state = next;
      return;
    }
    t.newid = nid;
    t.state = next;
    f.tab[nid].newid = -1;
    f.tab[nid].state = next;
  } else if (t.kind == 1) {
    // the predicted gap did not split the region: mean thresholds from now on
    t.mode = 1;
    t.newid = -1;
    t.state = next;
  } else if (f.collapse) {
    // a cut at ceil(mean) that leaves the region whole: every exact value of
    // the region lies within one quantum of the mean, so the region is final
    // (the cuts that would only spread the leftover excess are skipped)
    t.newid = -2;
    t.state = 2;
  } else {
    t.newid = -1;
    t.state = next;
  }
}

// floor(s / d) of a 128-bit two's complement s and d > 0
__device__ inline long long floor_div128(unsigned long long hi, unsigned long long lo, long long d) {
  const __int128 s = (__int128)(((unsigned __int128)hi << 64) | (unsigned __int128)lo);
  long long q = (long long)floor(i128_to_double(hi, lo) / (double)d);
  for (int it = 0; it < 16; ++it) {
    const __int128 rem = s - (__int128)q * (__int128)d;
    if (rem >= 0 && rem < (__int128)d) break;
    const unsigned __int128 ur = (unsigned __int128)rem;
    long long adj = (long long)floor(i128_to_double((unsigned long long)(ur >> 64), (unsigned long long)ur) / (double)d);
    if (adj == 0) adj = rem < 0 ? -1 : 1;
    q += adj;
  }
  return q;
}

// Output value of every node: the region's mean for the regions made final
// by k_pf_split, else r.  (r itself stays the flow-consistent state, which a
// merged region resumes from.)
__global__ __launch_bounds__(NTHREADS) void k_pf_values(PFP f, long long* __restrict__ out) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  const RegionStat& t = f.tab[f.reg[v]];
  out[v] = (t.state == 2 && t.newid == -2) ? floor_div128(t.sum_hi, t.sum_lo, (long long)t.cnt) : f.r[v];
}

// source-side nodes of a split region move to the new region
__global__ __launch_bounds__(NTHREADS) void k_pf_move(PFP f, const int* __restrict__ reach, int st) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  const int k = f.reg[v];
  const RegionStat& t = f.tab[k];
  if (t.state == st && t.newid >= 0 && reach[v]) f.reg[v] = t.newid;
}

// ---- asynchronous divide and conquer: per-region completion -------------------

// thresholds in use again (setup done): no region is fresh, no slot is held
__global__ __launch_bounds__(NTHREADS) void k_pf_unfresh(PFP f, int nreg) {
  const int k = blockIdx.x * NTHREADS + threadIdx.x;
  if (k >= nreg) return;
  f.tab[k].fresh = 0;
  f.tab[k].hslot = -1;
}

// regions (state 1) with an active node (excess and a finite label)
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_pf_busy(PFP f, KP<T> p) {
  for (long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x; v < f.N; v += (long long)gridDim.x * NTHREADS) {
    if (p.ex[v] > 0 && p.h[v] < INF_H) {
      RegionStat& t = f.tab[f.reg[v]];
      if (t.state == 1 && !t.busy) t.busy = 1;
    }
  }
}

// a region with no active node after an exact relabel holds a maximum flow of
// its cut: state 4 (n counts them); busy flags are cleared for the next relabel
__global__ __launch_bounds__(NTHREADS) void k_pf_mark_done(PFP f, int nreg, unsigned long long* n) {
  const int k = blockIdx.x * NTHREADS + threadIdx.x;
  if (k >= nreg) return;
  RegionStat& t = f.tab[k];
  if (t.state != 1) return;
  if (!t.busy) {
    t.state = 4;
    atomicAdd(n, 1ull);
  }
  t.busy = 0;
}

// S_min seeds: the excess nodes of the regions in state st
template <typename T>
__global__ __launch_bounds__(NTHREADS) void k_pf_flood_init(PFP f, KP<T> p, int* __restrict__ reach, int st) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  reach[v] = (p.ex[v] > 0 && f.tab[f.reg[v]].state == st) ? 1 : 0;
}

// zero every arc whose ends lie in different regions
template <int K, typename T>
__global__ __launch_bounds__(NTHREADS) void k_pf_cutarcs(PFP f, KP<T> p) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  int z, y, x;
  node_coords(v, p.H, p.W, z, y, x);
  const int k = f.reg[v];
  MaskT<K> cut = 0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int nz = z + p.dirs.dz[i], ny = y + p.dirs.dy[i], nx = x + p.dirs.dx[i];
    if (nz < 0 || nz >= p.D || ny < 0 || ny >= p.H || nx < 0 || nx >= p.W) continue;
    const long long w = ((long long)nz * p.H + ny) * p.W + nx;
    if (f.reg[w] != k) {
      T* a = &p.res[(long long)i * p.N + v];
      if (*a != 0) {
        *a = 0;
        cut |= (MaskT<K>)1 << i;
      }
    }
  }
  if (cut) reinterpret_cast<MaskT<K>*>(p.mask)[v] &= (MaskT<K>)~cut;
}

// a region check: regions [0, nchecked) take part (state 1), the rest is final
__global__ __launch_bounds__(NTHREADS) void k_pf_init_check(PFP f, int nreg, int nchecked) {
  const int k = blockIdx.x * NTHREADS + threadIdx.x;
  if (k < nreg) f.tab[k].state = k < nchecked ? 1 : 2;
  if (k == 0) *f.nreg = nreg;
}

__global__ __launch_bounds__(NTHREADS) void k_pf_cut(PFP f, int* __restrict__ out) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  out[v] = f.r[v] < 0 ? 1 : 0;
}

// TV network of a stack of `layers` images (one lambda per layer): reduction
// 2 uf (addE1(0, uf)) and, on every in-lattice arc of direction i,
// 2 pwf[layer][i mod k/2] (addE2(0, pwf, pwf, 0)) (tv_solver.hpp:71-112).
template <int K>
__global__ __launch_bounds__(NTHREADS) void k_tv_net(int D, int H, int W, long long N, DirTab dirs,
                                                     const long long* __restrict__ uf,
                                                     const long long* __restrict__ pwf,
                                                     long long* __restrict__ red0, long long* __restrict__ alpha0) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= N) return;
  const long long plane = (long long)H * W;
  const int z = (int)(v / plane);
  const long long q = v - (long long)z * plane;
  const int y = (int)(q / W), x = (int)(q % W);
  red0[v] = 2 * uf[q];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int ny = y + dirs.dy[i], nx = x + dirs.dx[i];
    const bool in = ny >= 0 && ny < H && nx >= 0 && nx < W;
    alpha0[(long long)i * N + v] = in ? 2 * pwf[(long long)z * (K / 2) + (i % (K / 2))] : 0;
  }
  (void)D;
}

__global__ __launch_bounds__(NTHREADS) void k_pf_init_layers(PFP f, const long long* __restrict__ red0,
                                                             long long plane, int layers) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  f.r[v] = red0[v];
  f.reg[v] = (int)(v / plane);
  if (v < layers) f.tab[v].state = 1;
}

// ---- guided divide and conquer: an initial partition from predicted values ----
//
// The lattice is partitioned into regions in which neighbouring predicted
// values differ by at most delta.  Every arc (v, u) between two regions is
// taken as saturated from the side assumed lower (smaller prediction) to the
// side assumed higher: r_v += c(v->u), r_u -= c(v->u), and the arc is cut.
// The regions are then independent problems for the divide and conquer.  When
// it is done, the assumption is checked on every cut arc (final r of the
// higher end >= final r of the lower end + 2, i.e. x_high >= x_low for the
// exact level values within +-1): where it holds on every arc, the optimality
// conditions of the whole problem hold (the subgradient of w |x_v - x_u| at
// x_v <= x_u includes the one used), so the result is its unique solution.
// Regions joined by a failing arc are merged and solved again from their
// original network.

__device__ inline int uf_load(const int* p, int x) {
  return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline int uf_find(int* p, int x) {
  for (;;) {
    const int px = uf_load(p, x);
    if (px == x) return x;
    const int gx = uf_load(p, px);
    // path halving: any ancestor is a valid parent
    if (gx != px) __hip_atomic_store(p + x, gx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = gx;
  }
}

// links the larger root below the smaller one, so parents only decrease
__device__ inline void uf_union(int* p, int a, int b) {
  for (;;) {
    a = uf_find(p, a);
    b = uf_find(p, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    int expected = a;
    if (__hip_atomic_compare_exchange_strong(p + a, &expected, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return;
  }
}

struct GP {
  const float* pred;  // predicted value per node
  const long long* val;  // output value per node (k_pf_values)
  int* preg;          // partition region per node (0 .. nparts-1)
  int* parent;        // union-find over nodes, later over partition regions
  int* idmap;         // root -> compact id, later partition region -> table id
  int* merged;        // per partition region: part of a merge
  int* layer_bad;     // per layer: a failing arc
  float delta;
};

__device__ inline bool assumed_lower(const float* pred, long long v, long long u) {
  const float a = pred[v], b = pred[u];
  return a < b || (a == b && v < u);
}

__global__ __launch_bounds__(NTHREADS) void k_iota(int* a, long long n) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v < n) a[v] = (int)v;
}

template <int K>
__global__ __launch_bounds__(NTHREADS) void k_part_join(PFP f, DirTab dirs, GP g) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  const int x = (int)(v % f.W), y = (int)((v / f.W) % f.H);
  const float pv = g.pred[v];
#pragma unroll
  for (int i = 0; i < K / 2; ++i) {
    const int ny = y + dirs.dy[i], nx = x + dirs.dx[i];
    if (ny < 0 || ny >= f.H || nx < 0 || nx >= f.W) continue;
    const long long u = v + (long long)dirs.dy[i] * f.W + dirs.dx[i];
    if (fabsf(pv - g.pred[u]) <= g.delta) uf_union(g.parent, (int)v, (int)u);
  }
}

__global__ __launch_bounds__(NTHREADS) void k_uf_flatten(int* parent, long long n) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v < n) parent[v] = uf_find(parent, (int)v);
}

__global__ __launch_bounds__(NTHREADS) void k_part_ids(PFP f, GP g) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  if (g.parent[v] == (int)v) g.idmap[v] = atomicAdd(f.nreg, 1);
}

__global__ __launch_bounds__(NTHREADS) void k_part_assign(PFP f, GP g) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  const int id = g.idmap[g.parent[v]];
  g.preg[v] = id;
  f.reg[v] = id;
  f.tab[id].state = 1;
}

// r = red0 + the saturated cut arcs of the partition
template <int K, typename T>
__global__ __launch_bounds__(NTHREADS) void k_part_network(PFP f, KP<T> p, GP g, const long long* __restrict__ red0,
                                                           const long long* __restrict__ alpha0) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  const int k = g.preg[v];
  const int x = (int)(v % f.W), y = (int)((v / f.W) % f.H);
  long long r = red0[v];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int ny = y + p.dirs.dy[i], nx = x + p.dirs.dx[i];
    if (ny < 0 || ny >= f.H || nx < 0 || nx >= f.W) continue;
    const long long u = v + (long long)p.dirs.dy[i] * f.W + p.dirs.dx[i];
    if (g.preg[u] == k) continue;
    if (assumed_lower(g.pred, v, u)) r += alpha0[(long long)i * f.N + v];
    else r -= alpha0[(long long)((i + K / 2) % K) * f.N + u];
  }
  f.r[v] = r;
}

// Merged regions continue from their current state: every cut arc inside a
// merged group (a partition arc or a divide-and-conquer arc) carries the flow
// that saturated it from its lower end, so its residuals are 0 from the lower
// end and c(lo->hi) + c(hi->lo) from the higher end; r already holds that flow.
// The lower end of a partition arc is the assumed one, of a divide-and-conquer
// arc the end with the smaller final r (a nested cut puts it strictly lower).
template <int K, typename T>
__global__ __launch_bounds__(NTHREADS) void k_part_restore(PFP f, KP<T> p, GP g, const long long* __restrict__ alpha0) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  const int k = g.preg[v];
  if (!g.merged[k]) return;
  const int grp = g.parent[k];
  const int rv = f.reg[v];
  const int x = (int)(v % f.W), y = (int)((v / f.W) % f.H);
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int ny = y + p.dirs.dy[i], nx = x + p.dirs.dx[i];
    if (ny < 0 || ny >= f.H || nx < 0 || nx >= f.W) continue;
    const long long u = v + (long long)p.dirs.dy[i] * f.W + p.dirs.dx[i];
    const int ku = g.preg[u];
    if (g.parent[ku] != grp || f.reg[u] == rv) continue;
    const bool lo = ku != k ? assumed_lower(g.pred, v, u) : g.val[v] < g.val[u];
    p.res[(long long)i * f.N + v] =
        lo ? (T)0 : (T)(alpha0[(long long)i * f.N + v] + alpha0[(long long)((i + K / 2) % K) * f.N + u]);
  }
}

// failing cut arcs: union of the two partition regions (g.parent over regions)
template <int K>
__global__ __launch_bounds__(NTHREADS) void k_part_check(PFP f, DirTab dirs, GP g, unsigned long long* n_bad) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  const int x = (int)(v % f.W), y = (int)((v / f.W) % f.H);
  const int k = g.preg[v];
  int bad = 0;
#pragma unroll
  for (int i = 0; i < K / 2; ++i) {
    const int ny = y + dirs.dy[i], nx = x + dirs.dx[i];
    if (ny < 0 || ny >= f.H || nx < 0 || nx >= f.W) continue;
    const long long u = v + (long long)dirs.dy[i] * f.W + dirs.dx[i];
    const int ku = g.preg[u];
    if (ku == k) continue;
    const bool lo = assumed_lower(g.pred, v, u);
    const long long gap = lo ? g.val[u] - g.val[v] : g.val[v] - g.val[u];
    if (gap < 2) {
      uf_union(g.parent, k, ku);
      ++bad;
    }
  }
  if (bad) {
    atomicAdd(n_bad, (unsigned long long)bad);
    g.layer_bad[v / ((long long)f.H * f.W)] = 1;
  }
}

// every region of a layer with a failing arc joins the layer's first region
__global__ __launch_bounds__(NTHREADS) void k_part_merge_layers(PFP f, GP g) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  const long long plane = (long long)f.H * f.W;
  const long long z = v / plane;
  if (g.layer_bad[z]) uf_union(g.parent, g.preg[v], g.preg[z * plane]);
}

__global__ __launch_bounds__(NTHREADS) void k_part_mark(GP g, int nparts) {
  const int k = blockIdx.x * NTHREADS + threadIdx.x;
  if (k >= nparts) return;
  const int root = g.parent[k];
  if (root != k) {
    g.merged[k] = 1;
    atomicOr(&g.merged[root], 1);
  }
}

// fresh table ids for the merged regions (roots of the region union-find)
__global__ __launch_bounds__(NTHREADS) void k_part_newids(PFP f, GP g, int nparts) {
  const int k = blockIdx.x * NTHREADS + threadIdx.x;
  if (k >= nparts) return;
  if (!g.merged[k] || g.parent[k] != k) return;
  const int nid = atomicAdd(f.nreg, 1);
  if (nid >= f.cap) {
    atomicOr(f.err, 1);
    g.idmap[k] = 0;
    return;
  }
  g.idmap[k] = nid;
  f.tab[nid].state = 1;
}

// nodes of merged regions join their group (table id from k_part_newids)
__global__ __launch_bounds__(NTHREADS) void k_part_relabel(PFP f, GP g) {
  const long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x;
  if (v >= f.N) return;
  const int k = g.preg[v];
  if (!g.merged[k]) return;
  const int grp = g.parent[k];
  g.preg[v] = grp;
  f.reg[v] = g.idmap[grp];
}

__global__ __launch_bounds__(NTHREADS) void k_max_label(const int* __restrict__ h, long long N, long long* part) {
  long long m = 0;
  for (long long v = (long long)blockIdx.x * NTHREADS + threadIdx.x; v < N; v += (long long)gridDim.x * NTHREADS) {
    const int x = h[v];
    if (x < INF_H && x > m) m = x;
  }
  m = block_reduce<true>(m);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + NTHREADS - 1) / NTHREADS); }

// ---- engine -----------------------------------------------------------------------

struct TileGrid {
  int ntx = 0, nty = 0, ntz = 0, ntiles = 0, nblocks = 0, nt = 0;
};

template <class G>
TileGrid make_grid(const LatticeGeom& geo) {
  TileGrid t;
  t.ntx = (int)((geo.W + G::TW - 1) / G::TW);
  t.nty = (int)((geo.H + G::TH - 1) / G::TH);
  t.ntz = (int)((geo.D + G::TD - 1) / G::TD);
  const long long n = (long long)t.ntx * t.nty * t.ntz;
  if (n > (1LL << 30)) fail(LF_ERR_BAD_ARG, "lattice too large");
  t.ntiles = (int)n;
  t.nblocks = (int)(((n + 7) / 8) * 8);
  t.nt = G::NT;
  return t;
}

// G0..G2: the push-tile shapes the tuner chooses from (same halo radius).
template <int K, class G0, class G1, class G2, class B, typename T, typename MT, class DIR>
class Engine final : public MinCutEngine {
 public:
  Engine(const LatticeGeom& geo) : g_(geo) {
    // the push kernel addresses a plane by 32-bit byte offsets
    if ((unsigned long long)geo.n() * std::max<size_t>(sizeof(T), 4) >= (1ull << 32))
      fail(LF_ERR_UNSUPPORTED, "lattice too large for one device engine: nodes x element bytes must stay below 4 GiB");
    LF_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    N_ = geo.n();
    p_.D = (int)geo.D; p_.H = (int)geo.H; p_.W = (int)geo.W; p_.N = N_;
    grid_[0] = make_grid<G0>(geo);
    grid_[1] = make_grid<G1>(geo);
    grid_[2] = make_grid<G2>(geo);
    ntiles_ = std::max(grid_[0].ntiles, std::max(grid_[1].ntiles, grid_[2].ntiles));
    set_grid(0);
    bp_.D = p_.D; bp_.H = p_.H; bp_.W = p_.W; bp_.N = N_;
    bp_.ntx = (int)((geo.W + B::TW - 1) / B::TW);
    bp_.nty = (int)((geo.H + B::TH - 1) / B::TH);
    bp_.ntz = (int)((geo.D + B::TD - 1) / B::TD);
    bntiles_ = (long long)bp_.ntx * bp_.nty * bp_.ntz;
    bp_.ntiles = (int)bntiles_;
    bp_.nblocks = (int)(((bntiles_ + 7) / 8) * 8);
    for (int i = 0; i < K; ++i) {
      p_.dirs.dz[i] = (int8_t)geo.dz[i];
      p_.dirs.dy[i] = (int8_t)geo.dy[i];
      p_.dirs.dx[i] = (int8_t)geo.dx[i];
    }
    bp_.dirs = p_.dirs;
    LF_HIP(hipMalloc(&red0_, sizeof(long long) * N_));
    LF_HIP(hipMalloc(&alpha0_, sizeof(long long) * N_ * K));
    LF_HIP(hipMalloc(&res_, sizeof(T) * N_ * K));
    LF_HIP(hipMalloc(&mbox_, sizeof(T) * N_ * K * 2));  // one mailbox per round parity
    LF_HIP(hipMalloc(&ex_, sizeof(T) * N_));
    LF_HIP(hipMalloc(&h_, sizeof(int) * N_));
    LF_HIP(hipMalloc(&mask_, sizeof(MT) * N_));
    LF_HIP(hipMalloc(&reach_, sizeof(int) * N_));
    LF_HIP(hipMalloc(&flags_[0], sizeof(int) * ntiles_));
    LF_HIP(hipMalloc(&flags_[1], sizeof(int) * ntiles_));
    LF_HIP(hipMalloc(&wl_[0], sizeof(int) * ntiles_));
    LF_HIP(hipMalloc(&wl_[1], sizeof(int) * ntiles_));
    LF_HIP(hipMalloc(&wl_cnt_, 4 * sizeof(unsigned)));
    LF_HIP(hipMalloc(&mail_, sizeof(unsigned) * ntiles_ * 2));
    LF_HIP(hipMemsetAsync(mail_, 0, sizeof(unsigned) * ntiles_ * 2, s_));
    LF_HIP(hipMalloc(&dmail_, sizeof(unsigned) * ntiles_));
    LF_HIP(hipMemsetAsync(dmail_, 0, sizeof(unsigned) * ntiles_, s_));
    LF_HIP(hipMemsetAsync(wl_cnt_, 0, 4 * sizeof(unsigned), s_));
    {
      // push rounds run as many blocks as the device holds at once
      int dev = 0, cus = 0;
      LF_HIP(hipGetDevice(&dev));
      LF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      const void* fn[3] = {reinterpret_cast<const void*>(&k_push<K, G0, T, DIR>),
                           reinterpret_cast<const void*>(&k_push<K, G1, T, DIR>),
                           reinterpret_cast<const void*>(&k_push<K, G2, T, DIR>)};
      for (int v = 0; v < 3; ++v) {
        int nb = 0;
        LF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn[v], NTHREADS, 0));
        wl_grid_[v] = std::max(1, nb) * std::max(1, cus);
      }
    }
    LF_HIP(hipMalloc(&bflags_[0], sizeof(int) * bntiles_));
    LF_HIP(hipMalloc(&bflags_[1], sizeof(int) * bntiles_));
    LF_HIP(hipMalloc(&dev_small_, 64));
    LF_HIP(hipMalloc(&part_, sizeof(long long) * 2 * RED_BLOCKS));
    if (bits_) {
      wp_.D = p_.D; wp_.H = p_.H; wp_.W = p_.W; wp_.ntx = bp_.ntx; wp_.nty = bp_.nty;
      wp_.ps = (long long)p_.D * bp_.nty * bp_.ntx * WT;
      LF_HIP(hipMalloc(&planes_, sizeof(unsigned long long) * wp_.ps * WPLANES));
      LF_HIP(hipMemsetAsync(planes_, 0, sizeof(unsigned long long) * wp_.ps * WPLANES, s_));
    }
    LF_HIP(hipHostMalloc(&host_small_, 128, hipHostMallocDefault));
    LF_HIP(hipEventCreate(&ev_[0]));
    LF_HIP(hipEventCreate(&ev_[1]));
    LF_HIP(hipEventCreate(&evb_[0]));
    LF_HIP(hipEventCreate(&evb_[1]));
    LF_HIP(hipMalloc(&bvisits_, 8 * sizeof(unsigned long long)));
    reset_energies();
  }
  ~Engine() override {
    hipStreamSynchronize(s_);
    hipFree(red0_); hipFree(alpha0_); hipFree(res_); hipFree(mbox_); hipFree(ex_); hipFree(h_);
    hipFree(mask_); hipFree(reach_); hipFree(flags_[0]); hipFree(flags_[1]); hipFree(bflags_[0]); hipFree(bflags_[1]); hipFree(dev_small_); hipFree(part_);
    hipFree(planes_); hipFree(wl_[0]); hipFree(wl_[1]); hipFree(wl_cnt_); hipFree(mail_); hipFree(dmail_);
    hipHostFree(host_small_);
    hipFree(pf_r_); hipFree(pf_reg_); hipFree(pf_tab_); hipFree(pf_nreg_); hipFree(pf_cnt_); hipFree(pf_hist_);
    hipFree(b_rects_); hipFree(b_blk_); hipFree(b_flows_);
    hipFree(g_preg_); hipFree(g_parent_); hipFree(g_idmap_); hipFree(g_merged_); hipFree(g_layer_bad_); hipFree(g_val_);
    hipEventDestroy(ev_[0]);
    hipEventDestroy(ev_[1]);
    hipEventDestroy(evb_[0]);
    hipEventDestroy(evb_[1]);
    hipFree(bvisits_);
    for (auto& kv : graphs_) hipGraphExecDestroy(kv.second);
    hipStreamDestroy(s_);
  }

  hipStream_t stream() const override { return s_; }

  void set_debug_checks(bool on) override { debug_checks_ = on; }

  // Host copies of the working state checked against the base network
  // (lf_verify.hpp).  Parametric solves change excesses and cut arcs between
  // regions, so there only residual signs, masks and labels are checked.
  void debug_check(bool labels_exact, const char* where) {
    if (!debug_checks_) return;
    LF_HIP(hipStreamSynchronize(s_));
    const size_t nk = (size_t)N_ * K;
    std::vector<T> res(nk), ex(N_), mb(2 * nk);
    std::vector<int32_t> h(N_);
    std::vector<MT> mk(N_);
    std::vector<int64_t> a0(nk), r0(N_);
    LF_HIP(hipMemcpy(res.data(), res_, sizeof(T) * nk, hipMemcpyDeviceToHost));
    LF_HIP(hipMemcpy(ex.data(), ex_, sizeof(T) * N_, hipMemcpyDeviceToHost));
    LF_HIP(hipMemcpy(mb.data(), mbox_, sizeof(T) * 2 * nk, hipMemcpyDeviceToHost));
    LF_HIP(hipMemcpy(h.data(), h_, sizeof(int32_t) * N_, hipMemcpyDeviceToHost));
    LF_HIP(hipMemcpy(mk.data(), mask_, sizeof(MT) * N_, hipMemcpyDeviceToHost));
    LF_HIP(hipMemcpy(a0.data(), alpha0_, sizeof(int64_t) * nk, hipMemcpyDeviceToHost));
    LF_HIP(hipMemcpy(r0.data(), red0_, sizeof(int64_t) * N_, hipMemcpyDeviceToHost));
    std::string err;
    if (pf_mode_) {
      for (size_t j = 0; j < nk && err.empty(); ++j)
        if (res[j] < 0) err = "negative residual at plane entry " + std::to_string(j);
      for (int64_t v = 0; v < N_ && err.empty() && mask_valid_; ++v)
        for (int i = 0; i < K; ++i)
          if ((((unsigned long long)mk[v] >> i) & 1ull) != (res[(size_t)i * N_ + v] > 0 ? 1ull : 0ull)) {
            err = "arc mask at node " + std::to_string(v) + " direction " + std::to_string(i);
            break;
          }
    } else {
      VerifyView vv;
      vv.D = p_.D; vv.H = p_.H; vv.W = p_.W; vv.K = K;
      vv.dz = g_.dz.data(); vv.dy = g_.dy.data(); vv.dx = g_.dx.data();
      vv.red0 = r0.data(); vv.alpha0 = a0.data();
      vv.cap_bytes = (int)sizeof(T);
      vv.res = res.data(); vv.ex = ex.data(); vv.mbox = mb.data(); vv.mbox_planes = 2;
      vv.mask = mask_valid_ ? mk.data() : nullptr;
      vv.mask_bytes = (int)sizeof(MT);
      vv.h = labels_exact ? h.data() : nullptr;
      vv.labels_exact = labels_exact;
      err = verify_state(vv);
    }
    if (!err.empty()) fail(LF_ERR_STATE, std::string("state check after ") + where + ": " + err);
  }

  // Push tiling of a solve: the tuning's tile shape; else, for parametric and
  // TV networks on 2-D radius-1 lattices, the 16x16 tiles, whose rounds are
  // the shortest (their long tails of rounds over few tiles are latency: the
  // 32-lambda 1024^2 TV path took 3.63 s on them, 3.99 s on 16x32, 5.6 s on
  // 32x32); else G0.
  int variant_for(const lf_tuning& tun) const {
    if (tun.tile_shape >= 1 && tun.tile_shape <= 3) return tun.tile_shape - 1;
    if (pf_mode_ && G1::TD == 1 && G1::TH == 16 && G1::TW == 16) return 1;
    return 0;
  }

  void set_grid(int v) {
    variant_ = v;
    p_.ntx = grid_[v].ntx;
    p_.nty = grid_[v].nty;
    p_.ntz = grid_[v].ntz;
    p_.ntiles = grid_[v].ntiles;
    p_.nblocks = grid_[v].nblocks;
  }
  template <class F>
  void with_geo(F&& f) {
    switch (variant_) {
      case 1: f(G1{}); break;
      case 2: f(G2{}); break;
      default: f(G0{}); break;
    }
  }

  void reset_energies() override {
    tv_mode_ = false;
    LF_HIP(hipMemsetAsync(red0_, 0, sizeof(long long) * N_, s_));
    LF_HIP(hipMemsetAsync(alpha0_, 0, sizeof(long long) * N_ * K, s_));
    solved = false;
  }

  void add_energies_device(const int64_t* du, const int64_t* db, hipStream_t s) override {
    wait_for(s);
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

  // working state (res, ex, zero mailbox) from the base network, with range checks
  // defer_check: the capacity checks are taken from the first global
  // relabel's read instead of a synchronisation of their own
  void load_working(bool defer_check = false) {
    KP<T> p = params();
    int* err = reinterpret_cast<int*>(dev_small_);
    unsigned long long* mx = reinterpret_cast<unsigned long long*>(dev_small_ + 32);
    LF_HIP(hipMemsetAsync(dev_small_, 0, 64, s_));
    const int lb = red_blocks();
    // a mailbox that every solve since the last clearing delivered is empty
    hipLaunchKernelGGL((k_load<K, T>), dim3(lb), dim3(NTHREADS), 0, s_, p, red0_, alpha0_, err, part_,
                       mbox_dirty_ ? 1 : 0);
    if (mbox_dirty_)
      hipLaunchKernelGGL((k_fill<unsigned>), dim3(blocks_for(ntiles_)), dim3(NTHREADS), 0, s_, dmail_, 0u, ntiles_);
    mbox_dirty_ = false;
    mask_valid_ = true;  // k_load wrote every node's arc mask
    hipLaunchKernelGGL((k_reduce<true>), dim3(1), dim3(NTHREADS), 0, s_, part_, lb, 2, 0, mx);
    hipLaunchKernelGGL((k_reduce<true>), dim3(1), dim3(NTHREADS), 0, s_, part_, lb, 2, 1, mx + 1);
    LF_HIP(hipGetLastError());
    if (defer_check) {
      load_check_pending_ = true;
      return;
    }
    LF_HIP(hipMemcpyAsync(host_small_, dev_small_, 64, hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    check_load(host_small_);
  }

  // Capacity checks of a load from a host copy of dev_small_[0, 64) taken
  // after k_load and its reductions: the error word at 0 (bit 1: negative
  // pairwise capacity) and the largest node and arc capacities at 32 and 40.
  // Until the first global relabel only the counter slot, the tile counter and
  // the active count (16, 24, 48) change.
  void check_load(const void* hs) {
    const int e = static_cast<const int*>(hs)[0];
    const unsigned long long* m = reinterpret_cast<const unsigned long long*>(static_cast<const char*>(hs) + 32);
    if (e & 2) fail(LF_ERR_NONSUBMODULAR, "negative pairwise capacity");
    max_red_ = m[0];
    max_alpha_ = m[1];
    const double bound = (double)m[0] + (double)K * (double)m[1] + 2.0 * (double)m[1];
    const double lim = sizeof(T) == 4 ? 2147483647.0 : 9.2e18;
    if (bound >= lim)
      fail(LF_ERR_OVERFLOW, sizeof(T) == 4 ? "capacities exceed the int32 range; use cap_bytes=8"
                                           : "capacities exceed the int64 range");
  }

  // Maximum preflow of the current working state (res, ex), then S_min.
  // Accumulates into stats.
  void solve_state(const lf_tuning& tun_in) {
    using clk = std::chrono::steady_clock;
    // LF_ROUNDS / LF_ITERS / LF_TILE_SHAPE fill tuning fields left at 0 (for studies)
    lf_tuning tun = tun_in;
    if (tun.rounds_per_relabel <= 0) tun.rounds_per_relabel = (int)env_double("LF_ROUNDS", 0);
    if (tun.local_iters <= 0) tun.local_iters = (int)env_double("LF_ITERS", 0);
    if (tun.tile_shape <= 0) tun.tile_shape = (int)env_double("LF_TILE_SHAPE", 0);
    if (tun.relabel_launches <= 0) tun.relabel_launches = (int)env_double("LF_RELABEL_CAP", 0);
    KP<T> p = params();
    // Fixed tuning when given; otherwise the number of push rounds between two
    // exact relabels follows the measured cost ratio of a relabel and a round
    // (a round over few active tiles is cheap, a relabel always covers the
    // lattice; pushing takes push_relabel_ratio() times the relabel's time),
    // pushing stops early once no tile is active, and the local iterations
    // per round grow when few nodes are active.
    const bool fixed_r = tun.rounds_per_relabel > 0;
    int R = fixed_r ? tun.rounds_per_relabel : 4;
    // a solve never switches tilings midway: the mailbox is empty here
    set_grid(variant_for(tun));
    p = params();
    bool pushed = false;
    long long cycles = 0;
    // Intermediate relabels stop after relabel_cap launches; the stopping
    // decision is always taken on an exact (uncapped) relabel.  Labels beyond
    // a short relabel's reach stay INF, which only idles their nodes until a
    // later relabel: min-cut networks with sinks everywhere measured fastest
    // at 2 launches (2048^2 segmentation 5.3 -> 2.2 ms, 256^3 blobs 31 -> 8.5
    // ms), parametric / TV networks, whose flow travels farther, at 4 (the
    // 32-lambda 1024^2 TV path on 16x16 tiles: 2 launches 3.75 s, 3-6 launches
    // 3.08-3.29 s, 8 3.27-3.32 s, 16 3.29-3.35 s).  A min-cut
    // solve that keeps going is one whose flow travels far too: its reach
    // doubles every 4 cycles up to 64 launches (256^3 unthresholded noise
    // volume: 2.7 s at a fixed 2 launches, 0.66-0.85 s at 8-16).
    const bool fixed_cap = tun.relabel_launches > 0;
    long long relabel_cap = fixed_cap ? tun.relabel_launches : (pf_mode_ ? 4 : 2);
    bool need_exact = false, resume = false, push_timed_later = false;
    for (;;) {
      const auto tb = clk::now();
      const long long bl0 = stats.bfs_launches;
      bool exact = false;
      const long long active = global_relabel(p, pushed, need_exact ? 0 : relabel_cap, &exact, resume);
      debug_check(exact, "a global relabel");
      resume = false;
      pushed = false;  // the mailbox was delivered
      double t_relabel = std::chrono::duration<double>(clk::now() - tb).count();
      if (push_timed_later) {
        // the relabel synchronised the stream: the rounds before it are done
        float ms = 0;
        LF_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
        stats.push_seconds += ms * 1e-3;
        t_relabel = std::max(0.0, t_relabel - ms * 1e-3);
        push_timed_later = false;
      }
      stats.bfs_seconds += t_relabel;
      ++stats.cycles;
      ++cycles;
      if (!fixed_cap && !pf_mode_ && cycles % 4 == 0 && relabel_cap < 64) relabel_cap *= 2;
      if (trace_) {
        hipLaunchKernelGGL((k_max_label), dim3(red_blocks()), dim3(NTHREADS), 0, s_, h_, N_, part_);
        hipLaunchKernelGGL((k_reduce<true>), dim3(1), dim3(NTHREADS), 0, s_, part_, red_blocks(), 1, 0, counter_ptr());
        const unsigned long long mxl = read_counter();
        std::fprintf(stderr, "[lf] cycle %lld active %lld bfs_launches %lld max_label %llu tiles_pushed %llu R %d\n",
                     cycles, active, stats.bfs_launches - bl0, mxl, (unsigned long long)read_tiles(), R);
      }
      if (active == 0) {
        if (exact) break;
        // confirm before stopping: carry the capped pass on to its fixpoint
        need_exact = true;
        resume = true;
        continue;
      }
      // once a confirmation finds only a few active nodes left, the last
      // cycles relabel exactly (a capped relabel would need a confirmation again)
      need_exact = need_exact && active * 1024 < N_;
      if (tun.max_cycles > 0 && cycles > tun.max_cycles)
        fail(LF_ERR_STATE, "max_cycles reached before the flow became maximum");
      p.iters = tun.local_iters > 0 ? tun.local_iters : (active * 64 < N_ ? 16 : 4);
      // adaptive schedule: chunks of 4, 8, then 16 rounds while tiles have
      // work (one host read per chunk; R is a multiple of 4)
      int chunk = fixed_r ? R : 4;
      int done_rounds = 0;
      double t_push = 0;
      unsigned long long tiles0 = tiles_seen_;  // no round since the relabel read it
      while (done_rounds < R) {
        LF_HIP(hipEventRecord(ev_[0], s_));
        const int n = std::min(chunk, R - done_rounds);
        push_rounds(p, n);
        LF_HIP(hipGetLastError());
        LF_HIP(hipEventRecord(ev_[1], s_));
        pushed = true;
        debug_check(false, "push rounds");
        if (fixed_r && done_rounds + n >= R) {
          // the cycle's last chunk of a fixed schedule: its time is read after
          // the next relabel's synchronisation, so the host does not wait here
          done_rounds += n;
          push_timed_later = true;
          break;
        }
        // one synchronisation: the tile counter (adaptive schedule) or the event
        unsigned long long tiles1 = 0;
        if (fixed_r)
          LF_HIP(hipEventSynchronize(ev_[1]));
        else
          tiles1 = read_tiles();
        float ms = 0;
        LF_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
        t_push += ms * 1e-3;
        done_rounds += n;
        if (!fixed_r) {
          if (tiles1 == tiles0) break;  // no tile had work
          tiles0 = tiles1;
          chunk = std::min(chunk * 2, 16);
        }
      }
      stats.push_seconds += t_push;
      if (!fixed_r) {
        const double per_round = t_push / std::max(done_rounds, 1);
        const double want = per_round > 0 ? push_relabel_ratio() * t_relabel / per_round : 4.0;
        R = 4 * (int)std::max(1.0, std::min(64.0, std::round(want / 4)));
      }
    }
    const auto tf = clk::now();
    hipLaunchKernelGGL((k_flood_init<T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, p, reach_);
    run_tiled([&](const BP& q) {
      launch_flood(q);
      ++stats.flood_launches;
    }, 0, false, nullptr, false, FLOOD_FIRST_BATCH);
    stats.flood_seconds += std::chrono::duration<double>(clk::now() - tf).count();
  }

  void begin_stats() {
    stats = lf_stats{};
    bfs_time_pending_ = false;
    fill_many({{u32(tiles_ptr()), 0u, 2}, {u32(bvisits_), 0u, 16}});
  }
  // One read for the statistics: the small device words (the counter slot,
  // which holds the last reduction, and the processed-tile counter) into
  // host_small_[0, 64) and the BFS visit counters into [64, 128).
  void end_stats() {
    stats.tile_nodes = grid_[variant_].nt;
    stats.bytes_per_node = 2 * ((long long)K * sizeof(T) + sizeof(T) + 4);
    stats.bfs_bytes_per_visit = (long long)K * sizeof(T) + 4LL * K + 4;
    char* hb = static_cast<char*>(host_small_);
    LF_HIP(hipMemcpyAsync(hb, dev_small_, 64, hipMemcpyDeviceToHost, s_));
    LF_HIP(hipMemcpyAsync(hb + 64, bvisits_, 64, hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    const unsigned long long* w = reinterpret_cast<const unsigned long long*>(hb);
    stats.tiles_pushed = (long long)w[3];  // tiles_ptr(): dev_small_ + 24
    stats.bfs_visits = 0;
    for (int i = 0; i < 8; ++i) stats.bfs_visits += (long long)w[8 + i];
  }

  void run(const lf_tuning& tun) override {
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    begin_stats();
    load_working(true);
    solve_state(tun);
    KP<T> p = params();
    hipLaunchKernelGGL((k_flow<T>), dim3(red_blocks()), dim3(NTHREADS), 0, s_, p, red0_, part_);
    hipLaunchKernelGGL((k_reduce<false>), dim3(1), dim3(NTHREADS), 0, s_, part_, red_blocks(), 1, 0, counter_ptr());
    end_stats();  // its read brings the flow too (the counter slot, dev_small_ + 16)
    flow = (long long)reinterpret_cast<const unsigned long long*>(host_small_)[2];
    stats.seconds = std::chrono::duration<double>(clk::now() - t0).count();
    solved = true;
    pf_solved = false;
  }

  // ---- parametric divide and conquer (LatticeLevelReductions / TV) ----------
  void pf_run(const lf_tuning& tun) override { pf_run_layers(tun, 1); }

  void set_tv_network(const int64_t* uf, const int64_t* pwf, int layers) override {
    tv_mode_ = true;
    const long long plane = N_ / layers;
    long long *duf = nullptr, *dpwf = nullptr;
    LF_HIP(hipMallocAsync((void**)&duf, sizeof(long long) * plane, s_));
    LF_HIP(hipMallocAsync((void**)&dpwf, sizeof(long long) * layers * (K / 2), s_));
    LF_HIP(hipMemcpyAsync(duf, uf, sizeof(long long) * plane, hipMemcpyHostToDevice, s_));
    LF_HIP(hipMemcpyAsync(dpwf, pwf, sizeof(long long) * layers * (K / 2), hipMemcpyHostToDevice, s_));
    hipLaunchKernelGGL((k_tv_net<K>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, p_.D, p_.H, p_.W, N_, p_.dirs,
                       duf, dpwf, red0_, alpha0_);
    LF_HIP(hipGetLastError());
    LF_HIP(hipFreeAsync(duf, s_));
    LF_HIP(hipFreeAsync(dpwf, s_));
    LF_HIP(hipStreamSynchronize(s_));
    solved = false;
    pf_solved = false;
  }

  void pf_run_layers(const lf_tuning& tun, int layers) override {
    PfMode pm{pf_mode_};
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    begin_stats();
    load_working();
    ensure_pf_buffers(N_);
    PFP pf = pf_params();
    clear_table();
    zero_int(pf.err, 1);
    // one region per layer
    hipLaunchKernelGGL((k_pf_init_layers), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, red0_, N_ / layers, layers);
    hipLaunchKernelGGL((k_fill<int>), dim3(1), dim3(NTHREADS), 0, s_, pf.nreg, layers, 1LL);
    long long levels = 0, sections = 0;
    dnc_loop(tun, levels, sections);
    if (read_int(pf.err)) fail(LF_ERR_STATE, "region table overflow in the divide and conquer");
    LF_HIP(hipStreamSynchronize(s_));
    end_stats();
    stats.reserved = levels;
    pf_levels = levels;
    pf_sections = sections;
    stats.seconds = std::chrono::duration<double>(clk::now() - t0).count();
    pf_solved = true;
    solved = false;
  }

  void pf_run_layers_guided(const lf_tuning& tun, int layers, const float* d_pred, float delta, int max_repairs,
                            double pscale, double poff, lf_guided_stats* gs) override {
    PfMode pm{pf_mode_};
    struct Reset {
      GuideParams& g;
      ~Reset() { g.pred = nullptr; }
    } reset_guide{guide_};
    guide_.pred = d_pred;
    guide_.pscale = pscale;
    guide_.poff = poff;
    // threshold choice: the emptiest of 64 bins of the predicted values with at
    // least 40% of the region on each side, when the two bins around it hold
    // at most 1% of the region (LF_GUIDE_* override these for studies).  On the
    // 32-lambda 1024^2 TV path with the shipped net: 1.00x the plain path's
    // time; a quarter on each side and 5% gaps (unbalanced cuts, taken on weaker
    // gaps) 0.91-0.97x, 20% gaps 0.85x
    guide_.min_range = (long long)(env_double("LF_GUIDE_RANGE", 0.02) * pscale);
    guide_.gap_frac = env_double("LF_GUIDE_GAP", 0.01);
    guide_.min_side = env_double("LF_GUIDE_SIDE", 0.4);
    guide_.nbins = (int)env_double("LF_GUIDE_BINS", 64);
    guide_.defer = env_double("LF_GUIDE_DEFER", 0) != 0;
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    lf_guided_stats g0{};
    if (N_ >= (1LL << 31) - 1) fail(LF_ERR_UNSUPPORTED, "guided divide and conquer needs fewer than 2^31 nodes");
    begin_stats();
    load_working();
    ensure_pf_buffers(2 * N_ + 1024);
    ensure_guide_buffers(layers);
    KP<T> p = params();
    PFP pf = pf_params();
    GP g = guide_params(d_pred, delta);
    clear_table();
    zero_int(pf.err, 1);
    // partition: components of the arcs whose predicted values differ by <= delta
    hipLaunchKernelGGL((k_iota), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, g.parent, N_);
    hipLaunchKernelGGL((k_part_join<K>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, p_.dirs, g);
    hipLaunchKernelGGL((k_uf_flatten), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, g.parent, N_);
    zero_int(pf.nreg, 1);
    hipLaunchKernelGGL((k_part_ids), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, g);
    hipLaunchKernelGGL((k_part_assign), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, g);
    hipLaunchKernelGGL((k_part_network<K, T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, p, g, red0_, alpha0_);
    hipLaunchKernelGGL((k_pf_cutarcs<K, T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, p);
    LF_HIP(hipGetLastError());
    const int nparts = read_int(pf.nreg);
    g0.regions = nparts;
    const double t_part = std::chrono::duration<double>(clk::now() - t0).count();
    long long levels = 0, sections = 0;
    bool overflow = false;
    for (int repair = 0;; ++repair) {
      dnc_loop(tun, levels, sections);
      if (read_int(pf.err)) {
        overflow = true;
        break;
      }
      // check every cut arc; failing arcs join their regions
      hipLaunchKernelGGL((k_pf_values), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, g_val_);
      hipLaunchKernelGGL((k_iota), dim3(blocks_for(nparts)), dim3(NTHREADS), 0, s_, g.parent, (long long)nparts);
      zero_int(g.layer_bad, layers);
      zero_counter();
      hipLaunchKernelGGL((k_part_check<K>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, p_.dirs, g, counter_ptr());
      LF_HIP(hipGetLastError());
      const long long bad = (long long)read_counter();
      if (trace_pf_) std::fprintf(stderr, "[lf-pf] check %d: %lld failing arcs\n", repair, bad);
      if (bad == 0) break;
      g0.failing_arcs += bad;
      ++g0.repairs;
      if (repair >= max_repairs) {
        // give up on the partition of every layer that still fails
        hipLaunchKernelGGL((k_part_merge_layers), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, g);
        ++g0.layers_merged;
      }
      hipLaunchKernelGGL((k_uf_flatten), dim3(blocks_for(nparts)), dim3(NTHREADS), 0, s_, g.parent, (long long)nparts);
      zero_int(g.merged, nparts);
      hipLaunchKernelGGL((k_part_mark), dim3(blocks_for(nparts)), dim3(NTHREADS), 0, s_, g, nparts);
      hipLaunchKernelGGL((k_part_newids), dim3(blocks_for(nparts)), dim3(NTHREADS), 0, s_, pf, g, nparts);
      hipLaunchKernelGGL((k_part_restore<K, T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, p, g, alpha0_);
      mask_valid_ = false;  // restored cut arcs: the next relabel rebuilds the masks
      hipLaunchKernelGGL((k_part_relabel), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, g);
      LF_HIP(hipGetLastError());
      if (read_int(pf.err)) {
        overflow = true;
        break;
      }
    }
    if (overflow) {
      // region table full: solve without the partition
      g0.fallback = 1;
      const lf_guided_stats keep = g0;
      guide_.pred = nullptr;
      pf_run_layers(tun, layers);
      if (gs) *gs = keep;
      return;
    }
    LF_HIP(hipStreamSynchronize(s_));
    end_stats();
    stats.reserved = levels;
    pf_levels = levels;
    pf_sections = sections;
    stats.seconds = std::chrono::duration<double>(clk::now() - t0).count();
    g0.partition_seconds = t_part;
    if (gs) *gs = g0;
    pf_solved = true;
    solved = false;
  }

  // Region statistics and thresholds (setToMeanReductionPoint) of the regions
  // in state st; they become active (state 1, fresh) or final.  With
  // predictions, regions with a gap in their predicted values cut there; with
  // `defer`, mean cuts then wait for a later depth.  Returns the number of
  // regions that cut (and in *predicted those at a predicted gap).
  long long thresholds(PFP& pf, int nreg, int st, bool defer, long long* predicted) {
    hipLaunchKernelGGL((k_pf_clear), dim3(blocks_for(nreg)), dim3(NTHREADS), 0, s_, pf, nreg, st);
    hipLaunchKernelGGL((k_pf_stats_impl<0>), dim3(red_blocks()), dim3(NTHREADS), 0, s_, pf, (const int*)nullptr, st);
    if (pf.pred)
      hipLaunchKernelGGL((k_pf_stats_impl<2>), dim3(red_blocks()), dim3(NTHREADS), 0, s_, pf, (const int*)nullptr, st);
    hipLaunchKernelGGL((k_fill<unsigned long long>), dim3(1), dim3(NTHREADS), 0, s_, pf_cnt_, 0ull, 4LL);
    hipLaunchKernelGGL((k_pf_thresh), dim3(blocks_for(nreg)), dim3(NTHREADS), 0, s_, pf, nreg, pf_cnt_,
                       guide_.min_range, st);
    LF_HIP(hipGetLastError());
    unsigned long long cnt[4];
    read_u64(pf_cnt_, cnt, 2);
    long long active = (long long)cnt[0];
    *predicted = 0;
    if (active > 0 && cnt[1] > 0) {
      ensure_hist((long long)cnt[1] * pf.nbins);
      pf.hist = pf_hist_;
      hipLaunchKernelGGL((k_fill<unsigned>), dim3(blocks_for((long long)cnt[1] * pf.nbins)), dim3(NTHREADS), 0, s_,
                         pf_hist_, 0u, (long long)cnt[1] * pf.nbins);
      hipLaunchKernelGGL((k_pf_hist), dim3(red_blocks()), dim3(NTHREADS), 0, s_, pf);
      hipLaunchKernelGGL((k_pf_pick), dim3(blocks_for(nreg)), dim3(NTHREADS), 0, s_, pf, nreg, guide_.gap_frac,
                         guide_.min_side, pf_cnt_ + 2);
      LF_HIP(hipGetLastError());
      read_u64(pf_cnt_ + 2, cnt + 2, 1);
      *predicted = (long long)cnt[2];
      if (defer && *predicted > 0) {
        hipLaunchKernelGGL((k_pf_defer), dim3(blocks_for(nreg)), dim3(NTHREADS), 0, s_, pf, nreg);
        active = *predicted;
      }
    }
    return active;
  }

  // excess t - r of the regions that cut (all nodes, or only the fresh regions)
  void setup_excess(const PFP& pf, const KP<T>& p, int only_fresh, int nreg) {
    int* err = reinterpret_cast<int*>(dev_small_);
    LF_HIP(hipMemsetAsync(err, 0, sizeof(int), s_));
    hipLaunchKernelGGL((k_pf_setup<T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, p, err, only_fresh);
    hipLaunchKernelGGL((k_pf_unfresh), dim3(blocks_for(nreg)), dim3(NTHREADS), 0, s_, pf, nreg);
    LF_HIP(hipMemcpyAsync(host_small_, err, sizeof(int), hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    if (*reinterpret_cast<int*>(host_small_))
      fail(LF_ERR_OVERFLOW, "parametric offsets exceed the capacity width; use cap_bytes=8");
  }

  // Depths of the divide and conquer until no region is active: every region
  // of a depth is cut by one lattice min-cut (the reference's order).
  void dnc_loop(const lf_tuning& tun, long long& levels, long long& sections) {
    if (tv_mode_ && async_) return dnc_async(tun, levels, sections);
    using clk = std::chrono::steady_clock;
    KP<T> p = params();
    PFP pf = pf_params();
    for (;;) {
      const int nreg = (int)std::min<long long>(read_int(pf.nreg), pf_cap_);
      long long predicted = 0;
      const long long active = thresholds(pf, nreg, 1, guide_.defer, &predicted);
      if (active == 0) break;
      ++levels;
      sections += active;
      // section min-cut of every active region at once
      setup_excess(pf, p, 0, nreg);
      const lf_stats st0 = stats;
      const auto td = clk::now();
      solve_state(tun);
      if (trace_pf_)
        std::fprintf(stderr, "[lf-pf] depth %lld regions %d active %lld predicted %lld solve %.4fs cycles %lld rounds %lld bfs %lld push %.4fs\n",
                     levels, nreg, active, predicted, std::chrono::duration<double>(clk::now() - td).count(),
                     stats.cycles - st0.cycles, stats.rounds - st0.rounds, stats.bfs_launches - st0.bfs_launches,
                     stats.push_seconds - st0.push_seconds);
      // r from the flow, split by S_min, cut the arcs between the new regions
      hipLaunchKernelGGL((k_pf_after<T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, p, 1);
      hipLaunchKernelGGL((k_pf_stats_impl<1>), dim3(red_blocks()), dim3(NTHREADS), 0, s_, pf, (const int*)reach_, 1);
      hipLaunchKernelGGL((k_pf_split), dim3(blocks_for(nreg)), dim3(NTHREADS), 0, s_, pf, nreg, 1, 1);
      hipLaunchKernelGGL((k_pf_move), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, (const int*)reach_, 1);
      hipLaunchKernelGGL((k_pf_cutarcs<K, T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, p);
      LF_HIP(hipGetLastError());
    }
  }

  // Asynchronous divide and conquer: after every exact relabel, each region
  // with no active node left holds a maximum flow of its cut and is split and
  // given its next threshold at once, while the other regions keep pushing.
  // The cuts are the same as in dnc_loop; only their order over regions
  // differs, and regions never exchange flow, so the result is the same.
  void dnc_async(const lf_tuning& tun, long long& levels, long long& sections) {
    using clk = std::chrono::steady_clock;
    KP<T> p = params();
    PFP pf = pf_params();
    int nreg = (int)std::min<long long>(read_int(pf.nreg), pf_cap_);
    long long predicted = 0;
    long long active = thresholds(pf, nreg, 1, false, &predicted);
    if (active == 0) return;
    setup_excess(pf, p, 0, nreg);
    ++levels;
    sections += active;
    set_grid(variant_for(tun));
    p = params();
    const bool fixed_r = tun.rounds_per_relabel > 0;
    int R = fixed_r ? tun.rounds_per_relabel : 4;
    bool pushed = false;
    for (;;) {
      const auto tb = clk::now();
      bool exact = false;
      const long long act = global_relabel(p, pushed, 0, &exact);
      pushed = false;
      // regions without active nodes are solved
      hipLaunchKernelGGL((k_pf_busy<T>), dim3(red_blocks()), dim3(NTHREADS), 0, s_, pf, p);
      hipLaunchKernelGGL((k_fill<unsigned long long>), dim3(1), dim3(NTHREADS), 0, s_, pf_cnt_ + 3, 0ull, 1LL);
      hipLaunchKernelGGL((k_pf_mark_done), dim3(blocks_for(nreg)), dim3(NTHREADS), 0, s_, pf, nreg, pf_cnt_ + 3);
      unsigned long long ndone = 0;
      read_u64(pf_cnt_ + 3, &ndone, 1);
      const double t_relabel = std::chrono::duration<double>(clk::now() - tb).count();
      stats.bfs_seconds += t_relabel;
      ++stats.cycles;
      if (ndone > 0) {
        const auto tf = clk::now();
        // r from the flow, S_min of the solved regions, split, next thresholds
        hipLaunchKernelGGL((k_pf_after<T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, p, 4);
        hipLaunchKernelGGL((k_pf_flood_init<T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, p, reach_, 4);
        run_tiled([&](const BP& q) {
          launch_flood(q);
          ++stats.flood_launches;
        }, 0, false, nullptr, false, FLOOD_FIRST_BATCH);
        hipLaunchKernelGGL((k_pf_stats_impl<1>), dim3(red_blocks()), dim3(NTHREADS), 0, s_, pf, (const int*)reach_, 4);
        hipLaunchKernelGGL((k_pf_split), dim3(blocks_for(nreg)), dim3(NTHREADS), 0, s_, pf, nreg, 4, 5);
        hipLaunchKernelGGL((k_pf_move), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, (const int*)reach_, 5);
        hipLaunchKernelGGL((k_pf_cutarcs<K, T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, p);
        LF_HIP(hipGetLastError());
        if (read_int(pf.err)) return;  // region table full: the caller falls back
        nreg = (int)std::min<long long>(read_int(pf.nreg), pf_cap_);
        const long long fresh = thresholds(pf, nreg, 5, false, &predicted);
        if (fresh > 0) setup_excess(pf, p, 1, nreg);
        ++levels;
        sections += fresh;
        stats.flood_seconds += std::chrono::duration<double>(clk::now() - tf).count();
        if (trace_pf_)
          std::fprintf(stderr, "[lf-pf] cycle %lld solved %llu new cuts %lld predicted %lld regions %d active nodes %lld\n",
                       stats.cycles, ndone, fresh, predicted, nreg, act);
        if (act == 0 && fresh == 0) break;  // every region final
        continue;                           // new excesses: relabel before pushing
      }
      if (act == 0) break;
      p.iters = tun.local_iters > 0 ? tun.local_iters : (act * 64 < N_ ? 16 : 4);
      // adaptive schedule: chunks of 4, 8, then 16 rounds while tiles have
      // work (one host read per chunk; R is a multiple of 4)
      int chunk = fixed_r ? R : 4;
      int done_rounds = 0;
      double t_push = 0;
      unsigned long long tiles0 = tiles_seen_;  // no round since the relabel read it
      while (done_rounds < R) {
        LF_HIP(hipEventRecord(ev_[0], s_));
        const int n = std::min(chunk, R - done_rounds);
        push_rounds(p, n);
        LF_HIP(hipGetLastError());
        LF_HIP(hipEventRecord(ev_[1], s_));
        pushed = true;
        // one synchronisation: the tile counter (adaptive schedule) or the event
        unsigned long long tiles1 = 0;
        if (fixed_r)
          LF_HIP(hipEventSynchronize(ev_[1]));
        else
          tiles1 = read_tiles();
        float ms = 0;
        LF_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
        t_push += ms * 1e-3;
        done_rounds += n;
        if (!fixed_r) {
          if (tiles1 == tiles0) break;  // no tile had work
          tiles0 = tiles1;
          chunk = std::min(chunk * 2, 16);
        }
      }
      stats.push_seconds += t_push;
      if (!fixed_r) {
        const double per_round = t_push / std::max(done_rounds, 1);
        const double want = per_round > 0 ? push_relabel_ratio() * t_relabel / per_round : 4.0;
        R = 4 * (int)std::max(1.0, std::min(64.0, std::round(want / 4)));
      }
    }
  }

  void pf_values_device(int64_t* d_out, hipStream_t s) override {
    hipLaunchKernelGGL((k_pf_values), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf_params(),
                       reinterpret_cast<long long*>(d_out));
    LF_HIP(hipGetLastError());
    sync_to(s);
  }
  void pf_cut_device(int32_t* d_out, hipStream_t s) override {
    hipLaunchKernelGGL((k_pf_cut), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf_params(), d_out);
    LF_HIP(hipGetLastError());
    sync_to(s);
  }
  void pf_regions_device(int32_t* d_reg, int64_t* d_r, hipStream_t s) override {
    if (!pf_solved) fail(LF_ERR_STATE, "no divide and conquer has run");
    LF_HIP(hipMemcpyAsync(d_reg, pf_reg_, sizeof(int) * N_, hipMemcpyDeviceToDevice, s_));
    LF_HIP(hipMemcpyAsync(d_r, pf_r_, sizeof(long long) * N_, hipMemcpyDeviceToDevice, s_));
    sync_to(s);
  }

  // Every region of reg (ids 0..nreg-1, -1: not checked) is cut at ceil of the
  // mean of its values r, with the arcs to other regions cut: the number of
  // its nodes on the low side of the minimal cut goes to cnt_low (0 when the
  // values span at most one unit), each node's side to low.  A region is a
  // level set of its own problem exactly when that side is empty or whole.
  void pf_check_regions(const int64_t* r, const int32_t* reg, int nreg, int64_t* cnt_low, uint8_t* low) override {
    if (nreg < 1) return;
    PfMode pm{pf_mode_};
    const auto t0 = std::chrono::steady_clock::now();
    load_working();
    ensure_pf_buffers(std::max<long long>(nreg + 1, 1024));
    std::vector<int> regd((size_t)N_);
    for (long long v = 0; v < N_; ++v) regd[v] = reg[v] < 0 ? nreg : reg[v];
    LF_HIP(hipMemcpyAsync(pf_r_, r, sizeof(long long) * N_, hipMemcpyHostToDevice, s_));
    LF_HIP(hipMemcpyAsync(pf_reg_, regd.data(), sizeof(int) * N_, hipMemcpyHostToDevice, s_));
    PFP pf = pf_params();
    KP<T> p = params();
    clear_table();
    zero_int(pf.err, 1);
    hipLaunchKernelGGL((k_pf_init_check), dim3(blocks_for(nreg + 1)), dim3(NTHREADS), 0, s_, pf, nreg + 1, nreg);
    hipLaunchKernelGGL((k_pf_cutarcs<K, T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, pf, p);
    LF_HIP(hipGetLastError());
    long long predicted = 0;
    const long long active = thresholds(pf, nreg + 1, 1, false, &predicted);
    if (active > 0) {
      setup_excess(pf, p, 0, nreg + 1);
      solve_state(lf_tuning{});
      hipLaunchKernelGGL((k_pf_stats_impl<1>), dim3(red_blocks()), dim3(NTHREADS), 0, s_, pf, (const int*)reach_, 1);
      LF_HIP(hipGetLastError());
    }
    std::vector<RegionStat> tab((size_t)nreg);
    LF_HIP(hipMemcpyAsync(tab.data(), pf_tab_, sizeof(RegionStat) * nreg, hipMemcpyDeviceToHost, s_));
    std::vector<int> reach(active > 0 ? (size_t)N_ : 0);
    if (active > 0) LF_HIP(hipMemcpyAsync(reach.data(), reach_, sizeof(int) * N_, hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    for (int k = 0; k < nreg; ++k) cnt_low[k] = tab[k].state == 1 ? (int64_t)tab[k].cnt1 : 0;
    for (long long v = 0; v < N_; ++v)
      low[v] = (reg[v] >= 0 && active > 0 && tab[reg[v]].state == 1 && reach[v]) ? 1 : 0;
    stats.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    pf_solved = false;
  }
  void set_base_network(const int64_t* red, const int64_t* alpha_lattice_major) override {
    tv_mode_ = false;
    // host arrays: red[N], alpha[N*K] (node-major, lattice direction order)
    std::vector<long long> planes((size_t)N_ * K);
    for (long long v = 0; v < N_; ++v)
      for (int i = 0; i < K; ++i) planes[(size_t)i * N_ + v] = alpha_lattice_major[(size_t)v * K + i];
    LF_HIP(hipMemcpyAsync(red0_, red, sizeof(long long) * N_, hipMemcpyHostToDevice, s_));
    LF_HIP(hipMemcpyAsync(alpha0_, planes.data(), sizeof(long long) * N_ * K, hipMemcpyHostToDevice, s_));
    LF_HIP(hipStreamSynchronize(s_));
    solved = false;
    pf_solved = false;
  }

  // ---- batched lattices ------------------------------------------------------
  void batch_set_layout(const std::vector<BatchRect>& rects) override {
    if (g_.ndim != 2 || g_.D != 1) fail(LF_ERR_UNSUPPORTED, "batches are 2-D only");
    const int n = (int)rects.size();
    std::vector<long long> blk((size_t)n + 1, 0);
    for (int i = 0; i < n; ++i) {
      const BatchRect& R = rects[i];
      if (R.h <= 0 || R.w <= 0 || R.y0 < 0 || R.x0 < 0 || R.y0 + R.h > g_.H || R.x0 + R.w > g_.W)
        fail(LF_ERR_BAD_ARG, "batch lattice outside the canvas");
      blk[i + 1] = blk[i] + (R.h * R.w + BCHUNK - 1) / BCHUNK;
    }
    if (blk[n] > 0x7fffffffLL) fail(LF_ERR_BAD_ARG, "batch too large");
    hipFree(b_rects_);
    hipFree(b_blk_);
    hipFree(b_flows_);
    b_rects_ = nullptr; b_blk_ = nullptr; b_flows_ = nullptr;
    LF_HIP(hipMalloc(&b_rects_, sizeof(BatchRect) * std::max(n, 1)));
    LF_HIP(hipMalloc(&b_blk_, sizeof(long long) * (n + 1)));
    LF_HIP(hipMalloc(&b_flows_, sizeof(unsigned long long) * std::max(n, 1)));
    LF_HIP(hipMemcpyAsync(b_rects_, rects.data(), sizeof(BatchRect) * n, hipMemcpyHostToDevice, s_));
    LF_HIP(hipMemcpyAsync(b_blk_, blk.data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice, s_));
    LF_HIP(hipStreamSynchronize(s_));
    b_rects_host_ = rects;
    b_nblocks_ = blk[n];
  }

  void batch_add_energies_device(int i, const int64_t* du, const int64_t* db, hipStream_t s) override {
    if (i < 0 || i >= (int)b_rects_host_.size()) fail(LF_ERR_OUT_OF_RANGE, "batch lattice index out of range");
    wait_for(s);
    const BatchRect& R = b_rects_host_[i];
    int* err = reinterpret_cast<int*>(dev_small_);
    LF_HIP(hipMemsetAsync(err, 0, sizeof(int), s_));
    hipLaunchKernelGGL((k_batch_reparam<K>), dim3(blocks_for(R.h * R.w)), dim3(NTHREADS), 0, s_, R, p_.W, N_,
                       p_.dirs, reinterpret_cast<const long long*>(du), reinterpret_cast<const long long*>(db),
                       red0_, alpha0_, err);
    LF_HIP(hipGetLastError());
    LF_HIP(hipMemcpyAsync(host_small_, err, sizeof(int), hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    if (reinterpret_cast<int*>(host_small_)[0] & 1)
      fail(LF_ERR_NONSUBMODULAR, "pairwise term with E00 + E11 > E01 + E10 (not submodular)");
    solved = false;
  }

  void batch_flows(int64_t* out) override {
    const int n = (int)b_rects_host_.size();
    if (n == 0) return;
    hipLaunchKernelGGL((k_fill<unsigned long long>), dim3(blocks_for(n)), dim3(NTHREADS), 0, s_, b_flows_, 0ull,
                       (long long)n);
    hipLaunchKernelGGL((k_batch_flow<T>), dim3((unsigned)b_nblocks_), dim3(NTHREADS), 0, s_, params(), red0_, b_rects_,
                       b_blk_, n, b_flows_);
    LF_HIP(hipGetLastError());
    LF_HIP(hipMemcpyAsync(out, b_flows_, sizeof(int64_t) * n, hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
  }

  void batch_cut_device(int32_t* d_out, int which, hipStream_t s) override {
    const int n = (int)b_rects_host_.size();
    if (n == 0) return;
    hipLaunchKernelGGL((k_batch_cut<T>), dim3((unsigned)b_nblocks_), dim3(NTHREADS), 0, s_, params(), reach_, which,
                       b_rects_, b_blk_, n, d_out);
    LF_HIP(hipGetLastError());
    sync_to(s);
  }

  void cut_device(int32_t* d_out, int which, hipStream_t s) override {
    KP<T> p = params();
    hipLaunchKernelGGL((k_cut<T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, p, reach_, which, d_out);
    LF_HIP(hipGetLastError());
    sync_to(s);
  }

 private:
  KP<T> params() const {
    KP<T> p = p_;
    p.res = res_; p.ex = ex_; p.h = h_; p.mbox = mbox_; p.mbox_in = mbox_ + N_ * K; p.mail = mail_;
    p.mail_in = mail_ + ntiles_; p.dmail = dmail_; p.mask = mask_;
    p.counter = counter_ptr();
    p.tiles = tiles_ptr();
    p.flag_in = flags_[cur_];
    p.flag_out = flags_[cur_ ^ 1];
    return p;
  }
  unsigned long long* counter_ptr() const { return reinterpret_cast<unsigned long long*>(dev_small_ + 16); }
  int red_blocks() const { return (int)std::min<long long>(RED_BLOCKS, std::max<long long>(1, blocks_for(N_))); }
  unsigned long long* tiles_ptr() const { return reinterpret_cast<unsigned long long*>(dev_small_ + 24); }
  unsigned long long tiles_seen_ = 0;  // processed-tile counter as the last relabel read it
  unsigned long long* active_ptr() const { return reinterpret_cast<unsigned long long*>(dev_small_ + 48); }
  void zero_counter() {
    hipLaunchKernelGGL((k_fill<unsigned long long>), dim3(1), dim3(NTHREADS), 0, s_, counter_ptr(), 0ull, 1LL);
  }
  unsigned long long read_tiles() {
    LF_HIP(hipMemcpyAsync(host_small_, tiles_ptr(), sizeof(unsigned long long), hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    return *reinterpret_cast<unsigned long long*>(host_small_);
  }
  int read_int(const int* d) {
    LF_HIP(hipMemcpyAsync(host_small_, d, sizeof(int), hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    return *reinterpret_cast<int*>(host_small_);
  }
  // Work queued on the consumer stream s (NULL = the null stream) after this
  // call waits for everything queued so far on the engine's stream.  The
  // engine's stream is non-blocking, so the null stream does not wait by
  // itself.
  void sync_to(hipStream_t s) {
    if (s == s_) return;
    hipEvent_t ev;
    LF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    LF_HIP(hipEventRecord(ev, s_));
    LF_HIP(hipStreamWaitEvent(s, ev, 0));
    LF_HIP(hipEventDestroy(ev));
  }
  // The engine's stream waits for everything queued so far on the producer
  // stream s (NULL = the null stream) before it reads inputs written there.
  void wait_for(hipStream_t s) {
    if (s == s_) return;
    hipEvent_t ev;
    LF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    LF_HIP(hipEventRecord(ev, s));
    LF_HIP(hipStreamWaitEvent(s_, ev, 0));
    LF_HIP(hipEventDestroy(ev));
  }
  void ensure_pf_buffers(long long cap) {
    if (!pf_r_) {
      LF_HIP(hipMalloc(&pf_cnt_, sizeof(unsigned long long) * 4));
      LF_HIP(hipMalloc(&pf_r_, sizeof(long long) * N_));
      LF_HIP(hipMalloc(&pf_reg_, sizeof(int) * N_));
      LF_HIP(hipMalloc(&pf_nreg_, sizeof(int) * 4));
    }
    if (cap > pf_cap_) {
      hipFree(pf_tab_);
      pf_tab_ = nullptr;
      LF_HIP(hipMalloc(&pf_tab_, sizeof(RegionStat) * cap));
      pf_cap_ = cap;
    }
  }
  void ensure_hist(long long n) {
    if (n <= pf_hist_cap_) return;
    hipFree(pf_hist_);
    pf_hist_ = nullptr;
    pf_hist_cap_ = std::max(n, 2 * pf_hist_cap_);
    LF_HIP(hipMalloc(&pf_hist_, sizeof(unsigned) * pf_hist_cap_));
  }
  void read_u64(const unsigned long long* d, unsigned long long* h, int n) {
    LF_HIP(hipMemcpyAsync(host_small_, d, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    std::memcpy(h, host_small_, sizeof(unsigned long long) * n);
  }
  static double env_double(const char* name, double dflt) {
    const char* v = std::getenv(name);
    return v ? std::atof(v) : dflt;
  }
  // Adaptive schedules choose the push rounds between two relabels so that
  // pushing takes about 2.5 times as long as the relabel before it.  Measured
  // with the ratio at 1 / 2 / 2.5 / 3 / 5: the 512-lattice batch 69.1 / 66.9 /
  // 67.7 / 73.0 / 85.3 ms; the 32-lambda 1024^2 TV path (16x32 tiles) 4.15 s
  // time-balanced, 3.93 s at 2, 3.92 s at 3, 4.23 s at 6, and 3.34-3.51 s at
  // 2-3 on 16x16 tiles; the 256^3 unthresholded noise volume 0.57 s balanced,
  // 0.52-0.53 s at 2-2.5, 0.50 s at 3.  LF_PUSH_RELABEL_RATIO overrides the
  // ratio (for studies).
  // S_min floods settle within about four launches on 2048^2 segmentation and
  // 256^2 grids (launches over settled tiles exit at their flag test, a few
  // microseconds each): the first check comes after four, one read instead of two.
  static constexpr int FLOOD_FIRST_BATCH = 4;
  static double push_relabel_ratio() {
    static const double r = env_double("LF_PUSH_RELABEL_RATIO", 2.5);
    return r;
  }
  void ensure_guide_buffers(int layers) {
    if (!g_preg_) {
      LF_HIP(hipMalloc(&g_preg_, sizeof(int) * N_));
      LF_HIP(hipMalloc(&g_parent_, sizeof(int) * N_));
      LF_HIP(hipMalloc(&g_idmap_, sizeof(int) * N_));
      LF_HIP(hipMalloc(&g_merged_, sizeof(int) * N_));
      LF_HIP(hipMalloc(&g_val_, sizeof(long long) * N_));
    }
    if (layers > g_layers_) {
      hipFree(g_layer_bad_);
      g_layer_bad_ = nullptr;
      LF_HIP(hipMalloc(&g_layer_bad_, sizeof(int) * layers));
      g_layers_ = layers;
    }
  }
  GP guide_params(const float* pred, float delta) const {
    GP g;
    g.pred = pred; g.preg = g_preg_; g.parent = g_parent_; g.idmap = g_idmap_; g.merged = g_merged_;
    g.layer_bad = g_layer_bad_; g.delta = delta; g.val = g_val_;
    return g;
  }
  void clear_table() {
    static_assert(sizeof(RegionStat) % 8 == 0, "RegionStat is cleared as 64-bit words");
    const long long words = (long long)(sizeof(RegionStat) / 8) * pf_cap_;
    hipLaunchKernelGGL((k_fill<unsigned long long>), dim3(blocks_for(words)), dim3(NTHREADS), 0, s_,
                       reinterpret_cast<unsigned long long*>(pf_tab_), 0ull, words);
  }
  void zero_int(int* a, long long n) {
    hipLaunchKernelGGL((k_fill<int>), dim3(blocks_for(n)), dim3(NTHREADS), 0, s_, a, 0, n);
  }
  PFP pf_params() const {
    PFP f;
    f.r = pf_r_; f.reg = pf_reg_; f.tab = pf_tab_; f.nreg = pf_nreg_; f.err = pf_nreg_ + 2;
    f.N = N_; f.cap = pf_cap_; f.D = p_.D; f.H = p_.H; f.W = p_.W;
    f.collapse = tv_mode_ ? 1 : 0;
    f.pred = guide_.pred;
    f.pscale = guide_.pscale;
    f.poff = guide_.poff;
    f.hist = pf_hist_;
    f.nbins = guide_.nbins;
    return f;
  }
  unsigned long long read_counter() {
    LF_HIP(hipMemcpyAsync(host_small_, counter_ptr(), sizeof(unsigned long long), hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    return *reinterpret_cast<unsigned long long*>(host_small_);
  }
  void fill_many(std::initializer_list<FillSeg> segs) {
    FillSet f{};
    long long mx = 0;
    for (const FillSeg& g : segs) {
      if (!g.p || g.n <= 0) continue;
      if (f.ns == 6) fail(LF_ERR_STATE, "fill_many: more than six segments");
      f.s[f.ns++] = g;
      mx = std::max(mx, g.n);
    }
    if (f.ns) hipLaunchKernelGGL(k_fill_multi, dim3(blocks_for(mx)), dim3(NTHREADS), 0, s_, f);
  }
  static unsigned* u32(void* ptr) { return reinterpret_cast<unsigned*>(ptr); }
  void fill_flags(int* f, int v, long long n = -1) {
    if (n < 0) n = ntiles_;
    hipLaunchKernelGGL((k_fill<int>), dim3(blocks_for(n)), dim3(NTHREADS), 0, s_, f, v, n);
  }

  // Repeat a tiled exact pass until no tile border changes.  Launches go in
  // batches; only the last launch of a batch is checked, launches over
  // converged tiles exit at their flag test.  resume: carry on from the tile
  // flags a capped pass left (its labels and flags are those of an uncapped
  // pass after the same launches).  deferred (capped passes): the batch that
  // reaches the cap is not checked here; *deferred tells the caller to read
  // the border counter with its own next read.
  // timed: the device time of the launch batches (hipEvents around each
  // batch, read once the batch is synchronised) goes to stats.bfs_device_seconds
  // first_batch: launches before the first check (then twice as many, up to 8)
  template <class F>
  bool run_tiled(F&& launch, long long cap = 0, bool resume = false, bool* deferred = nullptr, bool timed = false,
                 int first_batch = 2) {
    if (!resume) fill_many({{u32(bflags_[bcur_]), 1u, bntiles_}, {u32(bflags_[bcur_ ^ 1]), 0u, bntiles_}});
    long long launched = 0;
    for (int batch = first_batch;; batch = std::min(batch * 2, 8)) {
      if (cap > 0 && launched >= cap) return false;  // stopped before the fixpoint
      if (timed) LF_HIP(hipEventRecord(evb_[0], s_));
      for (int b = 0; b < batch; ++b) {
        if (b == batch - 1) zero_counter();
        BP q = bp_;
        q.flag_in = bflags_[bcur_];
        q.flag_out = bflags_[bcur_ ^ 1];
        q.counter = counter_ptr();
        launch(q);
        bcur_ ^= 1;
        ++launched;
      }
      LF_HIP(hipGetLastError());
      if (timed) {
        LF_HIP(hipEventRecord(evb_[1], s_));
        bfs_time_pending_ = true;
      }
      if (deferred && cap > 0 && launched >= cap) {
        *deferred = true;
        return false;
      }
      const unsigned long long border = read_counter();
      take_bfs_time();
      if (border == 0) return true;
      // labels / reach bits only move one way, so the fixpoint is reached;
      // a pass that does not get there within this many launches is a bug
      if (launched > 16 * bntiles_ + 4096) fail(LF_ERR_STATE, "tiled exact pass did not converge");
    }
  }

  // One push / relabel round over the current work list; the next round's
  // list is the other one, and the three list counters rotate (in, out, to clear).
  void launch_push_round(const KP<T>& p) {
    KP<T> q = p;
    // the round parity is cur_: this round fills mailbox cur_ and drains the
    // other, which the round before filled
    q.mbox = mbox_ + (cur_ ? N_ * K : 0);
    q.mbox_in = mbox_ + (cur_ ? 0 : N_ * K);
    q.mail = mail_ + (cur_ ? ntiles_ : 0);
    q.mail_in = mail_ + (cur_ ? 0 : ntiles_);
    q.flag_in = flags_[cur_];
    q.flag_out = flags_[cur_ ^ 1];
    q.wl_in = wl_[cur_];
    q.wl_out = wl_[cur_ ^ 1];
    q.wl_cnt_in = wl_cnt_ + wl_k_ % 3;
    q.wl_cnt_out = wl_cnt_ + (wl_k_ + 1) % 3;
    q.wl_cnt_zero = wl_cnt_ + (wl_k_ + 2) % 3;
    with_geo([&](auto g) {
      using GG = decltype(g);
      hipLaunchKernelGGL((k_push<K, GG, T, DIR>), dim3(wl_grid_[variant_]), dim3(NTHREADS), 0, s_, q);
    });
    cur_ ^= 1;
    wl_k_ = (wl_k_ + 1) % 3;
    mbox_dirty_ = true;
    ++stats.rounds;
  }

  // n push rounds in a row.  Their launch parameters depend only on the
  // tiling, the local iterations, the list phase (flag array x list counter,
  // 2 x 3) and n, so each such chunk is captured once as a hipGraph and
  // replayed; the graph holds kernel nodes only (LF_NO_GRAPHS=1: plain launches).
  void push_rounds(const KP<T>& p, int n) {
    if (!use_graphs_ || n < 2) {
      for (int r = 0; r < n; ++r) launch_push_round(p);
      return;
    }
    const auto key = std::make_tuple(variant_, p.iters, cur_ + 2 * wl_k_, n);
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
      const int cur0 = cur_, k0 = wl_k_;
      const long long rounds0 = stats.rounds;
      LF_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
      for (int r = 0; r < n; ++r) launch_push_round(p);
      hipGraph_t g = nullptr;
      LF_HIP(hipStreamEndCapture(s_, &g));
      hipGraphExec_t ex = nullptr;
      const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      LF_HIP(e);
      it = graphs_.emplace(key, ex).first;
      // the captured launches have not run: back to the chunk's first phase
      cur_ = cur0;
      wl_k_ = k0;
      stats.rounds = rounds0;
    }
    LF_HIP(hipGraphLaunch(it->second, s_));
    cur_ ^= n & 1;
    wl_k_ = (wl_k_ + n) % 3;
    stats.rounds += n;
    mbox_dirty_ = true;
  }

  // add the device time of the last timed batch once the stream passed it
  void take_bfs_time() {
    if (!bfs_time_pending_) return;
    float ms = 0;
    LF_HIP(hipEventElapsedTime(&ms, evb_[0], evb_[1]));
    stats.bfs_device_seconds += ms * 1e-3;
    bfs_time_pending_ = false;
  }

  void launch_flood(const BP& q) {
    if (bits_)
      hipLaunchKernelGGL(k_flood_wave, dim3(bp_.nblocks), dim3(WT), 0, s_, q, wp_, planes_, reach_);
    else
      hipLaunchKernelGGL((k_flood_tile<K, B, MT, DIR>), dim3(bp_.nblocks), dim3(BTHREADS), 0, s_, q, mask_, reach_);
  }

  // Deliver mailboxes, BFS labels, flag the tiles of active nodes; returns the
  // number of active nodes.  With cap > 0 the BFS may stop after cap launches:
  // the labels are then lengths of real residual paths (upper bounds of the
  // distances) and unreached nodes are INF, which only steers pushes; *exact
  // tells whether the fixpoint was reached.  resume: no push round since a
  // capped relabel, whose pass is carried on to the fixpoint.
  long long global_relabel(KP<T> p, bool pushed, long long cap, bool* exact, bool resume = false) {
    if (resume) pushed = false;
    bool deferred = false;
    // the delivery pass also sets the start labels when the masks are current
    const bool fused_init = pushed && !bits_ && mask_valid_;
    if (pushed) {
      // only the last round's mailbox holds flow: the tiles it was sent to
      // drained the earlier ones in the round after (and a relabel found
      // every mailbox empty)
      KP<T> q = p;
      q.mbox = mbox_ + (cur_ ? 0 : N_ * K);
      with_geo([&](auto g) {
        using GG = decltype(g);
        hipLaunchKernelGGL((k_deliver<K, GG, T, DIR>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, q,
                           fused_init ? 1 : 0);
      });
    }
    if (pushed) mbox_dirty_ = false;  // every slot the rounds filled is delivered and cleared
    if (bits_) {
      const long long nw = (long long)p_.D * p_.H * wp_.ntx;  // one wave per plane word
      if (!resume)
        hipLaunchKernelGGL((k_bits_init<K, T>), dim3((unsigned)((nw * 64 + NTHREADS - 1) / NTHREADS)),
                           dim3(NTHREADS), 0, s_, p, wp_, planes_, mask_valid_ ? 1 : 0);
      *exact = run_tiled([&](const BP& q) {
        hipLaunchKernelGGL(k_bfs_wave, dim3(bp_.nblocks), dim3(WT), 0, s_, q, wp_, planes_, h_);
        ++stats.bfs_launches;
      }, cap, resume, &deferred);
    } else {
      if (resume) {
        // labels and tile flags as the capped pass left them
      } else if (mask_valid_) {
        if (!fused_init) hipLaunchKernelGGL((k_label_init<T>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, p);
      } else {
        hipLaunchKernelGGL((k_bfs_init<K, T, MT>), dim3(blocks_for(N_)), dim3(NTHREADS), 0, s_, p, mask_);
        mask_valid_ = true;
      }
      *exact = run_tiled([&](const BP& q) {
        BP qv = q;
        qv.visits = bvisits_;
        hipLaunchKernelGGL((k_bfs_tile<K, B, MT, DIR>), dim3(bp_.nblocks), dim3(BTHREADS), 0, s_, qv, mask_, h_);
        ++stats.bfs_launches;
      }, cap, resume, &deferred, true);
    }
    // push rounds start with exactly the tiles holding active nodes, and with
    // empty mailboxes (delivered above: no mail or delivery bits left)
    fill_many({{mail_, 0u, 2 * ntiles_},
               {pushed ? dmail_ : nullptr, 0u, ntiles_},
               {u32(flags_[cur_]), 0u, ntiles_},
               {u32(flags_[cur_ ^ 1]), 0u, ntiles_},
               {wl_cnt_, 0u, 3}});
    KP<T> q = p;
    q.flag_out = flags_[cur_];
    with_geo([&](auto g) {
      using GG = decltype(g);
      hipLaunchKernelGGL((k_count_active<GG, T>), dim3(red_blocks()), dim3(NTHREADS), 0, s_, q, part_);
    });
    // the first round's work list: the flagged tiles
    hipLaunchKernelGGL((k_flags_to_list), dim3(blocks_for(p_.ntiles)), dim3(NTHREADS), 0, s_, flags_[cur_],
                       p_.ntiles, wl_[cur_], wl_cnt_ + wl_k_);
    hipLaunchKernelGGL((k_reduce<false>), dim3(1), dim3(NTHREADS), 0, s_, part_, red_blocks(), 1, 0, active_ptr());
    LF_HIP(hipGetLastError());
    // one read of the small device words: the border counter of the last BFS
    // launch (slot 0) up to the active count (slot 4), and the checks of a
    // load that deferred them
    LF_HIP(hipMemcpyAsync(host_small_, dev_small_, 64, hipMemcpyDeviceToHost, s_));
    LF_HIP(hipStreamSynchronize(s_));
    if (load_check_pending_) {
      load_check_pending_ = false;
      check_load(host_small_);
    }
    const unsigned long long* hs =
        reinterpret_cast<const unsigned long long*>(static_cast<const char*>(host_small_) + 16);
    take_bfs_time();
    if (deferred) *exact = hs[0] == 0;
    tiles_seen_ = hs[1];
    return (long long)hs[4];
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
  // LF_WAVE_BFS=1: 2D lattices with 3x3 neighbourhoods run the exact passes
  // on wave tiles over arc bit planes instead of the LDS-list tile kernels
  const bool bits_ = B::TD == 1 && B::TH == WT && B::TW == WT && B::RZ == 0 && B::RY == 1 && B::RX == 1 &&
                     std::getenv("LF_WAVE_BFS") != nullptr;
  WP wp_{};
  unsigned long long* planes_ = nullptr;
  int* reach_ = nullptr;
  int* flags_[2] = {nullptr, nullptr};
  int* wl_[2] = {nullptr, nullptr};  // push work lists (tile ids)
  unsigned* wl_cnt_ = nullptr;       // their lengths, rotating over three counters
  unsigned* mail_ = nullptr;         // per tile: sides that sent it flow
  unsigned* dmail_ = nullptr;        // per tile: sides that sent it flow since the last delivery
  bool mbox_dirty_ = true;           // the mailbox may hold flow (or was never cleared)
  bool debug_checks_ = false;        // invariant checks after push chunks and relabels
  bool mask_valid_ = false;          // mask_ matches res_ (kept by load, push, delivery, cuts)
  const bool use_graphs_ = std::getenv("LF_NO_GRAPHS") == nullptr;
  std::map<std::tuple<int, int, int, int>, hipGraphExec_t> graphs_;  // push chunks by (tiling, iters, phase, n)
  int wl_k_ = 0;                     // counter index of the next round's input list
  int wl_grid_[3] = {1, 1, 1};       // resident blocks of k_push per tiling
  int cur_ = 0;
  BP bp_{};
  TileGrid grid_[3];
  int variant_ = 0;
  long long bntiles_ = 0;
  int* bflags_[2] = {nullptr, nullptr};
  int bcur_ = 0;
  char* dev_small_ = nullptr;
  long long* part_ = nullptr;
  void* host_small_ = nullptr;
  hipEvent_t ev_[2] = {nullptr, nullptr};
  hipEvent_t evb_[2] = {nullptr, nullptr};  // around a batch of exact-BFS launches
  bool bfs_time_pending_ = false;
  unsigned long long* bvisits_ = nullptr;   // exact-BFS node visits, one slot per XCD
  unsigned long long max_red_ = 0, max_alpha_ = 0;
  bool load_check_pending_ = false;  // a load's capacity checks wait for the next relabel's read
  const bool trace_ = std::getenv("LF_TRACE") != nullptr;
  const bool trace_pf_ = std::getenv("LF_TRACE_PF") != nullptr;
  bool tv_mode_ = false;  // TV network: regions that a mean cut leaves whole are final
  bool pf_mode_ = false;  // a parametric divide and conquer is running
  struct PfMode {        // sets pf_mode_ for the scope of a parametric run
    bool& m;
    bool old;
    explicit PfMode(bool& f) : m(f), old(f) { m = true; }
    ~PfMode() { m = old; }
  };
  // LF_PF_ASYNC selects the asynchronous divide and conquer for TV networks
  // (same result; on the 1024^2 x 32-lambda path it measured slower than the
  // depth-by-depth order, so that order is the default)
  const bool async_ = std::getenv("LF_PF_ASYNC") != nullptr;
  long long* pf_r_ = nullptr;
  int* pf_reg_ = nullptr;
  RegionStat* pf_tab_ = nullptr;
  int* pf_nreg_ = nullptr;
  long long pf_cap_ = 0;
  unsigned long long* pf_cnt_ = nullptr;
  unsigned* pf_hist_ = nullptr;
  long long pf_hist_cap_ = 0;
  struct GuideParams {
    const float* pred = nullptr;
    double pscale = 0, poff = 0;
    long long min_range = 0;
    double gap_frac = 0, min_side = 0;
    int nbins = 32;
    bool defer = false;
  } guide_;
  int* g_preg_ = nullptr;
  int* g_parent_ = nullptr;
  int* g_idmap_ = nullptr;
  int* g_merged_ = nullptr;
  long long* g_val_ = nullptr;
  int* g_layer_bad_ = nullptr;
  int g_layers_ = 0;
  std::vector<BatchRect> b_rects_host_;
  BatchRect* b_rects_ = nullptr;
  long long* b_blk_ = nullptr;
  unsigned long long* b_flows_ = nullptr;
  long long b_nblocks_ = 0;
};

template <int K, class G0, class G1, class G2, class B, class DIR = RtDir>
std::unique_ptr<MinCutEngine> make_k(const LatticeGeom& g, int cap_bytes) {
  using MT = MaskT<K>;
  if (cap_bytes == 4) return std::unique_ptr<MinCutEngine>(new Engine<K, G0, G1, G2, B, int, MT, DIR>(g));
  return std::unique_ptr<MinCutEngine>(new Engine<K, G0, G1, G2, B, long long, MT, DIR>(g));
}

// Push tilings (G0..G2) and exact-pass tile (B) of the radius-1 neighbourhoods.
struct Tiles2d1 {
  using G0 = Geo<1, 16, 32, 0, 1, 1>;
  // one node per thread: the shortest rounds, for small lattices whose rounds
  // are one tile's latency (256^2 grid: 10.5 us per round against 19.6 us on
  // 32x32 tiles)
  using G1 = Geo<1, 16, 16, 0, 1, 1>;
  using G2 = Geo<1, 32, 32, 0, 1, 1>;
  using B = BGeo<1, 64, 64, 0, 1, 1>;
};
struct Tiles3d6 {
  using G0 = Geo<4, 8, 32, 1, 1, 1>;
  using G1 = Geo<2, 8, 32, 1, 1, 1>;
  using G2 = Geo<4, 16, 16, 1, 1, 1>;
  using B = BGeo<16, 16, 16, 1, 1, 1>;
};
struct Tiles3d26 {  // also the 18-neighbourhood
  using G0 = Geo<4, 4, 16, 1, 1, 1>;
  using G1 = Geo<2, 8, 16, 1, 1, 1>;
  using G2 = Geo<4, 8, 8, 1, 1, 1>;
  using B = BGeo<16, 16, 16, 1, 1, 1>;
};
template <int K, class TL, class DIR = RtDir>
std::unique_ptr<MinCutEngine> make_tiled(const LatticeGeom& g, int cap_bytes) {
  return make_k<K, typename TL::G0, typename TL::G1, typename TL::G2, typename TL::B, DIR>(g, cap_bytes);
}

int max_abs(const std::vector<int>& v) {
  int m = 0;
  for (int x : v) m = std::max(m, std::abs(x));
  return m;
}

}  // namespace

#ifndef LF_MINCUT_WIDE_TU
std::unique_ptr<MinCutEngine> make_mincut_engine(const LatticeGeom& g, int cap_bytes) {
  if (cap_bytes != 4 && cap_bytes != 8) fail(LF_ERR_BAD_ARG, "cap_bytes must be 4 or 8");
  if (g.ndim == 4) return make_generic_engine(g, cap_bytes);
  const int rz = max_abs(g.dz), ry = max_abs(g.dy), rx = max_abs(g.dx);
  const bool is3d = g.ndim == 3;
  // the lattice orders of the common radius-1 neighbourhoods have engines
  // with the direction table compiled in (lf_mincut_canon.hip)
  if (auto e = make_canon_engine(g, cap_bytes)) return e;
  if (!is3d && rz == 0 && ry <= 1 && rx <= 1) {
    if (g.k == 2) return make_tiled<2, Tiles2d1>(g, cap_bytes);
    if (g.k == 4) return make_tiled<4, Tiles2d1>(g, cap_bytes);
    if (g.k == 8) return make_tiled<8, Tiles2d1>(g, cap_bytes);
  }
  if (!is3d && rz == 0 && ry <= 2 && rx <= 2) {
    using G0 = Geo<1, 16, 32, 0, 2, 2>;
    using G1 = Geo<1, 8, 64, 0, 2, 2>;
    using G2 = Geo<1, 32, 32, 0, 2, 2>;
    using B = BGeo<1, 64, 64, 0, 2, 2>;
    if (g.k == 12) return make_k<12, G0, G1, G2, B>(g, cap_bytes);
    if (g.k == 16) return make_k<16, G0, G1, G2, B>(g, cap_bytes);
  }
  if (is3d && rz <= 1 && ry <= 1 && rx <= 1) {
    if (g.k == 6) return make_tiled<6, Tiles3d6>(g, cap_bytes);
    if (g.k == 18) return make_tiled<18, Tiles3d26>(g, cap_bytes);
    if (g.k == 26) return make_tiled<26, Tiles3d26>(g, cap_bytes);
  }
  if (auto e = make_wide_engine(g, cap_bytes)) return e;
  fail(LF_ERR_UNSUPPORTED, "no device path for this kernel (more than 48 directions or radius above 4)");
}
#endif

}  // namespace lfamd
