// Host invariant checks of a device min-cut's working state (lf_verify.hpp).
#include "lf_verify.hpp"

#include <cstdio>

namespace lfamd {
namespace {

constexpr int64_t INF_LABEL = 1 << 30;

int64_t load(const void* a, int bytes, int64_t i) {
  if (bytes == 4) return static_cast<const int32_t*>(a)[i];
  return static_cast<const int64_t*>(a)[i];
}

uint64_t load_mask(const void* a, int bytes, int64_t i) {
  switch (bytes) {
    case 1: return static_cast<const uint8_t*>(a)[i];
    case 2: return static_cast<const uint16_t*>(a)[i];
    case 4: return static_cast<const uint32_t*>(a)[i];
    default: return static_cast<const uint64_t*>(a)[i];
  }
}

std::string at(const char* what, int64_t v, int i, int64_t a, int64_t b) {
  char buf[256];
  std::snprintf(buf, sizeof buf, "%s at node %lld direction %d: %lld vs %lld", what, (long long)v, i, (long long)a,
                (long long)b);
  return buf;
}

}  // namespace

std::string verify_state(const VerifyView& s) {
  const int64_t N = (int64_t)s.D * s.H * s.W;
  const int K = s.K;
  auto nbr = [&](int64_t v, int i) -> int64_t {
    const int64_t x = v % s.W, y = (v / s.W) % s.H, z = v / ((int64_t)s.W * s.H);
    const int64_t nz = z + s.dz[i], ny = y + s.dy[i], nx = x + s.dx[i];
    if (nz < 0 || nz >= s.D || ny < 0 || ny >= s.H || nx < 0 || nx >= s.W) return -1;
    return (nz * s.H + ny) * s.W + nx;
  };
  auto inflight = [&](int i, int64_t v) -> int64_t {
    int64_t f = 0;
    for (int p = 0; p < s.mbox_planes; ++p) f += load(s.mbox, s.cap_bytes, ((int64_t)p * K + i) * N + v);
    return f;
  };
  for (int64_t v = 0; v < N; ++v) {
    int64_t dres = 0;
    for (int i = 0; i < K; ++i) {
      const int64_t r = load(s.res, s.cap_bytes, (int64_t)i * N + v);
      const int64_t a0 = s.alpha0[(int64_t)i * N + v];
      dres += r - a0;
      if (r < 0) return at("negative residual", v, i, r, 0);
      const int64_t u = nbr(v, i);
      if (u < 0) {
        if (r != 0 || inflight(i, v) != 0) return at("flow on an arc leaving the lattice", v, i, r, inflight(i, v));
        continue;
      }
      if (i < K / 2) {
        const int ri = (i + K / 2) % K;
        const int64_t have = r + load(s.res, s.cap_bytes, (int64_t)ri * N + u) + inflight(i, v) + inflight(ri, u);
        const int64_t cap = a0 + s.alpha0[(int64_t)ri * N + u];
        if (have != cap) return at("arc pair capacity", v, i, have, cap);
      }
      if (s.mask) {
        const bool bit = (load_mask(s.mask, s.mask_bytes, v) >> i) & 1u;
        if (bit != (r > 0)) return at("arc mask", v, i, bit ? 1 : 0, r);
      }
    }
    const int64_t e = load(s.ex, s.cap_bytes, v);
    if (e + s.red0[v] != dres) return at("flow conservation", v, -1, e + s.red0[v], dres);
  }
  if (s.h && s.labels_exact) {
    for (int64_t v = 0; v < N; ++v) {
      int64_t want;
      if (load(s.ex, s.cap_bytes, v) < 0) {
        want = 0;
      } else {
        int64_t m = INF_LABEL;
        for (int i = 0; i < K; ++i) {
          const int64_t u = nbr(v, i);
          if (u < 0 || load(s.res, s.cap_bytes, (int64_t)i * N + v) <= 0) continue;
          if (s.h[u] < m) m = s.h[u];
        }
        want = m >= INF_LABEL - 1 ? INF_LABEL : m + 1;
      }
      if (s.h[v] != want) return at("exact distance label", v, -1, s.h[v], want);
    }
  }
  return std::string();
}

}  // namespace lfamd
