// Invariant checks of a device min-cut's working state, on host copies
// (the analogue of the reference's debug verification,
// solver/network_flow/push_relabel.hpp:132-267: _debug_forceVerifyAll and
// its capacity / excess / height checks).  Arrays are in the engine's layout:
// plane-major residuals res[i*N + v] over the (D, H, W) lattice in C order.
#pragma once

#include <cstdint>
#include <string>

namespace lfamd {

struct VerifyView {
  int D = 1, H = 1, W = 1, K = 0;
  const int* dz = nullptr;
  const int* dy = nullptr;
  const int* dx = nullptr;
  const int64_t* red0 = nullptr;    // base network: reduction per node (excess = -red0)
  const int64_t* alpha0 = nullptr;  // base network: arc capacities, K planes
  int cap_bytes = 8;                // width of res / ex / mailboxes (4 or 8)
  const void* res = nullptr;        // K planes
  const void* ex = nullptr;
  const void* mbox = nullptr;       // flow in flight, `mbox_planes` planes of K (0 = none)
  int mbox_planes = 0;
  const void* mask = nullptr;       // per-node arc bits (nullptr: not checked)
  int mask_bytes = 1;
  const int32_t* h = nullptr;       // labels (nullptr: not checked)
  bool labels_exact = false;        // h must be the exact BFS distance to the sinks
};

// Empty when every invariant holds, else a description of the first violation:
//   capacities: every residual >= 0; arcs leaving the lattice hold nothing;
//   pairs: res(v->u) + res(u->v) + the flow in flight both ways = the pair's
//          capacity in the base network;
//   conservation: ex[v] + red0[v] = sum over i of (res[i][v] - alpha0[i][v]);
//   masks: bit i = (res[i][v] > 0);
//   labels (labels_exact): h = 0 on nodes with ex < 0, else 1 + the smallest
//          label over residual arcs (INF_H = 2^30 when none is finite).
std::string verify_state(const VerifyView& s);

}  // namespace lfamd
