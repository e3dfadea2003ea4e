// Boykov-Kolmogorov max-flow on a lattice network: the CPU cross-check solver
// "BKGCEnergyMinimizer".
//
// The reference wraps Kolmogorov's research code for this solver
// (solver/energy_minimization/bk_graphcuts.hpp:65-163, vendored under a
// research licence); this file is written from the published algorithm
// instead: Y. Boykov and V. Kolmogorov, "An Experimental Comparison of
// Min-Cut/Max-Flow Algorithms for Energy Minimization in Vision", IEEE TPAMI
// 26(9):1124-1137, 2004, Section 3 (search trees S and T grown from the
// terminals, augmentation along the path where they touch, adoption of the
// orphans it creates; the timestamp / distance marks of the adoption stage
// pick the candidate parent closest to its terminal and let later origin
// checks of the same stage stop at a verified node).  The growth-stage
// re-parenting heuristic is left out: it only shortens paths.
//
// Network: the same doubled ("raw") units as the device solver
// (NodeFiller::addE1 / addE2, network_flow_node.hpp:419-485): a node's
// terminal residual tr = -reduction (> 0: capacity from the source, < 0: to the
// sink) and res[v][i] the residual of the arc v -> v + dir[i].  Label 1 (the
// reference's on() = true) is the source side, so the minimal label-1 set is
// residual reachability from the source after the flow is maximum and the
// maximal one is the set of nodes that cannot reach the sink; both are taken
// by two searches over the final residual graph.

#include "lf_bk.hpp"

#include <algorithm>
#include <deque>
#include <limits>
#include <string>

#include "../../include/latticeflow_amd.h"
#include "lf_engine.hpp"

namespace lfamd {

namespace {
constexpr int64_t kNone = -1;      // free node
constexpr int64_t kTerminal = -2;  // tree root: its parent is the terminal
constexpr int64_t kOrphan = -3;    // lost its parent arc in the last augmentation
constexpr uint8_t kFree = 0, kS = 1, kT = 2;
}  // namespace

BkSolver::BkSolver(int ndim, const std::vector<int64_t>& shape, int k, const std::vector<int>& offs)
    : nd_(ndim), k_(k), shape_(shape), offs_(offs) {
  n_ = 1;
  stride_.assign(nd_, 1);
  for (int d = nd_ - 1; d >= 0; --d) {
    stride_[d] = n_;
    n_ *= shape_[d];
  }
  nbr_.assign((size_t)n_ * k_, -1);
  std::vector<int64_t> c(nd_);
  for (int64_t v = 0; v < n_; ++v) {
    int64_t rem = v;
    for (int d = 0; d < nd_; ++d) {
      c[d] = rem / stride_[d];
      rem -= c[d] * stride_[d];
    }
    for (int i = 0; i < k_; ++i) {
      int64_t w = 0;
      bool in = true;
      for (int d = 0; d < nd_ && in; ++d) {
        const int64_t cn = c[d] + offs_[(size_t)i * nd_ + d];
        in = cn >= 0 && cn < shape_[d];
        w += cn * stride_[d];
      }
      if (in) nbr_[(size_t)v * k_ + i] = w;
    }
  }
  unary_.assign((size_t)n_ * 2, 0);
  binary_.assign((size_t)n_ * (k_ / 2) * 4, 0);
}

void BkSolver::set_energies(const int64_t* unary, const int64_t* binary) {
  std::copy(unary, unary + unary_.size(), unary_.begin());
  std::copy(binary, binary + binary_.size(), binary_.begin());
  solved_ = false;
}

void BkSolver::run() {
  const int k = k_, kp = k_ / 2;
  auto rev = [k](int i) { return (i + k / 2) % k; };
  // energies -> network
  tr_.assign(n_, 0);
  res_.assign((size_t)n_ * k, 0);
  for (int64_t v = 0; v < n_; ++v) tr_[v] -= 2 * (unary_[2 * v + 1] - unary_[2 * v]);
  for (int64_t v = 0; v < n_; ++v)
    for (int e = 0; e < kp; ++e) {
      const int64_t w = nbr_[(size_t)v * k + e];
      if (w < 0) continue;
      const int64_t* E = &binary_[((size_t)v * kp + e) * 4];
      const int64_t q1 = E[2] - E[0], q2 = E[1] - E[0], q12 = E[0] + E[3] - E[1] - E[2];
      if (q12 > 0) fail(LF_ERR_NONSUBMODULAR, "pairwise term with E00 + E11 > E01 + E10 (not submodular)");
      tr_[v] -= 2 * q1 + q12;
      tr_[w] -= 2 * q2 + q12;
      res_[(size_t)v * k + e] += -q12;
      res_[(size_t)w * k + rev(e)] += -q12;
    }

  std::vector<uint8_t> tree(n_, kFree), inq(n_, 0);
  std::vector<int64_t> parent(n_, kNone);
  std::vector<int8_t> pdir(n_, -1);  // direction from a node to its parent
  std::vector<int64_t> ts(n_, 0), dist(n_, 0);
  std::deque<int64_t> active, orphans;
  int64_t time = 0;
  for (int64_t v = 0; v < n_; ++v)
    if (tr_[v] != 0) {
      tree[v] = tr_[v] > 0 ? kS : kT;
      parent[v] = kTerminal;
      dist[v] = 1;
      active.push_back(v);
      inq[v] = 1;
    }
  // residual capacity of the tree arc between p (in tree t) and its
  // neighbour along direction i: p -> q in S, q -> p in T
  auto tree_cap = [&](uint8_t t, int64_t p, int64_t q, int i) {
    return t == kS ? res_[(size_t)p * k + i] : res_[(size_t)q * k + rev(i)];
  };
  flow_ = 0;
  augmentations_ = 0;
  for (;;) {
    // ---- growth: the first active node that touches the other tree ----
    int64_t s_node = -1, t_node = -1;
    int s_dir = -1;  // direction of the arc s_node -> t_node
    while (!active.empty() && s_node < 0) {
      const int64_t p = active.front();
      if (tree[p] == kFree) {
        active.pop_front();
        inq[p] = 0;
        continue;
      }
      const uint8_t t = tree[p];
      for (int i = 0; i < k; ++i) {
        const int64_t q = nbr_[(size_t)p * k + i];
        if (q < 0 || tree_cap(t, p, q, i) <= 0) continue;
        if (tree[q] == kFree) {
          tree[q] = t;
          parent[q] = p;
          pdir[q] = (int8_t)rev(i);
          ts[q] = ts[p];
          dist[q] = dist[p] + 1;
          if (!inq[q]) {
            active.push_back(q);
            inq[q] = 1;
          }
        } else if (tree[q] != t) {
          if (t == kS) {
            s_node = p, t_node = q, s_dir = i;
          } else {
            s_node = q, t_node = p, s_dir = rev(i);
          }
          break;
        }
      }
      if (s_node < 0) {  // no path through p: it stops growing
        active.pop_front();
        inq[p] = 0;
      }
    }
    if (s_node < 0) break;  // the trees are separated by saturated arcs

    // ---- augmentation: bottleneck of s -> ... -> s_node -> t_node -> ... -> t ----
    int64_t d = res_[(size_t)s_node * k + s_dir];
    for (int64_t x = s_node; parent[x] != kTerminal; x = parent[x])
      d = std::min(d, res_[(size_t)parent[x] * k + rev(pdir[x])]);
    {
      int64_t x = s_node;
      while (parent[x] != kTerminal) x = parent[x];
      d = std::min(d, tr_[x]);
    }
    for (int64_t x = t_node; parent[x] != kTerminal; x = parent[x])
      d = std::min(d, res_[(size_t)x * k + pdir[x]]);
    {
      int64_t x = t_node;
      while (parent[x] != kTerminal) x = parent[x];
      d = std::min(d, -tr_[x]);
    }
    res_[(size_t)s_node * k + s_dir] -= d;
    res_[(size_t)t_node * k + rev(s_dir)] += d;
    for (int64_t x = s_node;;) {
      const int64_t par = parent[x];
      if (par == kTerminal) {
        tr_[x] -= d;
        if (tr_[x] == 0) {
          parent[x] = kOrphan;
          orphans.push_back(x);
        }
        break;
      }
      const int a = rev(pdir[x]);  // arc par -> x
      res_[(size_t)par * k + a] -= d;
      res_[(size_t)x * k + pdir[x]] += d;
      if (res_[(size_t)par * k + a] == 0) {
        parent[x] = kOrphan;
        orphans.push_back(x);
      }
      x = par;
    }
    for (int64_t x = t_node;;) {
      const int64_t par = parent[x];
      if (par == kTerminal) {
        tr_[x] += d;
        if (tr_[x] == 0) {
          parent[x] = kOrphan;
          orphans.push_back(x);
        }
        break;
      }
      const int a = pdir[x];  // arc x -> par
      res_[(size_t)x * k + a] -= d;
      res_[(size_t)par * k + rev(a)] += d;
      if (res_[(size_t)x * k + a] == 0) {
        parent[x] = kOrphan;
        orphans.push_back(x);
      }
      x = par;
    }
    flow_ += d;
    ++augmentations_;

    // ---- adoption ----
    ++time;
    while (!orphans.empty()) {
      const int64_t p = orphans.front();
      orphans.pop_front();
      const uint8_t t = tree[p];
      int64_t best = -1, best_d = std::numeric_limits<int64_t>::max();
      int best_dir = -1;
      for (int i = 0; i < k; ++i) {
        const int64_t q = nbr_[(size_t)p * k + i];
        if (q < 0 || tree[q] != t) continue;
        // the arc from the candidate parent q towards p (S) or from p to q (T)
        const int64_t cap = t == kS ? res_[(size_t)q * k + rev(i)] : res_[(size_t)p * k + i];
        if (cap <= 0) continue;
        // q qualifies when its parent chain reaches the terminal
        int64_t steps = 0, x = q, dq = -1;
        for (;;) {
          if (ts[x] == time) {
            dq = steps + dist[x];
            break;
          }
          if (parent[x] == kTerminal) {
            ts[x] = time;
            dist[x] = 1;
            dq = steps + 1;
            break;
          }
          if (parent[x] < 0) break;  // orphan or free: no origin
          x = parent[x];
          ++steps;
        }
        if (dq < 0) continue;
        if (dq < best_d) {
          best = q;
          best_d = dq;
          best_dir = i;
        }
        // cache the distances of the checked chain for this adoption stage
        for (int64_t y = q, dy = dq; ts[y] != time; y = parent[y], --dy) {
          ts[y] = time;
          dist[y] = dy;
        }
      }
      if (best >= 0) {
        parent[p] = best;
        pdir[p] = (int8_t)best_dir;
        ts[p] = time;
        dist[p] = best_d + 1;
        continue;
      }
      // no parent: p becomes free; neighbours of its tree may grow into it
      // again, and its children are orphans
      for (int i = 0; i < k; ++i) {
        const int64_t q = nbr_[(size_t)p * k + i];
        if (q < 0 || tree[q] != t) continue;
        const int64_t cap = t == kS ? res_[(size_t)q * k + rev(i)] : res_[(size_t)p * k + i];
        if (cap > 0 && !inq[q]) {
          active.push_back(q);
          inq[q] = 1;
        }
        if (parent[q] == p) {
          parent[q] = kOrphan;
          orphans.push_back(q);
        }
      }
      tree[p] = kFree;
      parent[p] = kNone;
    }
  }

  // canonical cuts from the final residual graph
  smin_.assign(n_, 0);
  smax_.assign(n_, 1);
  std::vector<int64_t> stack;
  for (int64_t v = 0; v < n_; ++v)
    if (tr_[v] > 0) {
      smin_[v] = 1;
      stack.push_back(v);
    }
  while (!stack.empty()) {
    const int64_t v = stack.back();
    stack.pop_back();
    for (int i = 0; i < k; ++i) {
      const int64_t w = nbr_[(size_t)v * k + i];
      if (w >= 0 && !smin_[w] && res_[(size_t)v * k + i] > 0) {
        smin_[w] = 1;
        stack.push_back(w);
      }
    }
  }
  for (int64_t v = 0; v < n_; ++v)
    if (tr_[v] < 0) {
      smax_[v] = 0;
      stack.push_back(v);
    }
  while (!stack.empty()) {
    const int64_t w = stack.back();
    stack.pop_back();
    for (int i = 0; i < k; ++i) {
      const int64_t v = nbr_[(size_t)w * k + i];
      if (v >= 0 && smax_[v] && res_[(size_t)v * k + rev(i)] > 0) {
        smax_[v] = 0;
        stack.push_back(v);
      }
    }
  }
  solved_ = true;
}

void BkSolver::cut(int32_t* out, int which) const {
  if (!solved_) fail(LF_ERR_STATE, "run() has not been called");
  const std::vector<uint8_t>& s = which == LF_CUT_SMAX ? smax_ : smin_;
  for (int64_t v = 0; v < n_; ++v) out[v] = s[v];
}

}  // namespace lfamd
