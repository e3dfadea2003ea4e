// The event-driven TV regularization path (TVSolver::run,
// solver/tv/tv_solver.hpp:114-149; _constructInitialRegPathsFromSolvedLattice
// :300-564; getRegularizationPath :181-244 and traceRegularizationPath
// :566-642; TVRegPathSegment, tv_regpath_node.hpp:67-1028).
//
// The TV network is built once at lambda_max; the point s = lambda /
// lambda_max of the path scales the function values by s (PFFlowNode::fv,
// pf_flow_node.hpp:227-230).  A level set A whose neighbours keep their order
// relative to A has the value
//
//   r_A(s) = (G_A + s Q_A) / n_A,  Q_A = sum of the raw function values on A,
//                                  G_A = net flow of A's saturated boundary arcs,
//
// linear in s (get_r_AtLambda / syncInformation, tv_regpath_node.hpp:146-150,
// 208-266).  From s = 1 down to 0 the partition changes by events:
//   * join: two neighbouring level sets whose lines meet (closed form,
//     lambdaOfJoin / setupFromJoin, :837-984); G and Q add;
//   * split: a level set stops being one (calculateSplit / applySplit,
//     :416-830).
//
// Splits are verified on the device, as a batch.  A level set A is one at s
// exactly when no subset S of A has a violation
//   g_S(s) = sum_S (r_A(s) - r_v(s)) - c(S, A \ S) > 0,
// which one minimum cut of A alone at its mean value decides
// (MinCutEngine::pf_check_regions).  Each g_S is linear in s and A is a level
// set at its birth, so A is one over its whole life [death, birth] exactly
// when it is one at its death.  Joins are therefore followed optimistically
// and every level set is checked at its death; checks are queued and packed
// into stacked copies of the lattice (overlapping level sets go to different
// copies), and when the copies are full, the whole live partition is checked
// at the current point as well, so that every level set alive above that
// point is verified.  A failed check locates the split point by Dinkelbach
// iterations (the root of the last violating subset's g_S is the next probe,
// until a probe passes).  The highest split point found is then recorded and
// the event pass runs again from s = 1 with that split applied; level sets
// born above it keep their identity and their verified status.
//
// Level sets are kept connected: the components of a level set, of a split's
// low part and of its high part each carry the value of the whole at that
// point (the violations of the parts add up to zero), so they are separate
// level sets from there on.
//
// The path is evaluated by tracing every pixel from its level set at s = 1
// through joins and splits to the one alive at s.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <map>
#include <numeric>
#include <queue>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/latticeflow_amd.h"
#include "lf_engine.hpp"
#include "lf_kernels.hpp"
#include "lf_tv.hpp"
#include "lf_tvpath.hpp"

namespace lfamd {

namespace {
typedef __int128 i128;
typedef long double ld;

struct Seg {
  int64_t n = 0;
  i128 Q = 0, G = 0;
  double s_hi = 1;   // born at
  double s_lo = -1;  // ended at (-1: lasts to s = 0)
  int end = 0;       // 1: joined into `up`; 2: split (parts in TvPath::Impl::children)
  int up = -1;
  int slot = -1;     // slot while alive
  ld value(ld s) const { return ((ld)G + s * (ld)Q) / (ld)n; }
};

struct Event {
  double s;
  int type;  // 0 join, 1 split
  int a, b;
};

// max-heap order: higher s first; at one s joins before splits; then by ids
struct EventLess {
  bool operator()(const Event& x, const Event& y) const {
    if (x.s != y.s) return x.s < y.s;
    if (x.type != y.type) return x.type > y.type;
    if (x.a != y.a) return x.a > y.a;
    return x.b > y.b;
  }
};

// neighbour slot of a level set and whether that neighbour lies above it
struct Adj {
  int32_t slot;
  uint8_t higher;
};

struct Forced {
  double s;
  std::vector<int32_t> low;  // sorted nodes of the low part
};

struct Cert {
  double birth, down_to;  // a verified level set on [down_to, birth]
};

// a level set under check
struct Region {
  int seg = 0;
  int layer = 0;
  double probe = 0;
  std::vector<int32_t> nodes;
  std::vector<int64_t> beta;      // boundary flow per node (constant over the life)
  std::vector<int32_t> last_low;  // violating subset of the last failed probe
  bool failed = false;
  bool split = false;
  double split_s = 0;
};

struct PassAbort {};  // a split was found: the event pass starts again
}  // namespace

struct TvPath::Impl {
  int64_t nx = 0, ny = 0, n = 0;
  int k = 0, L = 4;
  double lmax = 0;
  bool constant = false;  // a constant image is its own path
  TvBase b;
  std::vector<int64_t> fv;    // raw function value per node at s = 1 (2 uf)
  std::vector<int32_t> nbr;   // n * k neighbours, -1 outside the image
  std::vector<int64_t> acap;  // raw arc capacity per direction (2 pwf at lambda_max)
  std::unique_ptr<MinCutEngine> chk;

  // level sets at s = 1
  std::vector<std::vector<int32_t>> init_nodes;
  std::vector<i128> init_G, init_Q;

  // history of the last pass
  std::vector<Seg> segs;
  std::vector<int32_t> initial;                           // level set of every node at s = 1
  std::unordered_map<int, std::vector<int>> children;     // split level set -> parts
  std::unordered_map<int, std::vector<int32_t>> members;  // part of a split -> sorted nodes
  std::map<int, Forced> forced;
  std::unordered_map<int, Cert> cert;
  TvPathStats st;

  // event pass state
  std::vector<int32_t> parent, slot_seg;
  std::vector<uint8_t> dead;
  std::vector<std::vector<int32_t>> nodes;
  std::vector<std::vector<Adj>> adj;
  std::vector<int32_t> node_slot;
  std::priority_queue<Event, std::vector<Event>, EventLess> heap;
  std::vector<int64_t> smark, smark2;
  std::vector<uint8_t> sflag;
  int64_t sepoch = 0;
  std::vector<int64_t> nmark;
  int64_t nepoch = 0;

  // batched checks
  std::vector<Region> pending;
  std::vector<uint8_t> occ;  // L * n: node taken in a copy
  std::vector<int64_t> stage_r;
  std::vector<int32_t> stage_reg;
  std::vector<uint8_t> stage_low;

  // ---- slots (live level sets; union-find joins them) ------------------------
  int find(int x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  }
  int new_slot() {
    const int id = (int)parent.size();
    parent.push_back(id);
    slot_seg.push_back(-1);
    dead.push_back(0);
    nodes.emplace_back();
    adj.emplace_back();
    smark.push_back(0);
    smark2.push_back(0);
    sflag.push_back(0);
    return id;
  }
  int new_seg(i128 G, i128 Q, int64_t cnt, double s_hi, int slot) {
    Seg sg;
    sg.n = cnt;
    sg.G = G;
    sg.Q = Q;
    sg.s_hi = s_hi;
    sg.slot = slot;
    segs.push_back(sg);
    const int id = (int)segs.size() - 1;
    slot_seg[slot] = id;
    auto it = forced.find(id);
    if (it != forced.end()) heap.push(Event{it->second.s, 1, id, -1});
    return id;
  }

  // distinct live neighbours of root slot x (compacts its list); afterwards
  // smark[r] == sepoch and sflag[r] hold for each of them
  const std::vector<Adj>& live_adj(int x) {
    ++sepoch;
    std::vector<Adj>& a = adj[x];
    size_t w = 0;
    smark[x] = sepoch;
    for (size_t i = 0; i < a.size(); ++i) {
      const int r = find(a[i].slot);
      if (dead[r] || smark[r] == sepoch) continue;
      smark[r] = sepoch;
      sflag[r] = a[i].higher;
      a[w++] = Adj{r, a[i].higher};
    }
    a.resize(w);
    smark[x] = 0;
    return a;
  }

  // boundary flow of node v of root slot x (after live_adj(x))
  int64_t beta_of(int x, int32_t v) {
    int64_t beta = 0;
    for (int i = 0; i < k; ++i) {
      const int32_t u = nbr[(size_t)v * k + i];
      if (u < 0) continue;
      const int r = find(node_slot[u]);
      if (r == x) continue;
      if (smark[r] != sepoch) fail(LF_ERR_STATE, "TV path: boundary neighbour missing from the adjacency");
      beta += sflag[r] ? acap[i] : -acap[i];
    }
    return beta;
  }

  // ---- joins --------------------------------------------------------------------
  // Where a and o meet below s_cur, given whether o lies above a now; false
  // when they part.
  bool join_point(int a, int o, bool o_higher, double s_cur, double* s) const {
    const Seg &A = segs[a], &B = segs[o];
    // n_A n_B (r_A(s) - r_B(s)) = c0 + c1 s
    const i128 c0 = A.G * (i128)B.n - B.G * (i128)A.n;
    const i128 c1 = A.Q * (i128)B.n - B.Q * (i128)A.n;
    if (c1 == 0) {
      if (c0 != 0) return false;
      *s = s_cur;  // the same line
      return true;
    }
    // above the crossing a lies above o exactly when c1 > 0; they meet going
    // down only if that is their present order
    if ((c1 > 0) != !o_higher) return false;
    const ld sc = -(ld)c0 / (ld)c1;
    if (!(sc > 0)) return false;
    *s = (double)std::min<ld>(sc, (ld)s_cur);
    return true;
  }

  void push_joins(int id, double s_cur) {
    const int x = segs[id].slot;
    for (const Adj& e : live_adj(x)) {
      const int o = slot_seg[e.slot];
      double s;
      if (join_point(id, o, e.higher != 0, s_cur, &s)) heap.push(Event{s, 0, std::min(id, o), std::max(id, o)});
    }
  }

  // the live level set that id is part of at a join at s (a level set that an
  // earlier join at the same point absorbed takes part through its successor)
  int live_at(int id, double s) const {
    while (segs[id].s_lo >= 0) {
      const Seg& g = segs[id];
      if (g.end != 1) return -1;
      if (!(segs[g.up].s_hi <= s + 1e-12 * s)) return -1;
      id = g.up;
    }
    return id;
  }

  int apply_join(int a, int bs, double s) {
    const int sa = segs[a].slot, sb = segs[bs].slot;
    int big = sa, small = sb;
    if (nodes[sb].size() > nodes[sa].size()) std::swap(big, small);
    parent[small] = big;
    nodes[big].insert(nodes[big].end(), nodes[small].begin(), nodes[small].end());
    std::vector<int32_t>().swap(nodes[small]);
    adj[big].insert(adj[big].end(), adj[small].begin(), adj[small].end());
    std::vector<Adj>().swap(adj[small]);
    const i128 G = segs[a].G + segs[bs].G, Q = segs[a].Q + segs[bs].Q;
    const int64_t cnt = segs[a].n + segs[bs].n;
    const int c = new_seg(G, Q, cnt, s, big);
    for (int x : {a, bs}) {
      segs[x].s_lo = s;
      segs[x].end = 1;
      segs[x].up = c;
      segs[x].slot = -1;
    }
    push_joins(c, s);
    return c;
  }

  // ---- splits -------------------------------------------------------------------
  // connected components of the nodes of `set` marked nmark[v] == ep
  void components(const std::vector<int32_t>& set, int64_t ep, std::vector<std::vector<int32_t>>& out) {
    const int64_t seen = ++nepoch;
    for (int32_t s0 : set) {
      if (nmark[s0] != ep) continue;
      out.emplace_back();
      std::vector<int32_t>& comp = out.back();
      nmark[s0] = seen;
      comp.push_back(s0);
      for (size_t h = 0; h < comp.size(); ++h) {
        const int32_t v = comp[h];
        for (int i = 0; i < k; ++i) {
          const int32_t u = nbr[(size_t)v * k + i];
          if (u >= 0 && nmark[u] == ep) {
            nmark[u] = seen;
            comp.push_back(u);
          }
        }
      }
    }
  }

  void apply_split(int id, double s, const std::vector<int32_t>& low) {
    const int x = segs[id].slot;
    live_adj(x);
    const std::vector<int64_t> flag_mark(smark);  // the old level set's neighbours
    const std::vector<uint8_t> flag_val(sflag);
    const int64_t flag_ep = sepoch;
    const int64_t lo_ep = ++nepoch, hi_ep = ++nepoch;
    for (int32_t v : low) nmark[v] = lo_ep;
    std::vector<int32_t> all;
    all.swap(nodes[x]);
    std::vector<int32_t> hi_set;
    hi_set.reserve(all.size() - low.size());
    for (int32_t v : all)
      if (nmark[v] != lo_ep) {
        nmark[v] = hi_ep;
        hi_set.push_back(v);
      }
    std::vector<std::vector<int32_t>> parts;
    components(low, lo_ep, parts);
    const size_t n_low = parts.size();
    components(hi_set, hi_ep, parts);
    // part of every node of the level set
    const int64_t part_ep = ++nepoch;
    std::unordered_map<int32_t, int32_t> pidx;
    pidx.reserve(all.size() * 2);
    for (size_t p = 0; p < parts.size(); ++p)
      for (int32_t v : parts[p]) {
        pidx[v] = (int32_t)p;
        nmark[v] = part_ep;
      }
    dead[x] = 1;
    std::vector<int> slots(parts.size());
    for (size_t p = 0; p < parts.size(); ++p) slots[p] = new_slot();
    std::vector<int> ids(parts.size());
    for (size_t p = 0; p < parts.size(); ++p) {
      const bool is_low = p < n_low;
      const int me = slots[p];
      const int64_t add_ep = ++sepoch;
      i128 G = 0, Q = 0;
      for (int32_t v : parts[p]) {
        Q += fv[v];
        for (int i = 0; i < k; ++i) {
          const int32_t u = nbr[(size_t)v * k + i];
          if (u < 0) continue;
          if (nmark[u] == part_ep) {
            const int q = pidx[u];
            if (q == (int)p) continue;
            // another part: the arc is saturated from the low part to the high one
            const bool q_low = q < (int)n_low;
            if (is_low == q_low) fail(LF_ERR_STATE, "TV path: adjacent parts on one side of a split");
            G += is_low ? acap[i] : -acap[i];
            if (smark2[slots[q]] != add_ep) {
              smark2[slots[q]] = add_ep;
              adj[me].push_back(Adj{slots[q], (uint8_t)(is_low ? 1 : 0)});
            }
            continue;
          }
          const int r = find(node_slot[u]);
          if (flag_mark[r] != flag_ep) fail(LF_ERR_STATE, "TV path: boundary neighbour missing from the adjacency");
          const bool higher = flag_val[r] != 0;
          G += higher ? acap[i] : -acap[i];
          if (smark2[r] != add_ep) {
            smark2[r] = add_ep;
            adj[me].push_back(Adj{r, (uint8_t)higher});
            adj[r].push_back(Adj{me, (uint8_t)!higher});
          }
        }
      }
      for (int32_t v : parts[p]) node_slot[v] = me;
      ids[p] = new_seg(G, Q, (int64_t)parts[p].size(), s, me);
    }
    for (size_t p = 0; p < parts.size(); ++p) {
      std::vector<int32_t> m(parts[p]);
      std::sort(m.begin(), m.end());
      members[ids[p]] = std::move(m);
      nodes[slots[p]] = std::move(parts[p]);
    }
    std::vector<Adj>().swap(adj[x]);
    children[id] = ids;
    segs[id].s_lo = s;
    segs[id].end = 2;
    segs[id].slot = -1;
    for (int c : ids) push_joins(c, s);
  }

  // ---- checks ---------------------------------------------------------------------
  Region make_region(int id, double s, int layer) {
    Region R;
    R.seg = id;
    R.layer = layer;
    R.probe = s;
    const int x = segs[id].slot;
    live_adj(x);
    R.nodes = nodes[x];
    R.beta.resize(R.nodes.size());
    for (size_t j = 0; j < R.nodes.size(); ++j) R.beta[j] = beta_of(x, R.nodes[j]);
    return R;
  }

  bool certified(int id, double s) const {
    auto it = cert.find(id);
    return it != cert.end() && it->second.down_to <= s;
  }

  // level set id ends at s: queue its check (or check everything now)
  void queue_check(int id, double s) {
    if (segs[id].n <= 1 || certified(id, s)) return;
    const std::vector<int32_t>& nd = nodes[segs[id].slot];
    int layer = -1;
    for (int l = 1; l < L && layer < 0; ++l) {
      const uint8_t* o = occ.data() + (size_t)l * n;
      bool ok = true;
      for (int32_t v : nd)
        if (o[v]) {
          ok = false;
          break;
        }
      if (ok) layer = l;
    }
    if (layer < 0) {
      flush(s);  // the live partition at s includes this level set
      return;
    }
    uint8_t* o = occ.data() + (size_t)layer * n;
    for (int32_t v : nd) o[v] = 1;
    pending.push_back(make_region(id, s, layer));
  }

  // Checks the queued level sets and the live partition at t.  Throws
  // PassAbort after recording the highest split point found.
  void flush(double t) {
    for (int x = 0; x < (int)parent.size(); ++x) {
      if (parent[x] != x || dead[x] || slot_seg[x] < 0) continue;
      const int id = slot_seg[x];
      if (segs[id].s_lo >= 0 || segs[id].n <= 1 || certified(id, t)) continue;
      pending.push_back(make_region(id, t, 0));
    }
    std::fill(occ.begin(), occ.end(), 0);
    if (pending.empty()) return;
    const auto t0 = std::chrono::steady_clock::now();
    run_checks();
    st.check_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int best = -1;
    for (size_t q = 0; q < pending.size(); ++q) {
      const Region& R = pending[q];
      if (!R.split) {
        Cert& c = cert[R.seg];
        c.birth = segs[R.seg].s_hi;
        c.down_to = R.probe;
        continue;
      }
      if (best < 0 || R.split_s > pending[best].split_s ||
          (R.split_s == pending[best].split_s && R.seg < pending[best].seg))
        best = (int)q;
    }
    if (best < 0) {
      pending.clear();
      return;
    }
    Region& R = pending[best];
    const double F = R.split_s;
    std::sort(R.last_low.begin(), R.last_low.end());
    forced[R.seg] = Forced{F, std::move(R.last_low)};
    // level sets born at or below F may get other identities in the next pass
    for (auto it = cert.begin(); it != cert.end();) {
      if (it->second.birth <= F)
        it = cert.erase(it);
      else
        ++it;
    }
    cert[R.seg] = Cert{segs[R.seg].s_hi, F};
    pending.clear();
    throw PassAbort{};
  }

  // one batched minimum cut per round; failed regions probe again at the root
  // of their violating subset until a probe passes
  void run_checks() {
    std::vector<int> live(pending.size());
    std::iota(live.begin(), live.end(), 0);
    std::vector<int64_t> cnt;
    for (int iter = 0; !live.empty(); ++iter) {
      std::fill(stage_reg.begin(), stage_reg.end(), -1);
      for (size_t q = 0; q < live.size(); ++q) {
        const Region& R = pending[live[q]];
        const size_t base = (size_t)R.layer * n;
        for (size_t j = 0; j < R.nodes.size(); ++j) {
          const int32_t v = R.nodes[j];
          stage_reg[base + v] = (int32_t)q;
          stage_r[base + v] = (int64_t)std::llround((ld)R.probe * (ld)fv[v]) + R.beta[j];
        }
      }
      cnt.assign(live.size(), 0);
      chk->pf_check_regions(stage_r.data(), stage_reg.data(), (int)live.size(), cnt.data(), stage_low.data());
      ++st.device_calls;
      if (iter == 0)
        st.checks += (int64_t)live.size();
      else
        st.dinkelbach_steps += (int64_t)live.size();
      std::vector<int> next;
      for (size_t q = 0; q < live.size(); ++q) {
        Region& R = pending[live[q]];
        const Seg& g = segs[R.seg];
        const size_t base = (size_t)R.layer * n;
        bool whole = cnt[q] == 0 || cnt[q] == g.n;
        std::vector<int32_t> S;
        i128 A1 = 0, A0 = 0;
        if (!whole) {
          // n_A g_S(s) = A1 s + A0 over the low side S of the cut
          i128 QS = 0, BS = 0, cS = 0;
          for (size_t j = 0; j < R.nodes.size(); ++j) {
            const int32_t v = R.nodes[j];
            if (!stage_low[base + v]) continue;
            S.push_back(v);
            QS += fv[v];
            BS += R.beta[j];
            for (int i = 0; i < k; ++i) {
              const int32_t u = nbr[(size_t)v * k + i];
              if (u >= 0 && stage_reg[base + u] == (int32_t)q && !stage_low[base + u]) cS += acap[i];
            }
          }
          const i128 m = (i128)S.size();
          A1 = m * g.Q - (i128)g.n * QS;
          A0 = m * g.G - (i128)g.n * (BS + cS);
          // the cut runs at ceil(mean) on values rounded to integers: a
          // violation within 2 units per node of S is rounding
          const ld val = (ld)A1 * (ld)R.probe + (ld)A0;
          if (val <= 2.0L * (ld)m * (ld)g.n) whole = true;
        }
        if (whole) {
          if (R.failed) {
            R.split = true;
            R.split_s = R.probe;
          }
          continue;
        }
        const ld root = A1 < 0 ? -(ld)A0 / (ld)A1 : (ld)INFINITY;
        R.last_low = std::move(S);
        R.failed = true;
        if (!(root < (ld)g.s_hi)) {
          // violated at its birth already
          R.split = true;
          R.split_s = g.s_hi;
          continue;
        }
        if (!(root > (ld)R.probe) || iter >= 64) {
          R.split = true;
          R.split_s = (double)std::max<ld>(root, (ld)R.probe);
          continue;
        }
        R.probe = (double)root;
        next.push_back(live[q]);
      }
      live.swap(next);
    }
  }

  // ---- one event pass from s = 1 -------------------------------------------------
  void reset_pass() {
    parent.clear();
    slot_seg.clear();
    dead.clear();
    nodes.clear();
    adj.clear();
    smark.clear();
    smark2.clear();
    sflag.clear();
    segs.clear();
    children.clear();
    members.clear();
    heap = decltype(heap)();
    pending.clear();
    std::fill(occ.begin(), occ.end(), 0);
  }

  void run_pass() {
    reset_pass();
    const int m = (int)init_nodes.size();
    node_slot.assign((size_t)n, -1);
    for (int x = 0; x < m; ++x) {
      new_slot();
      nodes[x] = init_nodes[x];
      for (int32_t v : nodes[x]) node_slot[v] = x;
    }
    for (int x = 0; x < m; ++x) new_seg(init_G[x], init_Q[x], (int64_t)init_nodes[x].size(), 1.0, x);
    // adjacency with the order of the values at s = 1 (no two neighbours tie)
    for (int x = 0; x < m; ++x) {
      const int64_t ep = ++sepoch;
      for (int32_t v : nodes[x])
        for (int i = 0; i < k; ++i) {
          const int32_t u = nbr[(size_t)v * k + i];
          if (u < 0) continue;
          const int y = node_slot[u];
          if (y == x || smark2[y] == ep) continue;
          smark2[y] = ep;
          const i128 lhs = (segs[y].G + segs[y].Q) * (i128)segs[x].n;
          const i128 rhs = (segs[x].G + segs[x].Q) * (i128)segs[y].n;
          adj[x].push_back(Adj{y, (uint8_t)(lhs > rhs)});
        }
    }
    for (int x = 0; x < m; ++x) push_joins(x, 1.0);
    while (!heap.empty()) {
      const Event e = heap.top();
      heap.pop();
      if (e.type == 1) {
        if (segs[e.a].s_lo >= 0) continue;
        apply_split(e.a, e.s, forced.at(e.a).low);
        continue;
      }
      int a = e.a, bs = e.b;
      const bool direct = segs[a].s_lo < 0 && segs[bs].s_lo < 0;
      if (!direct) {
        a = live_at(a, e.s);
        bs = live_at(bs, e.s);
        if (a < 0 || bs < 0 || a == bs) continue;
        // simultaneous joins: the successor must meet the other one here
        const ld va = segs[a].value(e.s), vb = segs[bs].value(e.s);
        if (std::fabs(va - vb) > 4.0L + 1e-15L * (std::fabs(va) + std::fabs(vb))) continue;
      }
      queue_check(a, e.s);
      queue_check(bs, e.s);
      apply_join(a, bs, e.s);
    }
    flush(0.0);
  }

  // ---- set-up ---------------------------------------------------------------------
  void initial_level_sets(const std::vector<int32_t>& reg, const std::vector<int64_t>& r) {
    // connected components of the regions
    std::vector<int32_t> comp((size_t)n, -1);
    std::vector<std::vector<int32_t>> cs;
    for (int64_t s0 = 0; s0 < n; ++s0) {
      if (comp[s0] >= 0) continue;
      const int32_t id = (int32_t)cs.size();
      cs.emplace_back();
      std::vector<int32_t>& c = cs.back();
      comp[s0] = id;
      c.push_back((int32_t)s0);
      for (size_t h = 0; h < c.size(); ++h) {
        const int32_t v = c[h];
        for (int i = 0; i < k; ++i) {
          const int32_t u = nbr[(size_t)v * k + i];
          if (u >= 0 && comp[u] < 0 && reg[u] == reg[v]) {
            comp[u] = id;
            c.push_back(u);
          }
        }
      }
    }
    const int m = (int)cs.size();
    std::vector<i128> G(m, 0), Q(m, 0);
    std::vector<int64_t> cnt(m, 0);
    for (int64_t v = 0; v < n; ++v) {
      G[comp[v]] += r[v] - fv[v];
      Q[comp[v]] += fv[v];
      ++cnt[comp[v]];
    }
    // neighbouring components with exactly equal values are one level set
    std::vector<int32_t> uf(m);
    std::iota(uf.begin(), uf.end(), 0);
    auto root = [&](int32_t x) {
      while (uf[x] != x) {
        uf[x] = uf[uf[x]];
        x = uf[x];
      }
      return x;
    };
    for (int64_t v = 0; v < n; ++v)
      for (int i = 0; i < k / 2; ++i) {
        const int32_t u = nbr[(size_t)v * k + i];
        if (u < 0 || comp[u] == comp[v]) continue;
        const int32_t a = comp[v], c = comp[u];
        if ((G[a] + Q[a]) * (i128)cnt[c] != (G[c] + Q[c]) * (i128)cnt[a]) continue;
        const int32_t ra = root(a), rc = root(c);
        if (ra != rc) uf[std::max(ra, rc)] = std::min(ra, rc);
      }
    std::vector<int32_t> idx(m, -1);
    init_nodes.clear();
    init_G.clear();
    init_Q.clear();
    for (int32_t c = 0; c < m; ++c) {
      const int32_t rc = root(c);
      if (idx[rc] < 0) {
        idx[rc] = (int32_t)init_nodes.size();
        init_nodes.emplace_back();
        init_G.push_back(0);
        init_Q.push_back(0);
      }
      const int32_t t = idx[rc];
      init_nodes[t].insert(init_nodes[t].end(), cs[c].begin(), cs[c].end());
      init_G[t] += G[c];
      init_Q[t] += Q[c];
    }
    initial.assign((size_t)n, -1);
    for (size_t t = 0; t < init_nodes.size(); ++t)
      for (int32_t v : init_nodes[t]) initial[v] = (int32_t)t;
  }
};

TvPath::TvPath(const char* kernel, int64_t nx, int64_t ny, const double* f, double lambda_max, int layers)
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
  if (layers < 0 || layers == 1 || layers > 64) fail(LF_ERR_BAD_ARG, "layers must be 0 or 2..64");
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) fail(LF_ERR_DEVICE, "no HIP device available");
  P.nx = nx;
  P.ny = ny;
  P.n = nx * ny;
  P.k = (int)kd->edges.size();
  P.L = layers ? layers : 4;
  // tv_solver.hpp:91: the network is built at max(lambda, 1e-8)
  P.lmax = std::max(lambda_max, 1e-8);
  P.b = tv_base(P.n, f);
  if (!(P.b.S > 0)) {
    P.constant = true;
    P.st.seconds = std::chrono::duration<double>(clk::now() - t0).count();
    return;
  }
  std::vector<int> offs;
  std::vector<double> w;
  lattice_edge_order(*kd, {nx, ny}, offs, w);
  const int k = P.k;
  P.nbr.assign((size_t)P.n * k, -1);
  for (int64_t v = 0; v < P.n; ++v) {
    const int64_t y = v / ny, x = v % ny;
    for (int i = 0; i < k; ++i) {
      const int64_t yy = y + offs[2 * i], xx = x + offs[2 * i + 1];
      if (yy >= 0 && yy < nx && xx >= 0 && xx < ny) P.nbr[(size_t)v * k + i] = (int32_t)(yy * ny + xx);
    }
  }
  std::vector<int64_t> pwf((size_t)k / 2);
  P.acap.assign((size_t)k, 0);
  for (int e = 0; e < k / 2; ++e) {
    pwf[e] = tv_pwf(P.b, P.lmax, w[e]);
    if (pwf[e] > ((int64_t)1 << 56) / k) fail(LF_ERR_OVERFLOW, "lambda too small for the 64-bit fixed point");
    P.acap[e] = P.acap[e + k / 2] = 2 * pwf[e];
  }
  P.fv.resize((size_t)P.n);
  for (int64_t v = 0; v < P.n; ++v) P.fv[v] = 2 * P.b.uf[v];
  auto geom = [&](int layers_) {
    LatticeGeom g;
    g.ndim = 2;
    g.k = k;
    g.D = layers_;
    g.H = nx;
    g.W = ny;
    g.dims = {nx, ny};
    g.offsets = offs;
    for (int i = 0; i < k; ++i) {
      g.dz.push_back(0);
      g.dy.push_back(offs[i * 2]);
      g.dx.push_back(offs[i * 2 + 1]);
    }
    return g;
  };
  // the exact solution at lambda_max (s = 1)
  std::vector<int32_t> reg((size_t)P.n);
  std::vector<int64_t> r((size_t)P.n);
  {
    std::unique_ptr<MinCutEngine> eng = make_mincut_engine(geom(1), 8);
    eng->set_tv_network(P.b.uf.data(), pwf.data(), 1);
    eng->pf_run_layers(lf_tuning{}, 1);
    hipStream_t s = eng->stream();
    int32_t* dreg = nullptr;
    int64_t* dr = nullptr;
    LF_HIP(hipMallocAsync((void**)&dreg, sizeof(int32_t) * P.n, s));
    LF_HIP(hipMallocAsync((void**)&dr, sizeof(int64_t) * P.n, s));
    eng->pf_regions_device(dreg, dr, s);
    LF_HIP(hipMemcpyAsync(reg.data(), dreg, sizeof(int32_t) * P.n, hipMemcpyDeviceToHost, s));
    LF_HIP(hipMemcpyAsync(r.data(), dr, sizeof(int64_t) * P.n, hipMemcpyDeviceToHost, s));
    LF_HIP(hipFreeAsync(dreg, s));
    LF_HIP(hipFreeAsync(dr, s));
    LF_HIP(hipStreamSynchronize(s));
    P.st.initial_seconds = std::chrono::duration<double>(clk::now() - t0).count();
  }
  P.initial_level_sets(reg, r);
  P.st.initial_regions = (int64_t)P.init_nodes.size();
  // the checking engine: L stacked copies of the lattice at lambda_max
  {
    std::vector<int64_t> pw((size_t)P.L * (k / 2));
    for (int l = 0; l < P.L; ++l) std::copy(pwf.begin(), pwf.end(), pw.begin() + (size_t)l * (k / 2));
    P.chk = make_mincut_engine(geom(P.L), 8);
    P.chk->set_tv_network(P.b.uf.data(), pw.data(), P.L);
  }
  P.occ.assign((size_t)P.L * P.n, 0);
  P.stage_r.assign((size_t)P.L * P.n, 0);
  P.stage_reg.assign((size_t)P.L * P.n, -1);
  P.stage_low.assign((size_t)P.L * P.n, 0);
  P.nmark.assign((size_t)P.n, 0);
  for (;;) {
    ++P.st.passes;
    try {
      P.run_pass();
      break;
    } catch (const PassAbort&) {
    }
  }
  P.st.segments = (int64_t)P.segs.size();
  for (const Seg& g : P.segs) {
    if (g.end == 1) ++P.st.joins;
    if (g.end == 2) ++P.st.splits;
  }
  P.st.joins /= 2;  // two level sets end in each join
  P.chk.reset();
  P.stage_r = {};
  P.stage_reg = {};
  P.stage_low = {};
  P.occ = {};
  P.nodes = {};
  P.adj = {};
  P.st.seconds = std::chrono::duration<double>(clk::now() - t0).count();
}

TvPath::~TvPath() {}

const TvPathStats& TvPath::stats() const { return im_->st; }
double TvPath::lambda_max() const { return im_->lmax; }
int64_t TvPath::nx() const { return im_->nx; }
int64_t TvPath::ny() const { return im_->ny; }

// getRegularizationPath (tv_solver.hpp:181-244): every pixel follows its level
// set from s = 1 down to s (traceRegularizationPath, :566-642); the level set
// alive at s is found once per level set (memoized), so a lambda costs
// O(level sets + pixels).
void TvPath::eval(int64_t n_lambda, const double* lambdas, double* R) const {
  const Impl& P = *im_;
  if (n_lambda > 0 && (!lambdas || !R)) fail(LF_ERR_BAD_ARG, "null argument");
  for (int64_t l = 0; l < n_lambda; ++l)
    if (lambdas[l] > P.lmax * (1 + 1e-12))
      fail(LF_ERR_OUT_OF_RANGE, "Given lambda values not in calculated range.");
  if (P.constant) {
    std::fill(R, R + n_lambda * P.n, P.b.M);
    return;
  }
  const int nseg = (int)P.segs.size();
  std::vector<int32_t> memo((size_t)nseg);
  std::vector<int32_t> chain;
  std::vector<double> rp((size_t)P.n);
  for (int64_t l = 0; l < n_lambda; ++l) {
    const double s = std::min(std::max(lambdas[l], 0.0) / P.lmax, 1.0);
    std::fill(memo.begin(), memo.end(), INT32_MIN);
    // the level set alive at s reached from id through joins; -(id + 1) when
    // a split comes first
    auto resolve = [&](int id) {
      chain.clear();
      int res;
      for (;;) {
        if (memo[id] != INT32_MIN) {
          res = memo[id];
          break;
        }
        chain.push_back(id);
        const Seg& g = P.segs[id];
        if (!(g.s_lo > s)) {
          res = id;
          break;
        }
        if (g.end == 2) {
          res = -(id + 1);
          break;
        }
        id = g.up;
      }
      for (int c : chain) memo[c] = res;
      return res;
    };
    for (int64_t v = 0; v < P.n; ++v) {
      int id = P.initial[v];
      for (;;) {
        const int res = resolve(id);
        if (res >= 0) {
          id = res;
          break;
        }
        const std::vector<int>& ch = P.children.at(-res - 1);
        int next = -1;
        for (int c : ch) {
          const std::vector<int32_t>& mem = P.members.at(c);
          if (std::binary_search(mem.begin(), mem.end(), (int32_t)v)) {
            next = c;
            break;
          }
        }
        if (next < 0) fail(LF_ERR_STATE, "TV path: pixel lost at a split");
        id = next;
      }
      rp[v] = (double)P.segs[id].value(s);
    }
    tv_output_path(P.b, rp.data(), P.n, R + l * P.n);
  }
}

}  // namespace lfamd
