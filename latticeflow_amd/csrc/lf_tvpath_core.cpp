// The event loop of the TV regularization path (TVSolver::run,
// solver/tv/tv_solver.hpp:114-149; _constructInitialRegPathsFromSolvedLattice
// :300-564; TVRegPathSegment, tv_regpath_node.hpp:67-1028).
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
// 208-266).  From s = 1 down to 0 the partition changes by events, taken from
// a max-heap over s:
//   * join: two neighbouring level sets whose lines meet (closed form,
//     lambdaOfJoin / setupFromJoin, :837-984); G and Q add;
//   * split: a level set stops being one (calculateSplit / applySplit,
//     :416-830).
//
// Every level set is checked when it is created, as the reference does.  A is
// a level set at s exactly when no subset S of A has a violation
//   g_S(s) = sum_S (r_A(s) - r_v(s)) - c(S, A \ S) > 0,
// which one minimum cut of A alone at its mean value decides.  Each g_S is
// linear in s, and A is a level set at its birth, so A is one on [t, birth]
// exactly when it is one at t: one check covers an interval.  A new level set
// is checked down to its first join point t (its earliest possible end); if
// it outlives t (that join was void), a check event at t checks it again down
// to its next join point.  When a check fails, the split point is found by
// Dinkelbach iterations (the root of the last violating subset's g_S is the
// next probe, until a probe passes) and a split event is scheduled there.
// Level sets of at most host_max nodes are cut on the host (a Dinic max
// flow); larger ones go to the device in one batch per event.
//
// Level sets are kept connected: the components of a level set, of a split's
// low part and of its high part each carry the value of the whole at that
// point (the violations of the parts add up to zero), so they are separate
// level sets from there on.

#include "lf_tvpath_core.hpp"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <numeric>
#include <queue>
#include <unordered_map>
#include <vector>

#include "lf_error.hpp"

namespace lfamd {

namespace {
typedef __int128 i128;
typedef long double ld;
using clk = std::chrono::steady_clock;

struct Seg {
  int64_t n = 0;
  i128 Q = 0, G = 0;
  double s_hi = 1;   // born at
  double s_lo = -1;  // ended at (-1: lasts to s = 0)
  int end = 0;       // 1: joined into `up`; 2: split (parts in Impl::children)
  int up = -1;
  int slot = -1;     // slot while alive
  double cert = 1;   // checked to be a level set down to here
  ld value(ld s) const { return ((ld)G + s * (ld)Q) / (ld)n; }
};

struct Event {
  double s;
  int type;  // 0 join, 1 split, 2 check again (the level set outlived its checked range)
  int a, b;
};

// max-heap order: higher s first; at one s joins, then splits, then checks;
// then by ids
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

// a level set under check
struct Region {
  int seg = 0;
  double probe = 0;
  double top = 1;  // a level set down to here already
  std::vector<int32_t> nodes;
  std::vector<int64_t> beta;      // boundary flow per node (constant over the life)
  std::vector<int32_t> last_low;  // violating subset of the last failed probe
  bool failed = false;
  bool split = false;
  double split_s = 0;
};

// ceil(a / b), b > 0
int64_t ceil_div(i128 a, int64_t b) {
  i128 q = a / b;
  if (a % b != 0 && a > 0) ++q;
  return (int64_t)q;
}

// Minimum cut of one level set on the host, by the rule of RegionChecker:
// Dinic's max flow from the nodes below the threshold to those above it over
// the level set's own arcs; the low side is what the remaining source
// capacity reaches (the minimal source side).
class HostCut {
 public:
  HostCut(const int32_t* nbr, int k, const int64_t* acap, int64_t n) : nbr_(nbr), k_(k), acap_(acap), lid_((size_t)n, -1) {}

  int64_t cut(const std::vector<int32_t>& nodes, const int64_t* rv, std::vector<uint8_t>& low) {
    const int m = (int)nodes.size();
    low.assign((size_t)m, 0);
    if (m == 0) return 0;
    i128 sum = 0;
    int64_t mn = INT64_MAX, mx = INT64_MIN;
    for (int j = 0; j < m; ++j) {
      sum += rv[j];
      mn = std::min(mn, rv[j]);
      mx = std::max(mx, rv[j]);
    }
    if (mx - mn <= 1) return 0;
    int64_t thr = ceil_div(sum, m);
    if (thr == mx) --thr;
    for (int j = 0; j < m; ++j) lid_[nodes[j]] = j;
    cs_.assign((size_t)m, 0);
    ct_.assign((size_t)m, 0);
    res_.assign((size_t)m * k_, 0);
    nb_.assign((size_t)m * k_, -1);
    for (int j = 0; j < m; ++j) {
      const int64_t e = thr - rv[j];
      if (e > 0) cs_[j] = e;
      else ct_[j] = -e;
      for (int i = 0; i < k_; ++i) {
        const int32_t u = nbr_[(size_t)nodes[j] * k_ + i];
        if (u >= 0 && lid_[u] >= 0) {
          nb_[(size_t)j * k_ + i] = lid_[u];
          res_[(size_t)j * k_ + i] = acap_[i];
        }
      }
    }
    for (int j = 0; j < m; ++j) lid_[nodes[j]] = -1;
    dinic(m);
    // the low side: reachable from the remaining source capacity
    std::vector<int32_t>& q = q_;
    q.clear();
    for (int j = 0; j < m; ++j)
      if (cs_[j] > 0) {
        low[j] = 1;
        q.push_back(j);
      }
    for (size_t h = 0; h < q.size(); ++h) {
      const int j = q[h];
      for (int i = 0; i < k_; ++i) {
        const int l = nb_[(size_t)j * k_ + i];
        if (l >= 0 && res_[(size_t)j * k_ + i] > 0 && !low[l]) {
          low[l] = 1;
          q.push_back(l);
        }
      }
    }
    return (int64_t)q.size();
  }

 private:
  int rev(int i) const { return i < k_ / 2 ? i + k_ / 2 : i - k_ / 2; }

  void dinic(int m) {
    level_.assign((size_t)m, -1);
    it_.assign((size_t)m, 0);
    for (;;) {
      // levels from the source
      std::fill(level_.begin(), level_.end(), -1);
      q_.clear();
      for (int j = 0; j < m; ++j)
        if (cs_[j] > 0) {
          level_[j] = 1;
          q_.push_back(j);
        }
      bool reach = false;
      for (size_t h = 0; h < q_.size(); ++h) {
        const int j = q_[h];
        if (ct_[j] > 0) reach = true;
        for (int i = 0; i < k_; ++i) {
          const int l = nb_[(size_t)j * k_ + i];
          if (l >= 0 && level_[l] < 0 && res_[(size_t)j * k_ + i] > 0) {
            level_[l] = level_[j] + 1;
            q_.push_back(l);
          }
        }
      }
      if (!reach) return;
      std::fill(it_.begin(), it_.end(), 0);
      // blocking flow along level-increasing arcs
      for (int j0 = 0; j0 < m; ++j0) {
        if (level_[j0] != 1) continue;
        while (cs_[j0] > 0) {
          pn_.clear();
          pd_.clear();
          int u = j0;
          bool found = false;
          for (;;) {
            if (ct_[u] > 0) {
              found = true;
              break;
            }
            int i = it_[u];
            int v = -1;
            for (; i < k_; ++i) {
              const int l = nb_[(size_t)u * k_ + i];
              if (l >= 0 && res_[(size_t)u * k_ + i] > 0 && level_[l] == level_[u] + 1) {
                v = l;
                break;
              }
            }
            it_[u] = i;
            if (v >= 0) {
              pn_.push_back(u);
              pd_.push_back(i);
              u = v;
              continue;
            }
            level_[u] = -1;  // dead end for this phase
            if (pn_.empty()) break;
            u = pn_.back();
            pn_.pop_back();
            pd_.pop_back();
            ++it_[u];
          }
          if (!found) break;
          int64_t f = std::min(cs_[j0], ct_[u]);
          for (size_t t = 0; t < pn_.size(); ++t) f = std::min(f, res_[(size_t)pn_[t] * k_ + pd_[t]]);
          cs_[j0] -= f;
          ct_[u] -= f;
          for (size_t t = 0; t < pn_.size(); ++t) {
            const int a = pn_[t], i = pd_[t];
            const int b = nb_[(size_t)a * k_ + i];
            res_[(size_t)a * k_ + i] -= f;
            res_[(size_t)b * k_ + rev(i)] += f;
          }
        }
      }
    }
  }

  const int32_t* nbr_;
  int k_;
  const int64_t* acap_;
  std::vector<int32_t> lid_;
  std::vector<int64_t> cs_, ct_, res_;
  std::vector<int32_t> nb_, level_, it_, q_, pn_, pd_;
};
}  // namespace

struct TvPathCore::Impl {
  TvPathProblem P;
  RegionChecker* dev = nullptr;
  int64_t host_max = 0;
  TvPathStats* st = nullptr;
  int k = 0;
  int64_t n = 0;
  std::unique_ptr<HostCut> hc;

  std::vector<Seg> segs;
  std::vector<int32_t> initial;
  std::unordered_map<int, std::vector<int>> children;     // split level set -> parts
  std::unordered_map<int, std::vector<int32_t>> members;  // part of a split -> sorted nodes
  std::unordered_map<int, std::pair<double, std::vector<int32_t>>> plan;  // scheduled split: point, low part
  std::unordered_map<int, double> first_;  // new level set -> its first join point

  // live level sets: slots joined by union-find
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

  // device staging
  std::vector<int64_t> stage_r;
  std::vector<int32_t> stage_reg;
  std::vector<uint8_t> stage_low;

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
      const int32_t u = P.nbr[(size_t)v * k + i];
      if (u < 0) continue;
      const int r = find(node_slot[u]);
      if (r == x) continue;
      if (smark[r] != sepoch) fail(LF_ERR_STATE, "TV path: boundary neighbour missing from the adjacency");
      beta += sflag[r] ? P.acap[i] : -P.acap[i];
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

  // join events of a new level set; returns the highest (its earliest
  // possible end), 0 when none
  double push_joins(int id, double s_cur) {
    const int x = segs[id].slot;
    double first = 0;
    for (const Adj& e : live_adj(x)) {
      const int o = slot_seg[e.slot];
      double s;
      if (join_point(id, o, e.higher != 0, s_cur, &s)) {
        heap.push(Event{s, 0, std::min(id, o), std::max(id, o)});
        first = std::max(first, s);
      }
    }
    return first;
  }

  // the highest point below s_cur where id meets a present neighbour, 0 when none
  double next_join(int id, double s_cur) {
    const int x = segs[id].slot;
    double first = 0;
    for (const Adj& e : live_adj(x)) {
      double s;
      if (join_point(id, slot_seg[e.slot], e.higher != 0, s_cur, &s) && s < s_cur) first = std::max(first, s);
    }
    return first;
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
    ++st->joins;
    first_[c] = push_joins(c, s);
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
          const int32_t u = P.nbr[(size_t)v * k + i];
          if (u >= 0 && nmark[u] == ep) {
            nmark[u] = seen;
            comp.push_back(u);
          }
        }
      }
    }
  }

  std::vector<int> apply_split(int id, double s, const std::vector<int32_t>& low) {
    const int x = segs[id].slot;
    live_adj(x);
    const int64_t flag_ep = sepoch;
    std::unordered_map<int, uint8_t> flag;  // the old level set's neighbours and their order
    for (const Adj& e : adj[x]) flag[e.slot] = e.higher;
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
    (void)flag_ep;
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
        Q += P.fv[v];
        for (int i = 0; i < k; ++i) {
          const int32_t u = P.nbr[(size_t)v * k + i];
          if (u < 0) continue;
          if (nmark[u] == part_ep) {
            const int q = pidx[u];
            if (q == (int)p) continue;
            // another part: the arc is saturated from the low part to the high one
            const bool q_low = q < (int)n_low;
            if (is_low == q_low) fail(LF_ERR_STATE, "TV path: adjacent parts on one side of a split");
            G += is_low ? P.acap[i] : -P.acap[i];
            if (smark2[slots[q]] != add_ep) {
              smark2[slots[q]] = add_ep;
              adj[me].push_back(Adj{slots[q], (uint8_t)(is_low ? 1 : 0)});
            }
            continue;
          }
          const int r = find(node_slot[u]);
          auto f = flag.find(r);
          if (f == flag.end()) fail(LF_ERR_STATE, "TV path: boundary neighbour missing from the adjacency");
          const bool higher = f->second != 0;
          G += higher ? P.acap[i] : -P.acap[i];
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
    ++st->splits;
    for (int c : ids) first_[c] = push_joins(c, s);
    return ids;
  }

  // ---- checks -------------------------------------------------------------------
  Region make_region(int id, double probe, double top) {
    Region R;
    R.seg = id;
    R.probe = probe;
    R.top = top;
    const int x = segs[id].slot;
    live_adj(x);
    R.nodes = nodes[x];
    R.beta.resize(R.nodes.size());
    for (size_t j = 0; j < R.nodes.size(); ++j) R.beta[j] = beta_of(x, R.nodes[j]);
    return R;
  }

  int64_t value_at(const Region& R, size_t j) const {
    return (int64_t)std::llround((ld)R.probe * (ld)P.fv[R.nodes[j]]) + R.beta[j];
  }

  // After a check of R at R.probe whose minimal cut has the low side S:
  // decides R or moves its probe.  Returns true when R is to be checked again.
  bool step(Region& R, std::vector<int32_t>& S, int iter) {
    const Seg& g = segs[R.seg];
    bool whole = S.empty() || (int64_t)S.size() == g.n;
    i128 A1 = 0, A0 = 0;
    if (!whole) {
      // n_A g_S(s) = A1 s + A0
      const int64_t in_ep = ++nepoch, low_ep = ++nepoch;
      for (int32_t v : R.nodes) nmark[v] = in_ep;
      for (int32_t v : S) nmark[v] = low_ep;
      i128 QS = 0, BS = 0, cS = 0;
      for (size_t j = 0; j < R.nodes.size(); ++j) {
        const int32_t v = R.nodes[j];
        if (nmark[v] != low_ep) continue;
        QS += P.fv[v];
        BS += R.beta[j];
        for (int i = 0; i < k; ++i) {
          const int32_t u = P.nbr[(size_t)v * k + i];
          if (u >= 0 && nmark[u] == in_ep) cS += P.acap[i];
        }
      }
      const i128 m = (i128)S.size();
      A1 = m * g.Q - (i128)g.n * QS;
      A0 = m * g.G - (i128)g.n * (BS + cS);
      // the cut runs at ceil(mean) on values rounded to integers: a violation
      // within 2 units per node of S is rounding
      const ld val = (ld)A1 * (ld)R.probe + (ld)A0;
      if (val <= 2.0L * (ld)m * (ld)g.n) whole = true;
    }
    if (whole) {
      if (R.failed) {
        R.split = true;
        R.split_s = R.probe;
      }
      return false;
    }
    const ld root = A1 < 0 ? -(ld)A0 / (ld)A1 : (ld)INFINITY;
    R.last_low.swap(S);
    R.failed = true;
    if (!(root < (ld)R.top)) {
      // violated where it is known to be a level set: rounding at its birth
      R.split = true;
      R.split_s = R.top;
      return false;
    }
    if (!(root > (ld)R.probe) || iter >= 64) {
      R.split = true;
      R.split_s = (double)std::max<ld>(root, (ld)R.probe);
      return false;
    }
    R.probe = (double)root;
    return true;
  }

  void host_verify(Region& R) {
    std::vector<int64_t> rv(R.nodes.size());
    std::vector<uint8_t> low;
    std::vector<int32_t> S;
    for (int iter = 0;; ++iter) {
      for (size_t j = 0; j < R.nodes.size(); ++j) rv[j] = value_at(R, j);
      const auto t0 = clk::now();
      hc->cut(R.nodes, rv.data(), low);
      st->host_check_seconds += std::chrono::duration<double>(clk::now() - t0).count();
      ++st->checks;
      ++st->host_checks;
      if (iter > 0) ++st->dinkelbach_steps;
      S.clear();
      for (size_t j = 0; j < R.nodes.size(); ++j)
        if (low[j]) S.push_back(R.nodes[j]);
      if (!step(R, S, iter)) return;
    }
  }

  void device_verify(std::vector<Region>& regs, std::vector<int> live) {
    std::vector<int64_t> cnt;
    std::vector<int32_t> S;
    for (int iter = 0; !live.empty(); ++iter) {
      for (size_t q = 0; q < live.size(); ++q) {
        const Region& R = regs[live[q]];
        for (size_t j = 0; j < R.nodes.size(); ++j) {
          stage_reg[R.nodes[j]] = (int32_t)q;
          stage_r[R.nodes[j]] = value_at(R, j);
        }
      }
      cnt.assign(live.size(), 0);
      const auto t0 = clk::now();
      dev->check(stage_r.data(), stage_reg.data(), (int)live.size(), cnt.data(), stage_low.data());
      st->device_check_seconds += std::chrono::duration<double>(clk::now() - t0).count();
      ++st->device_calls;
      st->checks += (int64_t)live.size();
      if (iter > 0) st->dinkelbach_steps += (int64_t)live.size();
      std::vector<int> next;
      for (size_t q = 0; q < live.size(); ++q) {
        Region& R = regs[live[q]];
        S.clear();
        for (int32_t v : R.nodes) {
          if (stage_low[v]) S.push_back(v);
          stage_reg[v] = -1;
        }
        if (step(R, S, iter)) next.push_back(live[q]);
      }
      live.swap(next);
    }
  }

  // Checks level sets down to their earliest possible end (`down` each, at
  // most `top`); schedules the splits found, or a check again where a level
  // set that outlives that point has to be checked further.
  void verify(const std::vector<int>& ids, const std::vector<double>& down, const std::vector<double>& top) {
    std::vector<Region> regs;
    for (size_t q = 0; q < ids.size(); ++q) {
      const int id = ids[q];
      if (segs[id].n > 1) {
        regs.push_back(make_region(id, down[q], top[q]));
      } else {
        segs[id].cert = 0;  // a single node never splits
      }
    }
    std::vector<int> on_dev;
    for (size_t q = 0; q < regs.size(); ++q) {
      if (dev && (host_max < 0 || (int64_t)regs[q].nodes.size() > host_max))
        on_dev.push_back((int)q);
      else
        host_verify(regs[q]);
    }
    if (!on_dev.empty()) device_verify(regs, on_dev);
    for (Region& R : regs) {
      if (!R.split) {
        segs[R.seg].cert = R.probe;
        if (R.probe > 0) heap.push(Event{R.probe, 2, R.seg, -1});
        continue;
      }
      segs[R.seg].cert = R.split_s;
      std::sort(R.last_low.begin(), R.last_low.end());
      plan[R.seg] = std::make_pair(R.split_s, std::move(R.last_low));
      heap.push(Event{R.split_s, 1, R.seg, -1});
    }
  }

  // new level sets, checked down to their first join point
  void verify_new(const std::vector<int>& ids) {
    std::vector<double> down, top;
    for (int id : ids) {
      down.push_back(first_[id]);
      top.push_back(segs[id].s_hi);
    }
    verify(ids, down, top);
  }

  // ---- set-up and the event loop ----------------------------------------------------
  void initial_level_sets(std::vector<std::vector<int32_t>>& sets, std::vector<i128>& Gs, std::vector<i128>& Qs) {
    const std::vector<int32_t>& reg = P.reg;
    // connected components of the level sets
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
          const int32_t u = P.nbr[(size_t)v * k + i];
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
      G[comp[v]] += P.r[v] - P.fv[v];
      Q[comp[v]] += P.fv[v];
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
        const int32_t u = P.nbr[(size_t)v * k + i];
        if (u < 0 || comp[u] == comp[v]) continue;
        const int32_t a = comp[v], c = comp[u];
        if ((G[a] + Q[a]) * (i128)cnt[c] != (G[c] + Q[c]) * (i128)cnt[a]) continue;
        const int32_t ra = root(a), rc = root(c);
        if (ra != rc) uf[std::max(ra, rc)] = std::min(ra, rc);
      }
    std::vector<int32_t> idx(m, -1);
    for (int32_t c = 0; c < m; ++c) {
      const int32_t rc = root(c);
      if (idx[rc] < 0) {
        idx[rc] = (int32_t)sets.size();
        sets.emplace_back();
        Gs.push_back(0);
        Qs.push_back(0);
      }
      const int32_t t = idx[rc];
      sets[t].insert(sets[t].end(), cs[c].begin(), cs[c].end());
      Gs[t] += G[c];
      Qs[t] += Q[c];
    }
  }

  void run() {
    std::vector<std::vector<int32_t>> sets;
    std::vector<i128> Gs, Qs;
    initial_level_sets(sets, Gs, Qs);
    const int m = (int)sets.size();
    st->initial_regions = m;
    node_slot.assign((size_t)n, -1);
    initial.assign((size_t)n, -1);
    for (int x = 0; x < m; ++x) {
      new_slot();
      for (int32_t v : sets[x]) {
        node_slot[v] = x;
        initial[v] = x;
      }
      nodes[x] = std::move(sets[x]);
    }
    for (int x = 0; x < m; ++x) new_seg(Gs[x], Qs[x], (int64_t)nodes[x].size(), 1.0, x);
    // adjacency with the order of the values at s = 1 (no two neighbours tie)
    for (int x = 0; x < m; ++x) {
      const int64_t ep = ++sepoch;
      for (int32_t v : nodes[x])
        for (int i = 0; i < k; ++i) {
          const int32_t u = P.nbr[(size_t)v * k + i];
          if (u < 0) continue;
          const int y = node_slot[u];
          if (y == x || smark2[y] == ep) continue;
          smark2[y] = ep;
          const i128 lhs = (segs[y].G + segs[y].Q) * (i128)segs[x].n;
          const i128 rhs = (segs[x].G + segs[x].Q) * (i128)segs[y].n;
          adj[x].push_back(Adj{y, (uint8_t)(lhs > rhs)});
        }
    }
    for (int x = 0; x < m; ++x) first_[x] = push_joins(x, 1.0);
    {
      std::vector<int> all(m);
      std::iota(all.begin(), all.end(), 0);
      verify_new(all);
      first_.clear();
    }
    while (!heap.empty()) {
      const Event e = heap.top();
      heap.pop();
      if (e.type == 1) {
        auto it = plan.find(e.a);
        if (segs[e.a].s_lo >= 0 || it == plan.end() || it->second.first != e.s) continue;
        const std::vector<int32_t> low = std::move(it->second.second);
        plan.erase(it);
        verify_new(apply_split(e.a, e.s, low));
        first_.clear();
        continue;
      }
      if (e.type == 2) {
        // the level set lives on below the point it was checked to: check it
        // down to its next join point
        const Seg& g = segs[e.a];
        if (g.s_lo >= 0 || g.cert != e.s) continue;
        verify({e.a}, {next_join(e.a, e.s)}, {e.s});
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
      if (segs[a].cert > e.s || segs[bs].cert > e.s) fail(LF_ERR_STATE, "TV path: join below the checked range");
      const int c = apply_join(a, bs, e.s);
      verify_new({c});
      first_.clear();
    }
    st->segments = (int64_t)segs.size();
    // the live structure is no longer needed
    parent = {};
    slot_seg = {};
    dead = {};
    nodes = {};
    adj = {};
    node_slot = {};
    smark = {};
    smark2 = {};
    sflag = {};
    nmark = {};
    stage_r = {};
    stage_reg = {};
    stage_low = {};
    plan.clear();
  }
};

TvPathCore::TvPathCore(TvPathProblem prob, RegionChecker* device, int64_t host_max, TvPathStats* stats)
    : im_(new Impl) {
  Impl& I = *im_;
  I.P = std::move(prob);
  I.dev = device;
  I.host_max = host_max;
  I.st = stats;
  I.k = I.P.k;
  I.n = I.P.n;
  if (I.n >= ((int64_t)1 << 31)) fail(LF_ERR_UNSUPPORTED, "the event path handles fewer than 2^31 pixels");
  if ((int64_t)I.P.nbr.size() != I.n * I.k || (int64_t)I.P.fv.size() != I.n || (int64_t)I.P.reg.size() != I.n ||
      (int64_t)I.P.r.size() != I.n || (int)I.P.acap.size() != I.k)
    fail(LF_ERR_BAD_ARG, "TV path: inconsistent problem arrays");
  I.hc.reset(new HostCut(I.P.nbr.data(), I.k, I.P.acap.data(), I.n));
  I.nmark.assign((size_t)I.n, 0);
  if (device) {
    I.stage_r.assign((size_t)I.n, 0);
    I.stage_reg.assign((size_t)I.n, -1);
    I.stage_low.assign((size_t)I.n, 0);
  }
  I.run();
  I.hc.reset();
}

TvPathCore::~TvPathCore() {}

void TvPathCore::values(double s, double* rp) const {
  const Impl& I = *im_;
  const int nseg = (int)I.segs.size();
  std::vector<int32_t> memo((size_t)nseg, INT32_MIN);
  std::vector<int32_t> chain;
  // the level set alive at s reached from id through joins; -(id + 1) when a
  // split comes first
  auto resolve = [&](int id) {
    chain.clear();
    int res;
    for (;;) {
      if (memo[id] != INT32_MIN) {
        res = memo[id];
        break;
      }
      chain.push_back(id);
      const Seg& g = I.segs[id];
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
  for (int64_t v = 0; v < I.n; ++v) {
    int id = I.initial[v];
    for (;;) {
      const int res = resolve(id);
      if (res >= 0) {
        id = res;
        break;
      }
      int next = -1;
      for (int c : I.children.at(-res - 1)) {
        const std::vector<int32_t>& mem = I.members.at(c);
        if (std::binary_search(mem.begin(), mem.end(), (int32_t)v)) {
          next = c;
          break;
        }
      }
      if (next < 0) fail(LF_ERR_STATE, "TV path: pixel lost at a split");
      id = next;
    }
    rp[v] = (double)I.segs[id].value(s);
  }
}

TvTrace TvPathCore::trace_table() const {
  const Impl& I = *im_;
  TvTrace t;
  t.initial = I.initial;
  const size_t ns = I.segs.size();
  t.s_lo.resize(ns);
  t.a.resize(ns);
  t.b.resize(ns);
  t.up.resize(ns);
  t.mem_begin.push_back(0);
  for (size_t g = 0; g < ns; ++g) {
    const Seg& sg = I.segs[g];
    t.s_lo[g] = sg.s_lo;
    t.a[g] = (double)((ld)sg.G / (ld)sg.n);
    t.b[g] = (double)((ld)sg.Q / (ld)sg.n);
    if (sg.end == 1) {
      t.up[g] = sg.up;
    } else if (sg.end == 2) {
      const int si = (int)t.part_begin.size();
      t.up[g] = -(si + 1);
      t.part_begin.push_back((int32_t)t.part_seg.size());
      for (int c : I.children.at((int)g)) {
        t.part_seg.push_back(c);
        const std::vector<int32_t>& mem = I.members.at(c);
        t.members.insert(t.members.end(), mem.begin(), mem.end());
        t.mem_begin.push_back((int64_t)t.members.size());
      }
      t.part_end.push_back((int32_t)t.part_seg.size());
    } else {
      t.up[g] = INT32_MIN;
    }
  }
  return t;
}

}  // namespace lfamd
