/*
 * lf_oracle.c -- CPU oracle for the lattice push-relabel min-cut path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / the timed CPU baseline.
 *
 * It restates, in plain C with int64 arithmetic, the reference algorithm of
 * hoytak/latticeflow (paths relative to the reference root):
 *
 *   orc_reparam        NodeFiller::addE1/addE2
 *                      solver/network_flow/network_flow_node.hpp:419-485,
 *                      filling loop solver/interfacing/direct_interface.hpp:65-95,
 *                      in-bounds edge skip solver/lattices/kernellattice.hpp:304-334
 *   orc_prflow_solve   PRFlow::run   solver/network_flow/push_relabel.hpp:1303-1354
 *                        hotStart    :305-415     quickFlow  :417-509
 *                        _run        :1069-1278   nextInQueue :532-587
 *                        addToLevel  :589-652     flipNode    :677-711
 *                        finish      :717-745     push        :777-800
 *                      padded storage solver/lattices/lattice.hpp:102-111
 *   orc_dinic          an independent max-flow (Dinic's blocking flows) over
 *                      the same network; no reference counterpart, used to
 *                      cross-check the flow value and the canonical cuts.
 *   orc_cuts           canonical outputs of a final state: the flow value,
 *                      S_min (residual reachability from excess) and S_max
 *                      (complement of "can reach a sink").
 *   orc_pf_solve       ParametricFlowSolver::run (divide and conquer)
 *                      solver/parametricflow/pf_solver.hpp:33-208 with
 *                      setToMeanReductionPoint pf_flow_node.hpp:331-396
 *   orc_pf_solve_mode  the same, optionally splitting each section by its
 *                      S_min instead of the flipped set
 *   orc_tv_single      TVSolver::setup + runSingleLambda
 *                      solver/tv/tv_solver.hpp:71-112,151-176
 *   orc_tv_path        calculate2dTV(..., vector<double>) output,
 *                      tv_solver.hpp:181-244,659-681
 *
 * How each function is pinned (tests/test_oracle*.py):
 *   orc_reparam       the identity 2E(x) = K0 + sum red x + sum alpha |xi - xj|
 *                     on random labellings; planted q12 / E01-E10 mistakes fail
 *   orc_prflow_solve  flow = orc_dinic, S_min / S_max = exhaustive search on
 *   orc_dinic         lattices of <= 16 nodes; the min energy from the flow
 *   orc_cuts          equals the brute-force minimum
 *   orc_pf_solve      Condat's exact 1-D TV (one quantum), the FP64 ROF dual on
 *                     2-D lattices, exact conservation of sum r
 *   orc_pf_solve_mode the S_min split gives the flipped-set split's levels on
 *                     seeded 1-D / 2-D / 3-D cases, final ranges span <= 1
 *   orc_tv_single     the FP64 ROF dual through the lambda convention
 *   orc_tv_path       the FP64 ROF dual through the path conventions
 *                     (R - M = 2 s S (u - mean u), s = lambda / lambda_max) and
 *                     path = scaled single-lambda solves; planted convention
 *                     mistakes (halving, s = 1, lambda_max = last lambda) fail
 *
 * Network convention (the reference's doubled, "raw" units,
 * network_flow_node.hpp:234-240): per node a signed `red` (reduction: >0 is
 * remaining sink capacity, <0 is excess from the source) and per node and
 * kernel direction a residual alpha[n*k+i] of the arc n -> n+offset[i].
 * Kernel directions are in the reference's jump-sorted order: i < k/2 are the
 * positive jumps, reverse(i) = (i + k/2) mod k (kernel_base.hpp:66-71).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define ORC_MAXK 128
#define ORC_MAXD 4

/* ------------------------------------------------------------------------- */
/* small growable vectors                                                    */

typedef struct { int64_t *d; size_t n, cap; } vec_t;

static void vpush(vec_t *v, int64_t x) {
  if (v->n == v->cap) {
    v->cap = v->cap ? 2 * v->cap : 16;
    v->d = (int64_t *)realloc(v->d, v->cap * sizeof(int64_t));
  }
  v->d[v->n++] = x;
}
static void vfree(vec_t *v) { free(v->d); v->d = NULL; v->n = v->cap = 0; }

/* ------------------------------------------------------------------------- */
/* unpadded lattice helpers                                                  */

typedef struct {
  int nd, k;
  int64_t shape[ORC_MAXD];
  int64_t stride[ORC_MAXD];
  int64_t n;
  int off[ORC_MAXK][ORC_MAXD];
} ulat_t;

static int ulat_init(ulat_t *L, int nd, const int64_t *shape, int k, const int32_t *offs) {
  if (nd < 1 || nd > ORC_MAXD || k < 2 || k > ORC_MAXK || (k & 1)) return -1;
  L->nd = nd; L->k = k; L->n = 1;
  for (int d = 0; d < nd; ++d) {
    if (shape[d] < 0) return -1;
    L->shape[d] = shape[d]; L->n *= shape[d];
  }
  int64_t s = 1;
  for (int d = nd - 1; d >= 0; --d) { L->stride[d] = s; s *= shape[d]; }
  for (int i = 0; i < k; ++i)
    for (int d = 0; d < nd; ++d) L->off[i][d] = offs[i * nd + d];
  return 0;
}

/* neighbour of linear index v along direction i, or -1 when out of bounds */
static int64_t ulat_nbr(const ulat_t *L, int64_t v, int i) {
  int64_t w = 0, rem = v;
  for (int d = 0; d < L->nd; ++d) {
    int64_t c = rem / L->stride[d];
    rem -= c * L->stride[d];
    int64_t cn = c + L->off[i][d];
    if (cn < 0 || cn >= L->shape[d]) return -1;
    w += cn * L->stride[d];
  }
  return w;
}

static inline int rev_dir(int k, int i) { return (i + k / 2) % k; }

/* ------------------------------------------------------------------------- */
/* Energy -> network (NodeFiller::addE1 / addE2, network_flow_node.hpp:419-485) */

int orc_reparam(int nd, const int64_t *shape, int k, const int32_t *offs,
                const int64_t *unary, const int64_t *binary,
                int64_t *red, int64_t *alpha) {
  ulat_t L = {0};
  if (ulat_init(&L, nd, shape, k, offs)) return -1;
  const int kp = k / 2;
  memset(red, 0, sizeof(int64_t) * L.n);
  memset(alpha, 0, sizeof(int64_t) * L.n * k);
  for (int64_t v = 0; v < L.n; ++v) {
    /* addE1(e0, e1): reduction += 2 (e1 - e0) */
    red[v] += 2 * (unary[2 * v + 1] - unary[2 * v + 0]);
  }
  for (int64_t v = 0; v < L.n; ++v) {
    for (int e = 0; e < kp; ++e) {
      int64_t w = ulat_nbr(&L, v, e);
      if (w < 0) continue; /* out-of-bounds pairs are skipped (kernellattice.hpp:331) */
      const int64_t *E = binary + (v * kp + e) * 4;
      int64_t e00 = E[0], e01 = E[1], e10 = E[2], e11 = E[3];
      int64_t q1 = e10 - e00, q2 = e01 - e00, q12 = e00 + e11 - e01 - e10;
      if (q12 > 0) return -2; /* non-submodular term */
      int64_t e2 = -q12;
      red[v] += 2 * q1 + q12;
      red[w] += 2 * q2 + q12;
      alpha[v * k + e] += e2;
      alpha[w * k + rev_dir(k, e)] += e2;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Canonical outputs of a final state                                        */

/* flow = sum(max(red0,0)) - sum(max(redF,0)); S_min = nodes reachable along
 * residual arcs (res > 0) from nodes with excess (redF < 0); S_max = nodes that
 * cannot reach a node with remaining sink capacity (redF > 0). */
int orc_cuts(int nd, const int64_t *shape, int k, const int32_t *offs,
             const int64_t *red0, const int64_t *redF, const int64_t *res,
             int64_t *flow, uint8_t *smin, uint8_t *smax) {
  ulat_t L = {0};
  if (ulat_init(&L, nd, shape, k, offs)) return -1;
  int64_t f = 0;
  for (int64_t v = 0; v < L.n; ++v) {
    f += (red0[v] > 0 ? red0[v] : 0) - (redF[v] > 0 ? redF[v] : 0);
  }
  *flow = f;
  int64_t *q = (int64_t *)malloc(sizeof(int64_t) * (L.n + 1));
  if (smin) {
    int64_t qh = 0, qt = 0;
    memset(smin, 0, (size_t)L.n);
    for (int64_t v = 0; v < L.n; ++v)
      if (redF[v] < 0) { smin[v] = 1; q[qt++] = v; }
    while (qh < qt) {
      int64_t v = q[qh++];
      for (int i = 0; i < k; ++i) {
        if (res[v * k + i] <= 0) continue;
        int64_t w = ulat_nbr(&L, v, i);
        if (w < 0 || smin[w]) continue;
        smin[w] = 1; q[qt++] = w;
      }
    }
  }
  if (smax) {
    /* reach_t: nodes that can reach a sink-capable node along residual arcs */
    int64_t qh = 0, qt = 0;
    uint8_t *rt = (uint8_t *)calloc((size_t)L.n + 1, 1);
    for (int64_t v = 0; v < L.n; ++v)
      if (redF[v] > 0) { rt[v] = 1; q[qt++] = v; }
    while (qh < qt) {
      int64_t w = q[qh++];
      for (int j = 0; j < k; ++j) {
        int64_t v = ulat_nbr(&L, w, j); /* v = w + off[j]; arc v -> w is rev(j) */
        if (v < 0 || rt[v]) continue;
        if (res[v * k + rev_dir(k, j)] <= 0) continue;
        rt[v] = 1; q[qt++] = v;
      }
    }
    for (int64_t v = 0; v < L.n; ++v) smax[v] = rt[v] ? 0 : 1;
    free(rt);
  }
  free(q);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Faithful restatement of PRFlow (push_relabel.hpp)                         */

typedef struct { vec_t nodes, active; } level_t;

typedef struct {
  int k, nd;
  int64_t np;             /* padded node count */
  int64_t begin, end;     /* [begin, end): the iterated node range (lattice.hpp:108-109) */
  int64_t jumps[ORC_MAXK];
  int rev[ORC_MAXK];
  uint32_t *key_state, *height, *level_index;
  int64_t *red, *alpha;
  level_t *levels; size_t nlevels;
  size_t top_level, top_level_with_excess;
  uint32_t key;
  size_t num_flips;
  /* statistics */
  int64_t n_discharges, n_pushes, n_gaps;
} pr_t;

static void ensure_levels(pr_t *p, size_t need) {
  if (need <= p->nlevels) return;
  size_t nl = p->nlevels ? p->nlevels : 16;
  while (nl < need) nl *= 2;
  p->levels = (level_t *)realloc(p->levels, nl * sizeof(level_t));
  memset(p->levels + p->nlevels, 0, (nl - p->nlevels) * sizeof(level_t));
  p->nlevels = nl;
}

static inline int64_t pr_excess(const pr_t *p, int64_t n) { return -p->red[n]; }
static inline int64_t pr_sinkflow(const pr_t *p, int64_t n) { return p->red[n] > 0 ? p->red[n] : 0; }
static inline int64_t pr_pushcap(const pr_t *p, int64_t src, int ei) {
  int64_t a = p->alpha[src * p->k + ei];
  return a > 0 ? a : 0;
}
static inline int pr_eligible(const pr_t *p, int64_t n) { return p->key_state[n] == p->key; }

/* addToLevel<excess_status, ...> (push_relabel.hpp:589-652); status 0 none, 1 has, 2 check */
static void pr_add_to_level(pr_t *p, int64_t n, size_t new_level, int status) {
  int has_excess = (status == 2) ? (pr_excess(p, n) > 0) : (status == 1);
  if (new_level > p->top_level) ++p->top_level;
  ensure_levels(p, p->top_level + 2);
  p->height[n] = (uint32_t)new_level;
  level_t *lv = &p->levels[new_level];
  p->level_index[n] = (uint32_t)lv->nodes.n;
  vpush(&lv->nodes, n);
  if (has_excess) {
    vpush(&lv->active, n);
    if (p->top_level_with_excess < new_level) p->top_level_with_excess = new_level;
  }
}

static void pr_node_now_has_excess(pr_t *p, int64_t n) {
  vpush(&p->levels[p->height[n]].active, n);
  if (p->top_level_with_excess < p->height[n]) p->top_level_with_excess = p->height[n];
}

/* NetworkFlowNode::flipNode<0> (network_flow_node.hpp:313-354) */
static void pr_flip_raw(pr_t *p, int64_t n) {
  for (int i = 0; i < p->k; ++i) p->alpha[(n + p->jumps[i]) * p->k + p->rev[i]] *= -1;
  p->key_state[n] |= 1u;
}

static void pr_raise_and_flip(pr_t *p, size_t level);

/* PRFlow::flipNode (push_relabel.hpp:677-690) */
static void pr_flip(pr_t *p, int64_t n, int check_top_level) {
  if (check_top_level && p->levels[p->height[n]].nodes.n == 0) {
    if (p->top_level > p->height[n]) pr_raise_and_flip(p, p->height[n]);
    else p->top_level = p->height[n] - 1;
  }
  p->height[n] = 0;
  pr_flip_raw(p, n);
  ++p->num_flips;
}

/* raiseAndFlipLevelAndAbove (push_relabel.hpp:692-711) */
static void pr_raise_and_flip(pr_t *p, size_t level) {
  ++p->n_gaps;
  for (size_t lv = level; lv <= p->top_level; ++lv) {
    level_t *L = &p->levels[lv];
    for (size_t j = 0; j < L->nodes.n; ++j) pr_flip(p, L->nodes.d[j], 0);
    L->nodes.n = 0;
  }
  p->top_level = level - 1;
}

/* push<node_may_have_excess> (push_relabel.hpp:777-800, pushExcess network_flow_node.hpp:356-402) */
static void pr_push(pr_t *p, int64_t src, int64_t dest, int ei, int64_t amount, int may_have_excess) {
  if (amount == 0) return;
  ++p->n_pushes;
  p->red[src] += amount;
  p->alpha[src * p->k + ei] -= amount;
  p->red[dest] -= amount;
  p->alpha[dest * p->k + p->rev[ei]] += amount;
  int64_t dest_excess = pr_excess(p, dest);
  if (may_have_excess && dest_excess > 0 && dest_excess <= amount) pr_node_now_has_excess(p, dest);
}

/* hotStart (push_relabel.hpp:305-415) over the node list nodes[0..nn) */
static void pr_hot_start(pr_t *p, const int64_t *nodes, int64_t nn) {
  p->top_level = 1; p->top_level_with_excess = 0; p->num_flips = 0; /* initQueue :1283-1299 */
  ensure_levels(p, 4);
  vec_t orig = {0}, queue = {0};
  for (int64_t j = 0; j < nn; ++j) {
    int64_t n = nodes[j];
    if (pr_sinkflow(p, n) > 0) { vpush(&orig, n); p->height[n] = 1; }
  }
  size_t qh = 0;
#define ADD_NEIGHBORS(N_, LEVEL_)                                              \
  do {                                                                         \
    for (int i_ = 0; i_ < p->k; ++i_) {                                        \
      int64_t nn_ = (N_) + p->jumps[i_];                                       \
      if (pr_eligible(p, nn_) && p->height[nn_] == 0 &&                        \
          pr_pushcap(p, nn_, p->rev[i_]) > 0) {                                \
        pr_add_to_level(p, nn_, (LEVEL_) + 1, 2);                              \
        vpush(&queue, nn_);                                                    \
      }                                                                        \
    }                                                                          \
  } while (0)
  for (size_t j = 0; j < orig.n; ++j) ADD_NEIGHBORS(orig.d[j], 0);
  size_t current_level = 0;
  while (qh < queue.n) {
    int64_t n = queue.d[qh++];
    if (p->height[n] != current_level) { ++current_level; ensure_levels(p, current_level + 2); }
    ADD_NEIGHBORS(n, current_level);
  }
#undef ADD_NEIGHBORS
  for (int64_t j = 0; j < nn; ++j) {
    int64_t n = nodes[j];
    if (p->height[n] == 0) pr_flip(p, n, 0);
  }
  for (size_t j = 0; j < orig.n; ++j) p->height[orig.d[j]] = 0;
  vfree(&orig); vfree(&queue);
}

/* quickFlow (push_relabel.hpp:417-509) */
static int pr_quick_flow(pr_t *p) {
  for (size_t level = p->top_level; level != 1; --level) {
    level_t *lv = &p->levels[level];
    for (int64_t ni = 0; ni < (int64_t)lv->active.n; ++ni) {
      int64_t n = lv->active.d[ni];
      if (!pr_eligible(p, n) || p->height[n] != level || pr_excess(p, n) <= 0) continue;
      int64_t rem = pr_excess(p, n);
      for (int i = 0; i < p->k; ++i) {
        int64_t nn = n + p->jumps[i];
        int64_t cap = pr_pushcap(p, n, i);
        if (cap != 0 && pr_eligible(p, nn) && p->height[nn] == level - 1) {
          int64_t amount = rem < cap ? rem : cap;
          rem -= amount;
          pr_push(p, n, nn, i, amount, 1);
          lv = &p->levels[level]; /* push may grow another level's vectors only */
          if (rem == 0) {
            lv->active.d[ni] = lv->active.d[lv->active.n - 1];
            lv->active.n--;
            --ni;
            break;
          }
        }
      }
    }
    if (p->top_level_with_excess == level && lv->active.n == 0) --p->top_level_with_excess;
  }
  level_t *lv = &p->levels[1];
  for (int64_t ni = 0; ni < (int64_t)lv->active.n; ++ni) {
    int64_t n = lv->active.d[ni];
    if (!pr_eligible(p, n) || p->height[n] != 1 || pr_excess(p, n) <= 0) continue;
    int64_t rem = pr_excess(p, n);
    for (int i = 0; i < p->k; ++i) {
      int64_t nn = n + p->jumps[i];
      int64_t cap = pr_pushcap(p, n, i), sf;
      if (cap != 0 && pr_eligible(p, nn) && p->height[nn] == 0 && (sf = pr_sinkflow(p, nn)) > 0) {
        int64_t amount = rem < cap ? rem : cap;
        if (sf < amount) amount = sf;
        rem -= amount;
        pr_push(p, n, nn, i, amount, 0);
        if (rem == 0) {
          lv->active.d[ni] = lv->active.d[lv->active.n - 1];
          lv->active.n--;
          --ni;
          break;
        }
      }
    }
  }
  if (p->top_level_with_excess == 1 && lv->active.n == 0) {
    --p->top_level_with_excess;
    return 1;
  }
  return 0;
}

/* nextInQueue (push_relabel.hpp:532-587); returns 1 when done */
static int pr_next_in_queue(pr_t *p, int64_t *out, int *base_level_empty) {
  for (;;) {
    if (p->top_level_with_excess == 0) return 1;
    while (p->levels[p->top_level_with_excess].active.n == 0) {
      --p->top_level_with_excess;
      if (p->top_level_with_excess == 0) return 1;
    }
    level_t *lv = &p->levels[p->top_level_with_excess];
    int64_t n = lv->active.d[--lv->active.n];
    if (pr_excess(p, n) <= 0) continue;
    /* remove it from the level */
    uint32_t li = p->level_index[n];
    lv->nodes.d[li] = lv->nodes.d[lv->nodes.n - 1];
    p->level_index[lv->nodes.d[li]] = li;
    lv->nodes.n--;
    *base_level_empty = (lv->nodes.n == 0);
    *out = n;
    return 0;
  }
}

/* finish (push_relabel.hpp:717-745) */
static void pr_finish(pr_t *p) {
  for (size_t level = 1; level <= p->top_level; ++level) {
    level_t *lv = &p->levels[level];
    for (size_t j = 0; j < lv->nodes.n; ++j) p->height[lv->nodes.d[j]] = 0;
    lv->nodes.n = 0;
  }
  p->top_level = 1;
}

typedef struct { uint32_t h; int idx; } pnb_t;

/* _run (push_relabel.hpp:1069-1278) */
static void pr_run_loop(pr_t *p) {
  if (p->levels[1].nodes.n == 0) return;
  const int k = p->k;
  pnb_t hl[ORC_MAXK];
  int64_t nl[ORC_MAXK];
  int64_t n;
  int base_level_empty;
  for (;;) {
  next_iteration:
    if (pr_next_in_queue(p, &n, &base_level_empty)) break;
    ++p->n_discharges;
    uint32_t base_height = p->height[n];
    int64_t rem = pr_excess(p, n);
    int first_nn = k, last_valid = 0;
    for (int i = 0; i < k; ++i) {
      int64_t nn = nl[i] = n + p->jumps[i];
      int valid = (pr_pushcap(p, n, i) != 0) && pr_eligible(p, nn);
      hl[i].h = valid ? p->height[nn] + 1 : 0;
      hl[i].idx = i;
      if (valid) { --first_nn; last_valid = i; }
    }
    if (first_nn == k) { pr_flip(p, n, base_level_empty); goto next_iteration; }
    if (first_nn == k - 1) {
      hl[first_nn] = hl[last_valid];
    } else {
      /* ascending by height; the reference's tie order comes from a missing
       * generated sorting network and only changes which min cut F is chosen */
      for (int a = 1; a < k; ++a) {
        pnb_t x = hl[a];
        int b = a - 1;
        while (b >= 0 && hl[b].h > x.h) { hl[b + 1] = hl[b]; --b; }
        hl[b + 1] = x;
      }
    }
    if (base_height == 1 && hl[first_nn].h == 1) {
      for (int i = first_nn; i < k && hl[i].h == 1; ++i) {
        int index = hl[i].idx;
        int64_t nn = nl[index];
        int64_t pc = pr_pushcap(p, n, index);
        int64_t amount = pc < rem ? pc : rem;
        int64_t sf = pr_sinkflow(p, nn);
        if (sf < amount) amount = sf;
        rem -= amount;
        pr_push(p, n, nn, index, amount, 0);
        if (pc == amount) {
          pnb_t t = hl[i]; hl[i] = hl[first_nn]; hl[first_nn] = t;
          ++first_nn;
        }
        if (rem == 0) {
          pr_add_to_level(p, n, 1, 0);
          goto next_iteration;
        }
      }
      for (int i = first_nn; i < k && hl[i].h == 1; ++i) {
        pr_add_to_level(p, nl[hl[i].idx], 1, 2); /* initNodeHeight */
        hl[i].h++;
      }
      base_level_empty = 0;
    }
    for (int i = first_nn; i < k; ++i) {
      if (base_level_empty && hl[i].h > base_height) {
        pr_raise_and_flip(p, base_height);
        pr_flip(p, n, 0);
        goto next_iteration;
      }
      int index = hl[i].idx;
      int64_t nn = nl[index];
      int64_t pc = pr_pushcap(p, n, index);
      int64_t amount = pc < rem ? pc : rem;
      rem -= amount;
      pr_push(p, n, nn, index, amount, 1);
      if (rem == 0) {
        pr_add_to_level(p, n, hl[i].h, 0);
        goto next_iteration;
      }
    }
    pr_flip(p, n, base_level_empty);
  }
  pr_finish(p);
  p->top_level = 1;
  p->top_level_with_excess = 0;
}

/* Build the padded lattice (lattice.hpp:102-111): dims (n0 + 4 eb0, n1 + eb1, ...),
 * iterated range starts 2*eb0 slabs in. */
static int pr_build(pr_t *p, const ulat_t *L, int64_t *pstride, int64_t *eb) {
  memset(p, 0, sizeof(*p));
  p->k = L->k; p->nd = L->nd;
  for (int d = 0; d < L->nd; ++d) {
    eb[d] = 0;
    for (int i = 0; i < L->k; ++i) {
      int64_t a = L->off[i][d] < 0 ? -L->off[i][d] : L->off[i][d];
      if (a > eb[d]) eb[d] = a;
    }
  }
  int64_t pdims[ORC_MAXD];
  pdims[0] = L->shape[0] + 4 * eb[0];
  for (int d = 1; d < L->nd; ++d) pdims[d] = L->shape[d] + eb[d];
  int64_t s = 1;
  for (int d = L->nd - 1; d >= 0; --d) { pstride[d] = s; s *= pdims[d]; }
  p->np = s;
  p->begin = 2 * eb[0] * pstride[0];
  p->end = p->np - 2 * eb[0] * pstride[0];
  for (int i = 0; i < L->k; ++i) {
    int64_t j = 0;
    for (int d = 0; d < L->nd; ++d) j += (int64_t)L->off[i][d] * pstride[d];
    p->jumps[i] = j;
    p->rev[i] = rev_dir(L->k, i);
  }
  p->key_state = (uint32_t *)calloc(p->np, sizeof(uint32_t));
  p->height = (uint32_t *)calloc(p->np, sizeof(uint32_t));
  p->level_index = (uint32_t *)calloc(p->np, sizeof(uint32_t));
  p->red = (int64_t *)calloc(p->np, sizeof(int64_t));
  p->alpha = (int64_t *)calloc(p->np * L->k, sizeof(int64_t));
  ensure_levels(p, 64);
  return (p->key_state && p->height && p->level_index && p->red && p->alpha) ? 0 : -1;
}

static void pr_free(pr_t *p) {
  free(p->key_state); free(p->height); free(p->level_index); free(p->red); free(p->alpha);
  for (size_t i = 0; i < p->nlevels; ++i) { vfree(&p->levels[i].nodes); vfree(&p->levels[i].active); }
  free(p->levels);
}

static int64_t padded_index(const ulat_t *L, const int64_t *pstride, int64_t begin, int64_t v) {
  int64_t rem = v, w = begin;
  for (int d = 0; d < L->nd; ++d) {
    int64_t c = rem / L->stride[d];
    rem -= c * L->stride[d];
    w += c * pstride[d];
  }
  return w;
}

/* PRFlow::run on the energy network given by (red, alpha) over the unpadded lattice.
 * Outputs the final red, the true residuals |alpha| (the flip sign removed,
 * network_flow_node.hpp:274-278) and the flipped (= reference getCut) bits.
 * stats[0..3] = discharges, pushes, gaps, flips. */
int orc_prflow_solve(int nd, const int64_t *shape, int k, const int32_t *offs,
                     const int64_t *red_in, const int64_t *alpha_in,
                     int64_t *red_out, int64_t *res_out, uint8_t *flipped, int64_t *stats) {
  ulat_t L = {0};
  if (ulat_init(&L, nd, shape, k, offs)) return -1;
  pr_t p;
  int64_t pstride[ORC_MAXD] = {0}, eb[ORC_MAXD] = {0};
  if (pr_build(&p, &L, pstride, eb)) { pr_free(&p); return -3; }
  int64_t *map = (int64_t *)malloc(sizeof(int64_t) * L.n);
  for (int64_t v = 0; v < L.n; ++v) {
    int64_t w = map[v] = padded_index(&L, pstride, p.begin, v);
    p.red[w] = red_in[v];
    for (int i = 0; i < k; ++i) p.alpha[w * k + i] = alpha_in[v * k + i];
  }
  /* run(): key = makeKeyState(0, 0); hotStart + quickFlow over [begin, end); _run */
  p.key = 0;
  int64_t nn = p.end - p.begin;
  int64_t *all = (int64_t *)malloc(sizeof(int64_t) * (nn > 0 ? nn : 1));
  for (int64_t j = 0; j < nn; ++j) all[j] = p.begin + j;
  pr_hot_start(&p, all, nn);
  pr_quick_flow(&p);
  pr_run_loop(&p);
  for (int64_t v = 0; v < L.n; ++v) {
    int64_t w = map[v];
    red_out[v] = p.red[w];
    for (int i = 0; i < k; ++i) {
      int64_t a = p.alpha[w * k + i];
      res_out[v * k + i] = a < 0 ? -a : a;
    }
    flipped[v] = (uint8_t)(p.key_state[w] & 1u);
  }
  if (stats) {
    stats[0] = p.n_discharges; stats[1] = p.n_pushes; stats[2] = p.n_gaps; stats[3] = (int64_t)p.num_flips;
  }
  free(all); free(map); pr_free(&p);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Independent max-flow: Dinic on the same network (no reference counterpart) */

int orc_dinic(int nd, const int64_t *shape, int k, const int32_t *offs,
              const int64_t *red, const int64_t *alpha,
              int64_t *flow, int64_t *red_out, int64_t *res_out) {
  ulat_t L = {0};
  if (ulat_init(&L, nd, shape, k, offs)) return -1;
  const int64_t N = L.n, S = N, T = N + 1;
  int64_t *nbr = (int64_t *)malloc(sizeof(int64_t) * N * k);
  int64_t *res = res_out;
  int64_t *src = (int64_t *)malloc(sizeof(int64_t) * N);   /* s -> v residual */
  int64_t *srcb = (int64_t *)calloc(N, sizeof(int64_t));   /* v -> s residual */
  int64_t *snk = (int64_t *)malloc(sizeof(int64_t) * N);   /* v -> t residual */
  int64_t *snkb = (int64_t *)calloc(N, sizeof(int64_t));   /* t -> v residual */
  int32_t *lev = (int32_t *)malloc(sizeof(int32_t) * (N + 2));
  int32_t *it = (int32_t *)malloc(sizeof(int32_t) * (N + 2));
  int64_t *q = (int64_t *)malloc(sizeof(int64_t) * (N + 2));
  int64_t *path = (int64_t *)malloc(sizeof(int64_t) * (N + 3));
  int32_t *parc = (int32_t *)malloc(sizeof(int32_t) * (N + 3));
  for (int64_t v = 0; v < N; ++v) {
    src[v] = red[v] < 0 ? -red[v] : 0;
    snk[v] = red[v] > 0 ? red[v] : 0;
    for (int i = 0; i < k; ++i) {
      int64_t w = nbr[v * k + i] = ulat_nbr(&L, v, i);
      res[v * k + i] = (w >= 0) ? alpha[v * k + i] : 0;
    }
  }
  int64_t total = 0;
  /* arc numbering at a lattice node v: 0..k-1 lattice arcs, k -> t, k+1 -> s.
   * at s: arc index = node id. */
  for (;;) {
    for (int64_t v = 0; v < N + 2; ++v) lev[v] = -1;
    int64_t qh = 0, qt = 0;
    lev[S] = 0;
    for (int64_t v = 0; v < N; ++v)
      if (src[v] > 0) { lev[v] = 1; q[qt++] = v; }
    while (qh < qt) {
      int64_t v = q[qh++];
      if (snk[v] > 0 && lev[T] < 0) lev[T] = lev[v] + 1;
      for (int i = 0; i < k; ++i) {
        int64_t w = nbr[v * k + i];
        if (w < 0 || res[v * k + i] <= 0 || lev[w] >= 0) continue;
        lev[w] = lev[v] + 1; q[qt++] = w;
      }
    }
    if (lev[T] < 0) break;
    for (int64_t v = 0; v < N + 2; ++v) it[v] = 0;
    /* iterative DFS for blocking flow */
    for (;;) {
      int64_t depth = 0;
      path[0] = S;
      int found = 0;
      while (depth >= 0) {
        int64_t u = path[depth];
        if (u == T) { found = 1; break; }
        int advanced = 0;
        if (u == S) {
          while (it[S] < N) {
            int64_t v = it[S];
            if (src[v] > 0 && lev[v] == 1) { parc[depth] = (int32_t)v; path[++depth] = v; advanced = 1; break; }
            it[S]++;
          }
        } else {
          while (it[u] <= k) {
            int a = it[u];
            if (a < k) {
              int64_t w = nbr[u * k + a];
              if (w >= 0 && res[u * k + a] > 0 && lev[w] == lev[u] + 1) {
                parc[depth] = a; path[++depth] = w; advanced = 1; break;
              }
            } else if (snk[u] > 0 && lev[T] == lev[u] + 1) {
              parc[depth] = k; path[++depth] = T; advanced = 1; break;
            }
            it[u]++;
          }
        }
        if (!advanced) {
          lev[u] = -1; /* dead end */
          if (--depth >= 0) it[path[depth]]++;
        }
      }
      if (!found) break;
      /* bottleneck */
      int64_t b = INT64_MAX;
      for (int64_t d = 0; d < depth; ++d) {
        int64_t u = path[d]; int a = parc[d]; int64_t c;
        if (u == S) c = src[a];
        else if (a == k) c = snk[u];
        else c = res[u * k + a];
        if (c < b) b = c;
      }
      for (int64_t d = 0; d < depth; ++d) {
        int64_t u = path[d]; int a = parc[d];
        if (u == S) { src[a] -= b; srcb[a] += b; }
        else if (a == k) { snk[u] -= b; snkb[u] += b; }
        else {
          int64_t w = nbr[u * k + a];
          res[u * k + a] -= b;
          res[w * k + rev_dir(k, a)] += b;
        }
      }
      total += b;
    }
  }
  *flow = total;
  /* final state in the oracle's convention: red = remaining sink cap (>0) or
   * remaining excess that never left s (<0, i.e. unused source capacity) */
  for (int64_t v = 0; v < N; ++v) red_out[v] = snk[v] > 0 ? snk[v] : -src[v];
  free(nbr); free(src); free(srcb); free(snk); free(snkb); free(lev); free(it); free(q);
  free(path); free(parc);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Parametric divide and conquer (ParametricFlowSolver::run,                 */
/* solver/parametricflow/pf_solver.hpp:33-208) on PFFlowNode                  */
/* (solver/parametricflow/pf_flow_node.hpp:83-421).                           */
/*                                                                           */
/* Per node: r() = reduction + current_fv_offset.  Each range is set to the  */
/* threshold ceil(mean r) (minus one when it equals max r) by setOffset, cut */
/* by a keyed PRFlow section, split in place by the state bit, the "on" part  */
/* flipped back, and both parts queued (FIFO).  A range stops when            */
/* max r - min r <= 1 (setToMeanReductionPoint, pf_flow_node.hpp:331-396).    */

typedef struct { int64_t start, end; } range_t;

/* pf_flow_node.hpp:331-396; returns 1 when the range is at its mean point */
static int pf_set_to_mean(pr_t *p, int64_t *off, const int64_t *pts, int64_t s, int64_t e) {
  __int128 r_sum = 0;
  double count = 0;
  int64_t r_min = p->red[pts[s]] + off[pts[s]], r_max = r_min;
  for (int64_t i = s; i < e; ++i) {
    int64_t n = pts[i];
    int64_t r = p->red[n] + off[n];
    r_sum += r;
    if (r < r_min) r_min = r;
    if (r > r_max) r_max = r;
    count += 1;
  }
  if (r_max - r_min <= 1) return 1;
  int64_t r_new = (int64_t)ceil((double)r_sum / count);
  if (r_new == r_max) --r_new;
  for (int64_t i = s; i < e; ++i) {
    int64_t n = pts[i];
    /* setOffset (pf_flow_node.hpp:311-319): reduction += old_offset - new_offset */
    p->red[n] += off[n] - r_new;
    off[n] = r_new;
  }
  return 0;
}

/* NetworkFlowNode::flipNode<1> (network_flow_node.hpp:313-354) */
static void pr_flip_back_raw(pr_t *p, int64_t n) {
  for (int i = 0; i < p->k; ++i) p->alpha[(n + p->jumps[i]) * p->k + p->rev[i]] *= -1;
  p->key_state[n] &= ~1u;
}

/* S_min of the section just solved: the nodes reachable from nodes with
 * excess over residual arcs inside the section.  A flipped node w has the
 * arcs into it negated (pr_flip_raw), so the residual of u -> w is
 * alpha(u -> w) with its sign turned when w is flipped.  Marks in_s. */
static void pf_mark_smin(const pr_t *p, const int64_t *pts, int64_t s, int64_t e, uint8_t *in_s, int64_t *queue) {
  const uint32_t key = p->key;
  int64_t qh = 0, qt = 0;
  for (int64_t i = s; i < e; ++i) {
    int64_t n = pts[i];
    in_s[n] = 0;
  }
  for (int64_t i = s; i < e; ++i) {
    int64_t n = pts[i];
    if (pr_excess(p, n) > 0) { in_s[n] = 1; queue[qt++] = n; }
  }
  while (qh < qt) {
    int64_t u = queue[qh++];
    for (int i = 0; i < p->k; ++i) {
      int64_t w = u + p->jumps[i];
      if ((p->key_state[w] & ~1u) != key || in_s[w]) continue;
      int64_t r = p->alpha[u * p->k + i];
      if (p->key_state[w] & 1u) r = -r;
      if (r <= 0) continue;
      in_s[w] = 1;
      queue[qt++] = w;
    }
  }
}

/* smin_split: the sides of each section's cut are S_min and its complement
 * (the canonical minimal cut, as the device splits) instead of the flipped set
 * the run leaves (pf_solver.hpp:101-152). */
static int64_t pf_run(pr_t *p, int64_t *off, int64_t *pts, int64_t npts, int64_t *bound_out, int64_t *nbound,
                      int smin_split) {
  range_t *q = (range_t *)malloc(sizeof(range_t) * (size_t)(2 * npts + 4));
  uint8_t *in_s = smin_split ? (uint8_t *)calloc((size_t)p->np, 1) : NULL;
  int64_t *bq = smin_split ? (int64_t *)malloc(sizeof(int64_t) * (size_t)(npts + 1)) : NULL;
  int64_t qh = 0, qt = 0, nb = 0;
  unsigned key = 1;
  q[qt++] = (range_t){0, npts};
  bound_out[nb++] = 0;
  int64_t sections = 0;
  while (qh < qt) {
    range_t rg = q[qh++];
    if (pf_set_to_mean(p, off, pts, rg.start, rg.end)) continue;
    ++sections;
    /* prepareSection (push_relabel.hpp:1360-1381): setKey<0> */
    for (int64_t i = rg.start; i < rg.end; ++i) p->key_state[pts[i]] = (key << 1) | 0u;
    /* runSection (:1383-1463) */
    p->key = key << 1;
    pr_hot_start(p, pts + rg.start, rg.end - rg.start);
    pr_quick_flow(p);
    pr_run_loop(p);
    if (smin_split) pf_mark_smin(p, pts, rg.start, rg.end, in_s, bq);
    /* in-place separation by state (pf_solver.hpp:101-152) */
    int64_t fwd = rg.start, bck = rg.end - 1, mid = 0;
#define ST(i_) (smin_split ? (int)in_s[pts[(i_)]] : (int)(p->key_state[pts[(i_)]] & 1u))
    for (;;) {
      int done = 0;
      while (ST(fwd) == 0) {
        ++fwd;
        if (bck == fwd) { done = 1; break; }
      }
      if (!done) {
        while (ST(bck) == 1) {
          --bck;
          if (bck == fwd) { done = 1; break; }
        }
      }
      if (!done) {
        int64_t t = pts[fwd]; pts[fwd] = pts[bck]; pts[bck] = t;
        ++fwd;
        if (fwd == bck) done = 1;
      }
      if (done) {
        if (bck != rg.end && ST(bck) == 0) ++bck;
        mid = bck - rg.start;
        break;
      }
    }
#undef ST
    int64_t top = rg.end - rg.start;
    if (mid > 0) q[qt++] = (range_t){rg.start, rg.start + mid};
    if (mid < top) {
      q[qt++] = (range_t){rg.start + mid, rg.end};
      /* (S_min split: the cleanup below flips back exactly the flipped nodes) */
      if (!smin_split)
        for (int64_t i = rg.start + mid; i < rg.end; ++i) pr_flip_back_raw(p, pts[i]);
    }
    if (mid > 0 && mid < top) bound_out[nb++] = mid + rg.start;
    /* cleanupSection (:1468-1487) */
    for (int64_t i = rg.start; i < rg.end; ++i) {
      int64_t n = pts[i];
      if (p->key_state[n] & 1u) pr_flip_back_raw(p, n);
      p->key_state[n] &= 1u;
    }
    ++key;
    if (qt + 2 > 2 * npts + 4) { /* cannot happen: every queued range is nonempty */
      free(q); free(in_s); free(bq);
      return -1;
    }
  }
  /* reset every offset to 0 (pf_solver.hpp:181-182) */
  for (int64_t i = 0; i < npts; ++i) {
    int64_t n = pts[i];
    p->red[n] += off[n];
    off[n] = 0;
  }
  *nbound = nb;
  free(q); free(in_s); free(bq);
  return sections;
}

/* LatticeLevelReductions on the network (red, alpha): final r per node, the
 * level-set index of each node (sorted boundaries), and the number of
 * section solves. */
int orc_pf_solve_mode(int nd, const int64_t *shape, int k, const int32_t *offs,
                      const int64_t *red_in, const int64_t *alpha_in,
                      int64_t *r_out, int64_t *level_out, int64_t *stats, int smin_split) {
  ulat_t L = {0};
  if (ulat_init(&L, nd, shape, k, offs)) return -1;
  pr_t p;
  int64_t pstride[ORC_MAXD] = {0}, eb[ORC_MAXD] = {0};
  if (pr_build(&p, &L, pstride, eb)) { pr_free(&p); return -3; }
  int64_t *map = (int64_t *)malloc(sizeof(int64_t) * (L.n + 1));
  int64_t *pts = (int64_t *)malloc(sizeof(int64_t) * (L.n + 1));
  int64_t *off = (int64_t *)calloc((size_t)p.np, sizeof(int64_t));
  int64_t *bnd = (int64_t *)malloc(sizeof(int64_t) * (L.n + 2));
  for (int64_t v = 0; v < L.n; ++v) {
    int64_t w = map[v] = padded_index(&L, pstride, p.begin, v);
    pts[v] = w;
    p.red[w] = red_in[v];
    for (int i = 0; i < k; ++i) p.alpha[w * k + i] = alpha_in[v * k + i];
  }
  int64_t nb = 0;
  int64_t sections = L.n > 0 ? pf_run(&p, off, pts, L.n, bnd, &nb, smin_split) : 0;
  if (sections < 0) return -4;
  /* level index per node: sort boundaries, map ranges */
  for (int64_t a = 1; a < nb; ++a) {
    int64_t x = bnd[a], b = a - 1;
    while (b >= 0 && bnd[b] > x) { bnd[b + 1] = bnd[b]; --b; }
    bnd[b + 1] = x;
  }
  bnd[nb] = L.n;
  int64_t *lev_pad = (int64_t *)calloc((size_t)p.np, sizeof(int64_t));
  for (int64_t l = 0; l < nb; ++l)
    for (int64_t i = bnd[l]; i < bnd[l + 1]; ++i) lev_pad[pts[i]] = l;
  for (int64_t v = 0; v < L.n; ++v) {
    r_out[v] = p.red[map[v]];
    if (level_out) level_out[v] = lev_pad[map[v]];
  }
  if (stats) { stats[0] = sections; stats[1] = nb; }
  free(lev_pad); free(map); free(pts); free(off); free(bnd); pr_free(&p);
  return 0;
}

int orc_pf_solve(int nd, const int64_t *shape, int k, const int32_t *offs,
                 const int64_t *red_in, const int64_t *alpha_in,
                 int64_t *r_out, int64_t *level_out, int64_t *stats) {
  return orc_pf_solve_mode(nd, shape, k, offs, red_in, alpha_in, r_out, level_out, stats, 0);
}

/* ------------------------------------------------------------------------- */
/* Single-lambda TV (TVSolver::setup + runSingleLambda,                      */
/* solver/tv/tv_solver.hpp:71-112,151-176; calculate2dTV :650-655), int64:    */
/* function precision 40 bits (pf_flow_node.hpp:103-115).                     */

#define TV_FBITS 40

static int64_t tv_to_fv(double v) { return (int64_t)round(v * (double)((int64_t)1 << TV_FBITS)); }

/* shape (nx, ny); offs/w: k lattice-order directions and geocut weights.
 * Writes the reference's R (before the Python layer's factor lambda). */
int orc_tv_single(int64_t nx, int64_t ny, int k, const int32_t *offs, const double *w,
                  const double *f, double lambda, double *R, int64_t *stats) {
  const int64_t n = nx * ny;
  if (n <= 0) return -1;
  double v_min = f[0], v_max = f[0], v_sum = f[0];
  for (int64_t i = 1; i < n; ++i) {
    double v = f[i];
    v_min = v_min < v ? v_min : v;
    v_max = v_max > v ? v_max : v;
    v_sum += v;
  }
  const double M = v_sum / (double)n;
  v_min -= M;
  v_max -= M;
  const double S = 2 * (v_max > -v_min ? v_max : -v_min);
  const double max_lambda = lambda > 1e-8 ? lambda : 1e-8;
  int64_t shape[2] = {nx, ny};
  int64_t *red = (int64_t *)calloc((size_t)n, sizeof(int64_t));
  int64_t *alpha = (int64_t *)calloc((size_t)(n * k), sizeof(int64_t));
  int64_t *r = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    int64_t uf = tv_to_fv(0.5 + (f[i] - M) / S);
    red[i] += 2 * uf; /* addE1(0, uf) */
  }
  const double gsv = 0.5 / (S * max_lambda);
  ulat_t L = {0};
  if (ulat_init(&L, 2, shape, k, offs)) { free(red); free(alpha); free(r); return -1; }
  for (int64_t v = 0; v < n; ++v)
    for (int e = 0; e < k / 2; ++e) {
      int64_t u = ulat_nbr(&L, v, e);
      if (u < 0) continue;
      int64_t pwf = tv_to_fv((gsv * w[e]) / S); /* toUnshiftedFVDType */
      /* addE2(0, pwf, pwf, 0): q1 = q2 = pwf, q12 = -2 pwf */
      alpha[v * k + e] += 2 * pwf;
      alpha[u * k + rev_dir(k, e)] += 2 * pwf;
    }
  int rc = orc_pf_solve(2, shape, k, offs, red, alpha, r, NULL, stats);
  if (rc) { free(red); free(alpha); free(r); return rc; }
  double sum = 0;
  for (int64_t i = 0; i < n; ++i) {
    double v = ((double)(r[i] / 2) / (double)((int64_t)1 << TV_FBITS) - 0.5) * S + M;
    R[i] = v;
    sum += v;
  }
  for (int64_t i = 0; i < n; ++i) R[i] = (R[i] - sum / (double)n) * S + M;
  free(red); free(alpha); free(r);
  return 0;
}

/* Regularization path at given lambdas (calculate2dTV with a lambda vector,
 * tv_solver.hpp:659-681, output getRegularizationPath :181-244).  The
 * reference builds the path by events; each segment's value is linear in
 * s = lambda / lambda_max (get_r_AtLambda, tv_regpath_node.hpp:146-150), and by
 * homogeneity of the TV problem r_path(s) = s r(lambda) with r(lambda) the
 * single-lambda solution, so the path is restated from single-lambda divide
 * and conquer solves.  R: n_lambda * nx * ny. */
int orc_tv_path(int64_t nx, int64_t ny, int k, const int32_t *offs, const double *w,
                const double *f, int64_t n_lambda, const double *lambdas, double *R) {
  const int64_t n = nx * ny;
  if (n <= 0) return -1;
  if (n_lambda <= 0) return 0;
  double v_min = f[0], v_max = f[0], v_sum = f[0];
  for (int64_t i = 1; i < n; ++i) {
    double v = f[i];
    v_min = v_min < v ? v_min : v;
    v_max = v_max > v ? v_max : v;
    v_sum += v;
  }
  const double M = v_sum / (double)n;
  v_min -= M;
  v_max -= M;
  const double S = 2 * (v_max > -v_min ? v_max : -v_min);
  double lmax = 0;
  for (int64_t l = 0; l < n_lambda; ++l) {
    double lm = lambdas[l] < 0 ? 0 : lambdas[l];
    if (lm > lmax) lmax = lm;
  }
  if (lmax < 1e-8) lmax = 1e-8;
  int64_t shape[2] = {nx, ny};
  ulat_t L = {0};
  if (ulat_init(&L, 2, shape, k, offs)) return -1;
  int64_t *red = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  int64_t *alpha = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n * k));
  int64_t *r = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  for (int64_t l = 0; l < n_lambda; ++l) {
    double lam = lambdas[l] < 0 ? 0 : lambdas[l];
    double *out = R + l * n;
    if (lam <= 0) {
      for (int64_t i = 0; i < n; ++i) out[i] = M;
      continue;
    }
    for (int64_t i = 0; i < n; ++i) red[i] = 2 * tv_to_fv(0.5 + (f[i] - M) / S);
    memset(alpha, 0, sizeof(int64_t) * (size_t)(n * k));
    const double gsv = 0.5 / (S * (lam > 1e-8 ? lam : 1e-8));
    for (int64_t v = 0; v < n; ++v)
      for (int e = 0; e < k / 2; ++e) {
        int64_t u = ulat_nbr(&L, v, e);
        if (u < 0) continue;
        int64_t pwf = tv_to_fv((gsv * w[e]) / S);
        alpha[v * k + e] += 2 * pwf;
        alpha[u * k + rev_dir(k, e)] += 2 * pwf;
      }
    int rc = orc_pf_solve(2, shape, k, offs, red, alpha, r, NULL, NULL);
    if (rc) { free(red); free(alpha); free(r); return rc; }
    const double sc = lam / lmax;
    double sum = 0;
    for (int64_t i = 0; i < n; ++i) {
      double v = ((sc * (double)r[i]) / (double)((int64_t)1 << TV_FBITS) - 0.5) * S + M;
      out[i] = v;
      sum += v;
    }
    const double mean = sum / (double)n;
    for (int64_t i = 0; i < n; ++i) out[i] = (out[i] - mean) * S + M;
  }
  free(red); free(alpha); free(r);
  return 0;
}
