"""Pins the oracle's restatement of the parametric divide and conquer
(ParametricFlowSolver, pf_solver.hpp:33-208) and single-lambda TV
(TVSolver::setup/runSingleLambda, tv_solver.hpp:71-176) by the mathematics:

* the divide and conquer solves argmin 1/2 |x - r|^2 + sum_e alpha_e |x_i - x_j|
  (level-set characterisation of the cut at each threshold); on 1-D chains this
  is checked against Condat's exact direct algorithm, to one integer quantum;
* flows conserve sum r exactly;
* the TV output equals M + S (u - mean u), u = ROF(f, 0.5 w_e / (S lambda)),
  checked against an independent FP64 dual solve (box-constrained, L-BFGS-B);
* planted mistakes (floor instead of ceil threshold, a dropped pairwise weight)
  break these checks.
"""
import numpy as np


def _chain(rng, n, a):
    r = rng.integers(-10 ** 6, 10 ** 6, n)
    alpha = np.zeros((n, 2), np.int64)
    alpha[:-1, 0] = a
    alpha[1:, 1] = a
    return r, alpha


def test_dc_matches_condat_1d(oracle):
    rng = np.random.default_rng(0)
    for _ in range(30):
        n = int(rng.integers(2, 60))
        a = int(rng.integers(0, 3 * 10 ** 5))
        offs, _ = oracle.edge_order("Full1d_2", (n,))
        r, alpha = _chain(rng, n, a)
        out = oracle.pf_solve((n,), offs, r, alpha)
        x = oracle.tv1d_condat(r.astype(float), a)
        assert np.abs(out["r"] - x).max() <= 1.0
        assert out["r"].sum() == r.sum()
        # nodes of one level set share their value up to one quantum
        for l in np.unique(out["level"]):
            v = out["r"][out["level"] == l]
            assert v.max() - v.min() <= 1


def test_dc_no_edges_is_identity(oracle):
    rng = np.random.default_rng(1)
    offs, _ = oracle.edge_order("Full2d_4", (5, 6))
    r = rng.integers(-1000, 1000, 30)
    out = oracle.pf_solve((5, 6), offs, r, np.zeros((30, 4), np.int64))
    assert np.array_equal(out["r"], r)


def test_dc_2d_matches_rof(oracle):
    rng = np.random.default_rng(2)
    for kernel in ("Full2d_4", "Full2d_8"):
        shape = (5, 6)
        offs, _ = oracle.edge_order(kernel, shape)
        n = 30
        r = rng.integers(-10 ** 7, 10 ** 7, n)
        a = 10 ** 6
        alpha = np.zeros((n, len(offs)), np.int64)
        for e, v, w in oracle._pairs(shape, offs):
            alpha[v, e] = a
            alpha[w, (e + len(offs) // 2) % len(offs)] = a
        out = oracle.pf_solve(shape, offs, r, alpha)
        x = oracle.rof_dual(r.astype(float), shape, offs, np.full(len(offs) // 2, float(a)))
        assert np.abs(out["r"] - x).max() <= 2.0 + 1e-7 * np.abs(x).max()


def test_tv_single_matches_rof(oracle):
    rng = np.random.default_rng(3)
    for kernel in ("Full2d_8", "Star2d_16", "Full2d_4"):
        f = rng.random((6, 7))
        lam = 6.0
        R, st = oracle.tv_single(f, lam, kernel)
        assert st["n_levels"] > 3
        M = f.mean()
        S = 2 * np.abs(f - M).max()
        offs, w = oracle.edge_order(kernel, f.shape)
        u = oracle.rof_dual(f, f.shape, offs, 0.5 * w / (S * lam)).reshape(f.shape)
        Rx = M + S * (u - u.mean())
        assert np.abs(R - Rx).max() <= 1e-7 * np.abs(Rx).max()


def test_planted_mistakes_caught(oracle):
    rng = np.random.default_rng(4)
    n, a = 40, 150000
    offs, _ = oracle.edge_order("Full1d_2", (n,))
    r, alpha = _chain(rng, n, a)
    x = oracle.tv1d_condat(r.astype(float), a)
    # a dropped pairwise term: solve with half the weight
    bad = oracle.pf_solve((n,), offs, r, alpha // 2)["r"]
    assert np.abs(bad - x).max() > 1.0
    # a wrong sign of the unary term
    bad = oracle.pf_solve((n,), offs, -r, alpha)["r"]
    assert np.abs(bad - x).max() > 1.0
