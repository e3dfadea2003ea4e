"""Pins the oracle's restatement of the parametric divide and conquer
(ParametricFlowSolver, pf_solver.hpp:33-208) and single-lambda TV
(TVSolver::setup/runSingleLambda, tv_solver.hpp:71-176) by the mathematics:

* the divide and conquer solves argmin 1/2 |x - r|^2 + sum_e alpha_e |x_i - x_j|
  (level-set characterisation of the cut at each threshold); on 1-D chains this
  is checked against Condat's exact direct algorithm, to one integer quantum;
* flows conserve sum r exactly;
* the TV output equals M + S (u - mean u), u = ROF(f, 0.5 w_e / (S lambda)),
  checked against an independent FP64 dual solve (box-constrained, L-BFGS-B);
* the regularization-path output (getRegularizationPath, tv_solver.hpp:181-244)
  equals M + 2 s S (u - mean u) with s = lambda / lambda_max, against the same
  ROF solves, and agrees with single-lambda solves the way the reference's own
  path check compares them;
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


def _path_from_rof(oracle, f, lams, kernel):
    """The regularization-path output of the reference's conventions, computed
    from independent FP64 ROF solves.

    The path network is built once at lambda_max (tv_solver.hpp:71-112, pairwise
    0.5 w_e / (S lambda_max)); a point of the path at s = lambda / lambda_max
    scales the function values by s (PFFlowNode::fv(scale) = multFVScale(base_fv,
    scale), pf_flow_node.hpp:160-174,227-230), so by the homogeneity of the ROF
    problem its raw value is s times the raw single-lambda value at lambda
    (argmin 1/2|x - s y|^2 + c TV(x) = s argmin 1/2|x - y|^2 + (c/s) TV(x)).  The
    traced raw values go through the solver's toFValue without the raw-to-energy
    halving that runSingleLambda applies (tv_solver.hpp:219 vs :164), then are
    re-centred and scaled by S again (:229-240).  With u = ROF(f, 0.5 w / (S
    lambda)) in image units this gives R - M = 2 s S (u - mean u); lambda <= 0 is
    the constant M (tv_solver.hpp:669-674: negative lambdas count as 0)."""
    f = np.asarray(f, np.float64)
    M = f.mean()
    S = 2 * np.abs(f - M).max()
    lam = np.maximum(np.asarray(lams, np.float64), 0.0)
    lmax = max(lam.max(), 1e-8)
    offs, w = oracle.edge_order(kernel, f.shape)
    out = np.empty((lam.size,) + f.shape)
    for l, lm in enumerate(lam):
        if lm <= 0:
            out[l] = M
            continue
        u = oracle.rof_dual(f, f.shape, offs, 0.5 * w / (S * lm)).reshape(f.shape)
        out[l] = M + 2 * (lm / lmax) * S * (u - u.mean())
    return out


def _path_err(R, Rx, f):
    """max deviation of the path from its pin, relative to the image's scale
    S = 2 max|f - M| (a slice's spread is at most 2 s S)"""
    f = np.asarray(f)
    return np.abs(R - Rx).max() / (2 * np.abs(f - f.mean()).max())


def test_tv_path_matches_rof(oracle):
    """Every slice of the oracle's path restatement against independent ROF
    solves mapped through the reference's path conventions."""
    rng = np.random.default_rng(6)
    for kernel, shape in (("Full2d_8", (6, 7)), ("Star2d_16", (5, 6)), ("Full2d_4", (7, 5))):
        f = rng.random(shape)
        # unsorted, with a repeat, a zero and a negative (counted as 0)
        lams = [3.0, 12.0, 0.7, 6.0, 0.0, 3.0, -1.0]
        R = oracle.tv_path(f, lams, kernel)
        Rx = _path_from_rof(oracle, f, lams, kernel)
        assert _path_err(R, Rx, f) <= 1e-7, (kernel, _path_err(R, Rx, f))


def test_tv_path_consistent_with_single(oracle):
    """The reference's own check of the path (scripts/test_reg_path.py:37-88,
    solver/test_case2.cpp): each path slice against a single-lambda solve.
    Under its conventions R_path - M = 2 (lambda / lambda_max) (R_single - M)."""
    rng = np.random.default_rng(9)
    f = rng.random((9, 8))
    lams = [0.05, 0.9, 1.5, 2.5, 4.0]
    R = oracle.tv_path(f, lams, "Full2d_8")
    M = f.mean()
    S = 2 * np.abs(f - M).max()
    spread = 0.0
    for l, lm in enumerate(lams):
        Rs, _ = oracle.tv_single(f, lm, "Full2d_8")
        want = M + 2 * (lm / max(lams)) * (Rs - M)
        assert np.abs(R[l] - want).max() <= 1e-9 * S
        spread = max(spread, np.abs(want - M).max())
    assert spread > 0.1 * S  # not every slice is fully merged


def test_tv_path_planted_mistakes_caught(oracle):
    """Plausible mistakes in the path conventions, planted into the oracle's
    output, fail the ROF pin: the raw-unit halving of the single-lambda output
    applied to the path, s = 1 (no homogeneity scaling), lambda_max taken from
    the last lambda of the list instead of the largest."""
    rng = np.random.default_rng(7)
    f = rng.random((6, 6))
    lams = [4.0, 1.0, 2.0]
    R = oracle.tv_path(f, lams, "Full2d_8")
    Rx = _path_from_rof(oracle, f, lams, "Full2d_8")
    assert _path_err(R, Rx, f) <= 1e-7
    M = f.mean()
    lam = np.asarray(lams)
    halved = M + (R - M) / 2
    no_scale = M + (R - M) / (lam / lam.max())[:, None, None]
    last_max = M + (R - M) * (lam.max() / lam[-1])
    for bad in (halved, no_scale, last_max):
        assert _path_err(bad, Rx, f) > 1e-3


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


def test_pf_smin_split_mode(oracle):
    """The oracle's divide and conquer with sections split by S_min (the
    device's rule) gives the same levels as the reference's split by the
    flipped set on these seeded cases, the same exact sums, and final ranges
    whose values differ by at most one (the stopping rule)."""
    rng = np.random.default_rng(11)
    for kernel, shape in [("Full1d_2", (140,)), ("Full2d_4", (21, 26)), ("Full2d_8", (25, 19)),
                          ("Star3d_26", (6, 5, 7))]:
        for _ in range(3):
            offs, _ = oracle.edge_order(kernel, shape)
            n = int(np.prod(shape))
            kp = len(offs) // 2
            un = rng.integers(0, 10 ** 6, (n, 2))
            b = np.zeros((n, kp, 4), np.int64)
            w = rng.integers(0, 2 * 10 ** 5, (n, kp))
            b[:, :, 1] = w
            b[:, :, 2] = w
            red, alpha = oracle.reparam(shape, offs, un, b)
            f = oracle.pf_solve(shape, offs, red, alpha)
            s = oracle.pf_solve(shape, offs, red, alpha, smin_split=True)
            assert s["r"].sum() == red.sum()
            assert np.array_equal(f["r"], s["r"]), (kernel, shape)
            for lv in np.unique(s["level"]):
                rr = s["r"][s["level"] == lv]
                assert rr.max() - rr.min() <= 1

