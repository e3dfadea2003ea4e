"""Device parametric divide and conquer (LatticeLevelReductions) and single-lambda
TV against the oracle's restatement of ParametricFlowSolver / TVSolver.

The device splits each region by the canonical minimal cut S_min, the reference
by its schedule-dependent flipped set; both are minimum cuts, so the level
values agree up to the integer quantum (|r_gpu - r_ref| <= 2 raw units) and the
TV output to well within the 1e-9 relative bound."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _levels_gpu(kernel, shape, red_alpha_unary_binary):
    from latticeflow_amd import LatticeFlow
    un, b = red_alpha_unary_binary
    lf = LatticeFlow(kernel, "LatticeLevelReductions", shape, numerical_precision=1.0, cap_bytes=8)
    lf.addEnergyInt(un, b)
    lf.run()
    return lf.levels().reshape(-1), lf.getCut().reshape(-1), lf.stats()


@pytest.mark.parametrize("kernel,nd", [("Full1d_2", 1), ("Full2d_4", 2), ("Full2d_8", 2), ("Star3d_26", 3)])
def test_level_reductions_vs_oracle(oracle, kernel, nd):
    rng = np.random.default_rng(7 + nd)
    for trial in range(4):
        if nd == 1:
            shape = (int(rng.integers(2, 200)),)
        elif nd == 2:
            shape = tuple(int(x) for x in rng.integers(2, 40, 2))
        else:
            shape = tuple(int(x) for x in rng.integers(2, 10, 3))
        offs, _ = oracle.edge_order(kernel, shape)
        n = int(np.prod(shape))
        kp = len(offs) // 2
        un = rng.integers(0, 10 ** 6, (n, 2))
        b = np.zeros((n, kp, 4), np.int64)
        w = rng.integers(0, 2 * 10 ** 5, (n, kp))
        b[:, :, 1] = w
        b[:, :, 2] = w
        print("case", kernel, shape, flush=True)
        r, cut, st = _levels_gpu(kernel, shape, (un, b))
        red, alpha = oracle.reparam(shape, offs, un, b)
        ref = oracle.pf_solve(shape, offs, red, alpha)
        assert r.sum() == ref["r"].sum() == red.sum()
        assert np.abs(r - ref["r"]).max() <= 2, (kernel, shape, np.abs(r - ref["r"]).max())
        assert np.array_equal(cut, (r < 0).astype(np.int32))


@pytest.mark.parametrize("kernel,nd", [("Full1d_2", 1), ("Full2d_4", 2), ("Full2d_8", 2), ("Star3d_26", 3)])
def test_level_regions_vs_oracle_smin(oracle, kernel, nd):
    """Against the oracle splitting every section by S_min as the device does:
    on each final range of the oracle the device values have the same exact
    sum and, sorted, are the same values (a final range holds two adjacent
    values, so its count and sum fix the multiset; only which node holds which
    value depends on the max flow each solver found)."""
    rng = np.random.default_rng(21 + nd)
    for trial in range(4):
        if nd == 1:
            shape = (int(rng.integers(2, 200)),)
        elif nd == 2:
            shape = tuple(int(x) for x in rng.integers(2, 40, 2))
        else:
            shape = tuple(int(x) for x in rng.integers(2, 10, 3))
        offs, _ = oracle.edge_order(kernel, shape)
        n = int(np.prod(shape))
        kp = len(offs) // 2
        un = rng.integers(0, 10 ** 6, (n, 2))
        b = np.zeros((n, kp, 4), np.int64)
        w = rng.integers(0, 2 * 10 ** 5, (n, kp))
        b[:, :, 1] = w
        b[:, :, 2] = w
        print("case", kernel, shape, flush=True)
        r, cut, st = _levels_gpu(kernel, shape, (un, b))
        red, alpha = oracle.reparam(shape, offs, un, b)
        ref = oracle.pf_solve(shape, offs, red, alpha, smin_split=True)
        for lv in np.unique(ref["level"]):
            sel = ref["level"] == lv
            assert r[sel].sum() == ref["r"][sel].sum(), (kernel, shape, lv)
            assert np.array_equal(np.sort(r[sel]), np.sort(ref["r"][sel])), (kernel, shape, lv)
        assert np.abs(r - ref["r"]).max() <= 1


@pytest.mark.parametrize("kernel", ["Full2d_8", "Star2d_16", "Full2d_4"])
def test_tv_single_vs_oracle(oracle, kernel):
    from latticeflow_amd.pylatticeflow import _tv_single
    rng = np.random.default_rng(3)
    for lam in (0.5, 2.0, 8.0, 40.0):
        f = rng.random((37, 45))
        R, st = _tv_single(f, lam, kernel)
        Rr, _ = oracle.tv_single(f, lam, kernel)
        rel = np.abs(R - Rr).max() / np.abs(Rr).max()
        print(kernel, lam, rel, st["reserved"], flush=True)
        assert rel <= 1e-9


def test_calculate2dTV_python_layer(oracle):
    from latticeflow_amd import calculate2dTV
    rng = np.random.default_rng(4)
    f = rng.random((20, 24))
    R = calculate2dTV(f, 3.0)
    Rr, _ = oracle.tv_single(f, 3.0, "Star2d_16")
    assert np.abs(R - 3.0 * Rr).max() <= 1e-9 * np.abs(3.0 * Rr).max()


@pytest.mark.parametrize("kernel", ["Full2d_8", "Star2d_16"])
def test_tv_path_vs_oracle(oracle, kernel):
    from latticeflow_amd.pylatticeflow import _tv_path
    from latticeflow_amd import workloads as W
    f = W.tv_voronoi(seed=11, shape=(48, 40), cells=12)
    lams = [0.02, 0.3, 1.0, 2.0, 0.0, 5.0]
    R, st = _tv_path(f, lams, kernel)
    Rr = oracle.tv_path(f, lams, kernel)
    rel = np.abs(R - Rr).max() / np.abs(Rr).max()
    print(kernel, rel, st, flush=True)
    assert rel <= 1e-9


def test_calculate2dTVPath_python_layer(oracle):
    from latticeflow_amd import calculate2dTVPath
    rng = np.random.default_rng(5)
    f = rng.random((16, 18))
    lams = np.array([4.0, 1.0, 2.0])
    R = calculate2dTVPath(f, lams)
    Rr = oracle.tv_path(f, lams, "Star2d_16") / lams[-1]
    assert np.abs(R - Rr).max() <= 1e-9 * np.abs(Rr).max()
    assert calculate2dTVPath(f, []).shape == (0, 16, 18)
