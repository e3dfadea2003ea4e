"""The Cython extension latticeflow_amd/ext/pylatticeflow (the reference's
extensions/python/pylatticeflow.pyx API over the C++ interface headers): the
module's names, errors and shapes without a GPU, the CPU solver through it, and
on the device the min-cut and both TV calls against the oracle and against the
ctypes layer."""
import numpy as np
import pytest


def _mod():
    from latticeflow_amd.ext import pylatticeflow
    return pylatticeflow


def test_module_surface():
    P = _mod()
    assert P.getKernelDimension("Star3d_26") == 3
    with pytest.raises(ValueError, match="not a valid kernel name"):
        P.getKernelDimension("Nope2d_3")
    with pytest.raises(ValueError, match="not a valid solver name"):
        P.LatticeFlow("Full2d_4", "NoSolver", (3, 3))
    with pytest.raises(ValueError, match="2-dimensional; a shape of 3"):
        P.LatticeFlow("Full2d_4", "BKGCEnergyMinimizer", (3, 3, 3))
    lf = P.LatticeFlow("Full2d_8", "BKGCEnergyMinimizer", (4, 5))
    assert (lf.ndim, lf.shape, lf.size, lf.nFillingEdges) == (2, (4, 5), 20, 4)
    assert lf.edges.tolist() == [[0, 1], [1, -1], [1, 0], [1, 1]]
    with pytest.raises(ValueError, match="Unary energy"):
        lf.addEnergy(np.zeros((19, 2)), np.zeros((20, 4, 4)))
    with pytest.raises(ValueError, match="wrong number of edges"):
        lf.addEnergy(np.zeros((20, 2)), np.zeros((20, 3, 4)))


def test_cpu_solver_matches_oracle(oracle):
    P = _mod()
    rng = np.random.default_rng(3)
    shape = (23, 31)
    offs, _ = oracle.edge_order("Full2d_8", shape)
    un = rng.integers(0, 50, (23 * 31, 2)) * 1e-6
    b = np.zeros((23 * 31, 4, 4))
    b[:, :, 1] = b[:, :, 2] = rng.integers(0, 20, (23 * 31, 4)) * 1e-6
    lf = P.LatticeFlow("Full2d_8", "BKGCEnergyMinimizer", shape)
    lf.addEnergy(un, b)
    lf.run()
    ui, bi = np.round(un * 1e6).astype(np.int64), np.round(b * 1e6).astype(np.int64)
    red, alpha = oracle.reparam(shape, offs, ui, bi)
    ref = oracle.prflow(shape, offs, red, alpha)
    assert np.array_equal(lf.getCut().reshape(-1), ref["smin"].astype(np.int32))


@pytest.mark.gpu
def test_device_solver_and_tv(oracle):
    P = _mod()
    from latticeflow_amd import workloads as W
    import latticeflow_amd as L
    w = W.seg8(seed=21, shape=(200, 150))
    lf = P.LatticeFlow("Full2d_8", "LatticeEnergyMinimizer", w["shape"], numerical_precision=1.0)
    lf.addEnergy(w["unary"], w["pairwise"])
    lf.run()
    red, alpha = oracle.reparam(w["shape"], oracle.edge_order("Full2d_8", w["shape"])[0], w["unary"], w["pairwise"])
    ref = oracle.prflow(w["shape"], oracle.edge_order("Full2d_8", w["shape"])[0], red, alpha)
    assert np.array_equal(lf.getCut().reshape(-1), ref["smin"].astype(np.int32))
    assert lf.runTime() > 0
    f = W.tv_voronoi(seed=8, shape=(27, 33), cells=7)
    R = P.calculate2dTV(f, 1.5)
    Rr, _ = oracle.tv_single(f, 1.5, "Star2d_16")
    assert np.abs(R - 1.5 * Rr).max() <= 1e-9 * np.abs(1.5 * Rr).max()
    assert np.abs(R - L.calculate2dTV(f, 1.5)).max() == 0
    lams = np.array([0.3, 2.0, 0.9])
    Rp = P.calculate2dTVPath(f, lams)
    Rpr = oracle.tv_path(f, lams, "Star2d_16") / lams[-1]
    assert np.abs(Rp - Rpr).max() <= 1e-9 * np.abs(Rpr).max()
    assert P.calculate2dTVPath(f, []).shape == (0, 27, 33)
