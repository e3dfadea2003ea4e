"""The CPU cross-check solver "BKGCEnergyMinimizer" (Boykov-Kolmogorov max-flow,
written from the 2004 paper; the reference wraps Kolmogorov's code for it,
solver/energy_minimization/bk_graphcuts.hpp:65-163).

It runs on the host, so these tests need no GPU: through the C ABI it must give
the oracle's flow value and both canonical cuts bit for bit (the oracle's
PRFlow restatement is itself pinned by Dinic and brute force), on every kernel
dimension, with the reference's replace-not-add energy semantics of the BK
filler (bk_graphcuts.hpp:41-59)."""
import numpy as np
import pytest


def _energies(rng, shape, offs, cap=30, wcap=12):
    n = int(np.prod(shape))
    kp = len(offs) // 2
    un = rng.integers(0, cap, (n, 2))
    b = np.zeros((n, kp, 4), np.int64)
    b[:, :, 0] = rng.integers(0, 3, (n, kp))
    b[:, :, 3] = rng.integers(0, 3, (n, kp))
    b[:, :, 1] = rng.integers(0, wcap, (n, kp)) + b[:, :, 0]
    b[:, :, 2] = rng.integers(0, wcap, (n, kp)) + b[:, :, 3]
    return un, b


def _bk(kernel, shape, un, b):
    from latticeflow_amd import LatticeFlow
    lf = LatticeFlow(kernel, "BKGCEnergyMinimizer", shape, numerical_precision=1.0)
    lf.addEnergyInt(un, b)
    lf.run()
    return lf


@pytest.mark.parametrize("kernel,nd", [("Full1d_2", 1), ("Full2d_4", 2), ("Full2d_8", 2), ("Star2d_16", 2),
                                       ("Full2d_12", 2), ("Full3d_6", 3), ("Star3d_26", 3)])
def test_bk_matches_oracle(oracle, kernel, nd):
    rng = np.random.default_rng(abs(hash(kernel)) % 9973)
    for trial in range(5):
        if nd == 1:
            shape = (int(rng.integers(1, 200)),)
        elif nd == 2:
            shape = tuple(int(x) for x in rng.integers(1, 40, 2))
        else:
            shape = tuple(int(x) for x in rng.integers(1, 12, 3))
        offs, _ = oracle.edge_order(kernel, shape)
        un, b = _energies(rng, shape, offs)
        lf = _bk(kernel, shape, un, b)
        red, alpha = oracle.reparam(shape, lf.directions(), un, b)
        ref = oracle.prflow(shape, lf.directions(), red, alpha)
        assert lf.flow() == ref["flow"], (kernel, shape)
        assert np.array_equal(lf.getCut("smin").reshape(-1), ref["smin"].astype(np.int32)), (kernel, shape)
        assert np.array_equal(lf.getCut("smax").reshape(-1), ref["smax"].astype(np.int32)), (kernel, shape)


def test_bk_brute_force(oracle):
    rng = np.random.default_rng(3)
    for _ in range(30):
        shape = (int(rng.integers(1, 4)), int(rng.integers(1, 5)))
        offs, _ = oracle.edge_order("Full2d_8", shape)
        un, b = _energies(rng, shape, offs, cap=9, wcap=6)
        lf = _bk("Full2d_8", shape, un, b)
        emin, smin, smax = oracle.brute_force(shape, offs, un, b)
        cut = lf.getCut().reshape(-1)
        assert oracle.energy(shape, offs, un, b, cut) == emin
        assert np.array_equal(cut, smin.astype(np.int32))
        assert np.array_equal(lf.getCut("smax").reshape(-1), smax.astype(np.int32))


def test_bk_segmentation_lattice(oracle):
    from latticeflow_amd import workloads as W
    w = W.seg8(seed=5, shape=(96, 130))
    lf = _bk(w["kernel"], w["shape"], w["unary"], w["pairwise"])
    red, alpha = oracle.reparam(w["shape"], lf.directions(), w["unary"], w["pairwise"])
    ref = oracle.prflow(w["shape"], lf.directions(), red, alpha)
    assert lf.flow() == ref["flow"]
    assert np.array_equal(lf.getCut().reshape(-1), ref["smin"].astype(np.int32))


def test_bk_energies_replace(oracle):
    """A second addEnergy replaces the first (the BK filler assigns)."""
    from latticeflow_amd import LatticeFlow
    rng = np.random.default_rng(11)
    shape = (13, 17)
    offs, _ = oracle.edge_order("Full2d_4", shape)
    un1, b1 = _energies(rng, shape, offs)
    un2, b2 = _energies(rng, shape, offs)
    lf = LatticeFlow("Full2d_4", "BKGCEnergyMinimizer", shape, numerical_precision=1.0)
    lf.addEnergyInt(un1, b1)
    lf.addEnergyInt(un2, b2)
    lf.run()
    ref = _bk("Full2d_4", shape, un2, b2)
    assert lf.flow() == ref.flow()
    assert np.array_equal(lf.getCut(), ref.getCut())
    assert lf.runTime() >= 0


def test_bk_rejects_nonsubmodular():
    from latticeflow_amd import LatticeFlow
    lf = LatticeFlow("Full2d_4", "BKGCEnergyMinimizer", (2, 2), numerical_precision=1.0)
    un = np.zeros((4, 2), np.int64)
    b = np.zeros((4, 2, 4), np.int64)
    b[0, 0] = (5, 0, 0, 5)
    lf.addEnergyInt(un, b)
    with pytest.raises(ValueError):
        lf.run()


def test_bk_4d(oracle):
    rng = np.random.default_rng(4)
    for _ in range(4):
        shape = tuple(int(x) for x in rng.integers(1, 6, 4))
        offs, _ = oracle.edge_order("Full4d_8", shape)
        un, b = _energies(rng, shape, offs)
        lf = _bk("Full4d_8", shape, un, b)
        red, alpha = oracle.reparam(shape, lf.directions(), un, b)
        ref = oracle.prflow(shape, lf.directions(), red, alpha)
        assert lf.flow() == ref["flow"]
        assert np.array_equal(lf.getCut("smin").reshape(-1), ref["smin"].astype(np.int32))
        assert np.array_equal(lf.getCut("smax").reshape(-1), ref["smax"].astype(np.int32))
