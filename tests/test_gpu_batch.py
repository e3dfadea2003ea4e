"""Batched min-cuts (many lattices, one device solve) against the CPU oracle.

Every lattice of a batch must give the flow value and the S_min / S_max cuts
that the oracle's PRFlow restatement gives for that lattice alone, bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_energies(rng, n, kp, cap=20, wcap=10):
    un = rng.integers(0, cap, (n, 2))
    b = np.zeros((n, kp, 4), np.int64)
    b[:, :, 0] = rng.integers(0, 3, (n, kp))
    b[:, :, 3] = rng.integers(0, 3, (n, kp))
    b[:, :, 1] = rng.integers(0, wcap, (n, kp)) + b[:, :, 0]
    b[:, :, 2] = rng.integers(0, wcap, (n, kp)) + b[:, :, 3]
    return un, b


def _check_batch(oracle, bt, energies, shapes):
    flows = bt.flows()
    smin = bt.getCuts("smin")
    smax = bt.getCuts("smax")
    assert bt.stats()["cycles"] >= 1
    tot = 0
    for i, (shape, (un, b)) in enumerate(zip(shapes, energies)):
        offs = oracle.edge_order(bt._kernel, shape)[0]
        red, alpha = oracle.reparam(shape, offs, un, b)
        ref = oracle.prflow(shape, offs, red, alpha)
        assert flows[i] == ref["flow"], (i, shape, flows[i], ref["flow"])
        assert np.array_equal(smin[i].reshape(-1), ref["smin"].astype(np.int32)), (i, shape)
        assert np.array_equal(smax[i].reshape(-1), ref["smax"].astype(np.int32)), (i, shape)
        tot += ref["flow"]
    from latticeflow_amd._lib import lib
    import ctypes
    f = ctypes.c_int64()
    assert lib().lf_get_flow(bt._h, ctypes.byref(f)) == 0
    assert f.value == tot


@pytest.mark.parametrize("kernel", ["Full2d_4", "Full2d_8", "Star2d_16"])
@pytest.mark.parametrize("cap_bytes", [4, 8])
def test_batch_random(oracle, kernel, cap_bytes):
    from latticeflow_amd import LatticeFlowBatch
    rng = np.random.default_rng(hash((kernel, cap_bytes, "batch")) & 0xffff)
    shapes = [tuple(int(x) for x in rng.integers(1, 80, 2)) for _ in range(23)] + [(1, 1), (3, 120)]
    print("batch", kernel, cap_bytes, shapes, flush=True)
    bt = LatticeFlowBatch(kernel, shapes, numerical_precision=1.0, cap_bytes=cap_bytes)
    energies = []
    for i, s in enumerate(shapes):
        offs = oracle.edge_order(kernel, s)[0]
        assert np.array_equal(bt.edges(i), offs[: len(offs) // 2])
        un, b = _random_energies(rng, s[0] * s[1], len(offs) // 2)
        energies.append((un, b))
        bt.addEnergyInt(i, un, b)
    bt.run()
    _check_batch(oracle, bt, energies, shapes)


def test_batch_seg8_autotune(oracle):
    """Segmentation lattices of uneven sizes (the batched workload at small sides);
    every tuning the autotuner tries keeps the results."""
    from latticeflow_amd import LatticeFlowBatch, workloads as W
    inst = W.batch_instances(seed=7, count=12, lo=32, hi=256)
    shapes = [s for s, _ in inst]
    print("batch seg8", shapes, flush=True)
    bt = LatticeFlowBatch("Full2d_8", shapes, numerical_precision=1.0, cap_bytes=4)
    energies = []
    for i, (s, seed) in enumerate(inst):
        w = W.seg8(seed=seed, shape=s)
        energies.append((w["unary"], w["pairwise"]))
        bt.addEnergyInt(i, w["unary"], w["pairwise"])
    bt.run()
    _check_batch(oracle, bt, energies, shapes)
    bt.autotune(reps=1)
    bt.run()
    _check_batch(oracle, bt, energies, shapes)


def test_batch_matches_single(oracle):
    """A batch of one lattice equals the plain handle on the same lattice."""
    from latticeflow_amd import LatticeFlow, LatticeFlowBatch, workloads as W
    w = W.seg8(seed=9, shape=(200, 333))
    bt = LatticeFlowBatch("Full2d_8", [w["shape"]], numerical_precision=1.0, cap_bytes=4)
    bt.addEnergyInt(0, w["unary"], w["pairwise"])
    bt.run()
    lf = LatticeFlow("Full2d_8", "LatticeEnergyMinimizer", w["shape"], numerical_precision=1.0, cap_bytes=4)
    lf.addEnergyInt(w["unary"], w["pairwise"])
    lf.run()
    assert bt.flows()[0] == lf.flow()
    assert np.array_equal(bt.getCuts()[0], lf.getCut())
