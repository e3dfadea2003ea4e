"""Capacity checks of a device min-cut solve.  The working state is loaded at
the start of every run and its checks (negative arc capacities, capacities
beyond the element width) are read with the first global relabel's read: an
int32 network whose capacities overflow fails every run with the overflow
error, the same network solves on int64 capacities, and other handles are
unaffected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_int32_overflow_fails_every_run(oracle):
    from latticeflow_amd import LatticeFlow, workloads as W
    shape = (48, 80)
    w = W.grid4_256(seed=3, shape=shape)
    un, b = w["unary"].copy(), w["pairwise"]
    # one node's terminal capacity far beyond the int32 range of the residuals
    un[17, 1] = un[17, 0] + (1 << 31)
    lf = LatticeFlow(w["kernel"], "LatticeEnergyMinimizer", shape, numerical_precision=1.0, cap_bytes=4)
    lf.addEnergyInt(un, b)
    for _ in range(2):
        print("overflowing int32 run", flush=True)
        with pytest.raises(ValueError, match="int32 range"):
            lf.run()
    # the same network on int64 capacities is exact
    l8 = LatticeFlow(w["kernel"], "LatticeEnergyMinimizer", shape, numerical_precision=1.0, cap_bytes=8)
    l8.addEnergyInt(un, b)
    print("int64 run of the same network", flush=True)
    l8.run()
    offs = l8.directions()
    red, alpha = oracle.reparam(shape, offs, un, b)
    ref = oracle.prflow(shape, offs, red, alpha)
    assert l8.flow() == ref["flow"]
    assert np.array_equal(l8.getCut("smin").reshape(-1), ref["smin"].astype(np.int32))
    # and a fresh int32 handle on a network that fits solves as usual
    l4 = LatticeFlow(w["kernel"], "LatticeEnergyMinimizer", shape, numerical_precision=1.0, cap_bytes=4)
    l4.addEnergyInt(w["unary"], b)
    print("int32 run of the plain network", flush=True)
    l4.run()
    red, alpha = oracle.reparam(shape, offs, w["unary"], b)
    ref = oracle.prflow(shape, offs, red, alpha)
    assert l4.flow() == ref["flow"]
    assert np.array_equal(l4.getCut("smin").reshape(-1), ref["smin"].astype(np.int32))
