"""Device min-cut parity against the CPU oracle (flow value and canonical cuts, bit-exact).

Every case goes through the C ABI of liblatticeflow_amd.so (via the Python host
mirror).  The oracle is the PRFlow restatement in oracle/lf_oracle.c, itself
pinned against Dinic and brute force in tests/test_oracle.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_energies(rng, shape, offs, cap=20, wcap=10):
    n = int(np.prod(shape))
    kp = len(offs) // 2
    un = rng.integers(0, cap, (n, 2))
    b = np.zeros((n, kp, 4), np.int64)
    b[:, :, 0] = rng.integers(0, 3, (n, kp))
    b[:, :, 3] = rng.integers(0, 3, (n, kp))
    b[:, :, 1] = rng.integers(0, wcap, (n, kp)) + b[:, :, 0]
    b[:, :, 2] = rng.integers(0, wcap, (n, kp)) + b[:, :, 3]
    return un, b


def _solve_gpu(kernel, shape, un, b, cap_bytes=8, tuning=None):
    from latticeflow_amd import LatticeFlow
    lf = LatticeFlow(kernel, "LatticeEnergyMinimizer", shape, numerical_precision=1.0, cap_bytes=cap_bytes)
    if tuning:
        lf.setTuning(**tuning)
    lf.addEnergyInt(un, b)
    lf.run()
    return lf.flow(), lf.getCut("smin").reshape(-1), lf.getCut("smax").reshape(-1), lf


def _check(oracle, kernel, shape, un, b, cap_bytes=8, tuning=None):
    lfl = _solve_gpu(kernel, shape, un, b, cap_bytes, tuning)
    offs = lfl[3].directions()
    red, alpha = oracle.reparam(shape, offs, un, b)
    ref = oracle.prflow(shape, offs, red, alpha)
    flow, smin, smax, _ = lfl
    assert flow == ref["flow"], (kernel, shape, flow, ref["flow"])
    assert np.array_equal(smin, ref["smin"].astype(np.int32)), (kernel, shape)
    assert np.array_equal(smax, ref["smax"].astype(np.int32)), (kernel, shape)
    return ref


@pytest.mark.parametrize("kernel,nd", [("Full2d_4", 2), ("Full2d_8", 2), ("Full2d_12", 2), ("Star2d_16", 2),
                                       ("Full3d_6", 3), ("Full3d_18", 3), ("Star3d_26", 3), ("Full1d_2", 1)])
@pytest.mark.parametrize("cap_bytes", [4, 8])
def test_random_small(oracle, kernel, nd, cap_bytes):
    rng = np.random.default_rng(hash((kernel, cap_bytes)) & 0xffff)
    for trial in range(6):
        if nd == 1:
            shape = (int(rng.integers(1, 300)),)
        elif nd == 2:
            shape = tuple(int(x) for x in rng.integers(1, 90, 2))
        else:
            shape = tuple(int(x) for x in rng.integers(1, 20, 3))
        offs, _ = oracle.edge_order(kernel, shape)
        un, b = _random_energies(rng, shape, offs)
        print("case", kernel, cap_bytes, shape, flush=True)
        _check(oracle, kernel, shape, un, b, cap_bytes, tuning={"tile_shape": 1 + trial % 3})


def test_grid4_256_config(oracle):
    from latticeflow_amd import workloads as W
    w = W.grid4_256(seed=1)
    _check(oracle, w["kernel"], w["shape"], w["unary"], w["pairwise"], cap_bytes=4)


def test_seg8_512(oracle):
    from latticeflow_amd import workloads as W
    w = W.seg8(seed=5, shape=(512, 512))
    _check(oracle, w["kernel"], w["shape"], w["unary"], w["pairwise"], cap_bytes=4)


def test_blobs26_small(oracle):
    from latticeflow_amd import workloads as W
    w = W.blobs26(seed=7, shape=(40, 48, 56))
    _check(oracle, w["kernel"], w["shape"], w["unary"], w["pairwise"], cap_bytes=8)


@pytest.mark.parametrize("tile_shape", [1, 2, 3])
@pytest.mark.parametrize("kernel,shape", [("Full2d_8", (150, 210)), ("Star2d_16", (77, 130)), ("Star3d_26", (20, 33, 41))])
def test_tile_shapes(oracle, tile_shape, kernel, shape):
    rng = np.random.default_rng(tile_shape * 31 + len(shape))
    offs, _ = oracle.edge_order(kernel, shape)
    un, b = _random_energies(rng, shape, offs, cap=200, wcap=60)
    for tun in ({"tile_shape": tile_shape}, {"tile_shape": tile_shape, "rounds_per_relabel": 3, "local_iters": 1}):
        _check(oracle, kernel, shape, un, b, cap_bytes=4, tuning=tun)


def test_autotune_keeps_results(oracle):
    from latticeflow_amd import workloads as W
    w = W.seg8(seed=9, shape=(300, 400))
    flow, smin, smax, lf = _solve_gpu(w["kernel"], w["shape"], w["unary"], w["pairwise"], cap_bytes=4)
    best = lf.autotune(reps=1)
    assert 1 <= best["tile_shape"] <= 3
    lf.run()
    assert lf.flow() == flow
    assert np.array_equal(lf.getCut().reshape(-1), smin)


@pytest.mark.parametrize("relabel_launches", [1, 2, 4, 16, 64])
def test_relabel_reach_keeps_results(oracle, relabel_launches):
    """However far the relabels between push rounds reach, the solve stops only
    on an exact relabel: flow and both canonical cuts stay bit-exact."""
    from latticeflow_amd import workloads as W
    w = W.seg8(seed=11, shape=(257, 381))
    for tun in ({"relabel_launches": relabel_launches},
                {"relabel_launches": relabel_launches, "tile_shape": 2, "rounds_per_relabel": 2}):
        _check(oracle, w["kernel"], w["shape"], w["unary"], w["pairwise"], cap_bytes=4, tuning=tun)


def test_repeated_solves_reuse_state(oracle):
    """Solves on one handle reuse the device state (mailboxes left empty by the
    previous solve, arc masks kept current): every repeat, across tile shapes,
    matches the oracle."""
    from latticeflow_amd import LatticeFlow, workloads as W
    w = W.blobs26(seed=13, shape=(24, 30, 36))
    lf = LatticeFlow(w["kernel"], "LatticeEnergyMinimizer", w["shape"], numerical_precision=1.0, cap_bytes=8)
    lf.addEnergyInt(w["unary"], w["pairwise"])
    offs = lf.directions()
    red, alpha = oracle.reparam(w["shape"], offs, w["unary"], w["pairwise"])
    ref = oracle.prflow(w["shape"], offs, red, alpha)
    for ts in (1, 3, 2, 1):
        lf.setTuning(tile_shape=ts)
        for _ in range(2):
            lf.run()
            assert lf.flow() == ref["flow"]
            assert np.array_equal(lf.getCut("smin").reshape(-1), ref["smin"].astype(np.int32))


def test_graph_replays_match_plain_launches(oracle):
    """Push chunks replayed from captured hipGraphs give the results of plain
    launches on every replay, with the solver's own eager work (relabels,
    delivery, flood) between replays."""
    import os
    from latticeflow_amd import LatticeFlow, workloads as W
    w = W.seg8(seed=17, shape=(333, 270))
    handles = {}
    for mode in ("graphs", "plain"):
        old = os.environ.get("LF_NO_GRAPHS")
        if mode == "plain":
            os.environ["LF_NO_GRAPHS"] = "1"
        try:
            lf = LatticeFlow(w["kernel"], "LatticeEnergyMinimizer", w["shape"], numerical_precision=1.0, cap_bytes=4)
        finally:
            if mode == "plain":
                if old is None:
                    os.environ.pop("LF_NO_GRAPHS", None)
                else:
                    os.environ["LF_NO_GRAPHS"] = old
        lf.setTuning(rounds_per_relabel=6, local_iters=2, tile_shape=3)
        lf.addEnergyInt(w["unary"], w["pairwise"])
        handles[mode] = lf
    offs = handles["plain"].directions()
    red, alpha = oracle.reparam(w["shape"], offs, w["unary"], w["pairwise"])
    ref = oracle.prflow(w["shape"], offs, red, alpha)
    for rep in range(4):
        for mode, lf in handles.items():
            lf.run()
            assert lf.flow() == ref["flow"], (mode, rep)
            assert np.array_equal(lf.getCut("smin").reshape(-1), ref["smin"].astype(np.int32)), (mode, rep)
            assert np.array_equal(lf.getCut("smax").reshape(-1), ref["smax"].astype(np.int32)), (mode, rep)


@pytest.mark.parametrize("kernel,shape", [("Full2d_4", (200, 131)), ("Full2d_8", (130, 257)), ("Full2d_8", (64, 64)),
                                          ("Full2d_8", (1, 300)), ("Full1d_2", (500,))])
def test_wave_tile_exact_passes(oracle, kernel, shape):
    """The wave-tile exact BFS and S_min flood over arc bit planes (selected
    with LF_WAVE_BFS=1 when the handle is created) give the bit-exact flow and
    cuts, on lattices with partial tiles and single rows."""
    import os
    from latticeflow_amd import LatticeFlow
    rng = np.random.default_rng(len(shape) * 1000 + shape[0])
    offs, _ = oracle.edge_order(kernel, shape)
    un, b = _random_energies(rng, shape, offs, cap=200, wcap=60)
    old = os.environ.get("LF_WAVE_BFS")
    os.environ["LF_WAVE_BFS"] = "1"
    try:
        lf = LatticeFlow(kernel, "LatticeEnergyMinimizer", shape, numerical_precision=1.0, cap_bytes=4)
    finally:
        if old is None:
            os.environ.pop("LF_WAVE_BFS", None)
        else:
            os.environ["LF_WAVE_BFS"] = old
    lf.addEnergyInt(un, b)
    red, alpha = oracle.reparam(shape, lf.directions(), un, b)
    ref = oracle.prflow(shape, lf.directions(), red, alpha)
    for tun in ({}, {"relabel_launches": 16, "tile_shape": 2}):
        lf.setTuning(**tun)
        lf.run()
        assert lf.flow() == ref["flow"], (kernel, shape, tun)
        assert np.array_equal(lf.getCut("smin").reshape(-1), ref["smin"].astype(np.int32))
        assert np.array_equal(lf.getCut("smax").reshape(-1), ref["smax"].astype(np.int32))


@pytest.mark.parametrize("kernel,nd", [("Full1d_4", 1), ("Full1d_6", 1), ("Full1d_8", 1), ("Full2d_20", 2),
                                       ("Full2d_24", 2), ("Star2d_24", 2), ("Full2d_28", 2), ("Star2d_32", 2),
                                       ("Full2d_36", 2), ("Full2d_48", 2), ("Full3d_32", 3)])
@pytest.mark.parametrize("cap_bytes", [4, 8])
def test_wide_neighbourhoods(oracle, kernel, nd, cap_bytes):
    """Radius 2-4 kernels of the reference family (code_setup.py:71-155), up
    to 48 directions, on the device: flow and both canonical cuts bit-exact,
    every tile shape."""
    rng = np.random.default_rng(hash((kernel, cap_bytes, 7)) & 0xffff)
    for trial in range(3):
        if nd == 1:
            shape = (int(rng.integers(5, 400)),)
        elif nd == 2:
            shape = tuple(int(x) for x in rng.integers(5, 70, 2))
        else:
            shape = tuple(int(x) for x in rng.integers(3, 18, 3))
        offs, _ = oracle.edge_order(kernel, shape)
        un, b = _random_energies(rng, shape, offs, cap=60, wcap=12)
        print("case", kernel, cap_bytes, shape, flush=True)
        _check(oracle, kernel, shape, un, b, cap_bytes, tuning={"tile_shape": 1 + trial % 3})


def test_bfs_list_overflow(oracle):
    """A sink on every (even, even) node of an 8-connected lattice: the first
    exact-BFS step of a 64x64 tile lists three quarters of its nodes, more than
    a list holds (half a tile), so that step walks the listed-node bits; flow
    and both canonical cuts stay bit-exact."""
    shape = (150, 203)
    offs, _ = oracle.edge_order("Full2d_8", shape)
    rng = np.random.default_rng(23)
    n = shape[0] * shape[1]
    yy, xx = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    sink = ((yy % 2 == 0) & (xx % 2 == 0)).reshape(-1)
    kp = len(offs) // 2
    b = np.zeros((n, kp, 4), np.int64)
    b[:, :, 1] = rng.integers(0, 20, (n, kp))
    b[:, :, 2] = rng.integers(0, 20, (n, kp))
    un = np.zeros((n, 2), np.int64)
    un[sink, 1] = 1000 + rng.integers(0, 50, int(sink.sum()))
    un[~sink, 0] = 300 + rng.integers(0, 50, int((~sink).sum()))
    red, _ = oracle.reparam(shape, offs, un, b)
    # sink-capable nodes (reduced unary > 0) exactly on the pattern
    assert np.array_equal(red > 0, sink) and np.all(red[~sink] < 0)
    for tun in ({}, {"relabel_launches": 1, "tile_shape": 2}):
        _check(oracle, "Full2d_8", shape, un, b, cap_bytes=4, tuning=tun)


@pytest.mark.parametrize("kernel", ["Full4d_8", "Star4d_8"])
@pytest.mark.parametrize("cap_bytes", [4, 8])
def test_4d_lattices(oracle, kernel, cap_bytes):
    """4-D kernels (install/code_setup.py:141-153, the reference's --4d
    build) on the node-parallel device engine: flow and both canonical cuts
    bit-exact, including degenerate extents of 1."""
    rng = np.random.default_rng(hash((kernel, cap_bytes, 4)) & 0xffff)
    for trial in range(5):
        shape = tuple(int(x) for x in rng.integers(1, 9, 4))
        if trial == 0:
            shape = (1, 5, 1, 6)
        offs, _ = oracle.edge_order(kernel, shape)
        un, b = _random_energies(rng, shape, offs, cap=40, wcap=15)
        print("case", kernel, cap_bytes, shape, flush=True)
        _check(oracle, kernel, shape, un, b, cap_bytes)


@pytest.mark.parametrize("kernel,nd", [("Full2d_8", 2), ("Full2d_4", 2), ("Star3d_26", 3), ("Full3d_6", 3),
                                       ("Star2d_16", 2)])
@pytest.mark.parametrize("cap_bytes", [4, 8])
def test_debug_checks_hold(oracle, kernel, nd, cap_bytes):
    """Every push chunk and exact relabel of the solve passes the state checks
    (capacities, arc pairs with the flow in flight, conservation, masks, exact
    labels), and the results are the oracle's."""
    from latticeflow_amd import LatticeFlow
    rng = np.random.default_rng(hash((kernel, cap_bytes, "dbg")) & 0xffff)
    for trial in range(3):
        shape = tuple(int(x) for x in rng.integers(8, 70 if nd == 2 else 14, nd))
        offs, _ = oracle.edge_order(kernel, shape)
        un, b = _random_energies(rng, shape, offs, cap=60, wcap=20)
        lf = LatticeFlow(kernel, "LatticeEnergyMinimizer", shape, numerical_precision=1.0, cap_bytes=cap_bytes)
        lf.setDebugChecks(True)
        lf.setTuning(rounds_per_relabel=2 + trial, local_iters=1 + 2 * trial, tile_shape=1 + trial % 3)
        lf.addEnergyInt(un, b)
        print("case", kernel, cap_bytes, shape, flush=True)
        lf.run()
        red, alpha = oracle.reparam(shape, offs, un, b)
        ref = oracle.prflow(shape, offs, red, alpha)
        assert lf.flow() == ref["flow"]
        assert np.array_equal(lf.getCut("smin").reshape(-1), ref["smin"].astype(np.int32))
