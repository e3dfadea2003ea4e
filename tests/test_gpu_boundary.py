"""The drop-in surface beyond the plain host-array calls, on the device:

* ``addEnergy`` with the reference's ``numerical_precision`` scaling
  (pylatticeflow.pyx:159-162: energies times 1/precision, rounded to long);
* ``addEnergyDevice`` / ``getCutDevice`` on device-resident torch tensors,
  written and read on a torch stream that is not the library's own (the
  library waits for the producer stream and the consumer waits for the cut);
* the same for a batch handle and for ``LatticeLevelReductions``.

Every result is compared with the oracle (bit-exact flow and cuts)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ref(oracle, shape, offs, un, b):
    red, alpha = oracle.reparam(shape, offs, un, b)
    return oracle.prflow(shape, offs, red, alpha)


@pytest.mark.parametrize("precision", [1e-6, 1e-3, 0.25])
def test_numerical_precision_scaling(oracle, precision):
    from latticeflow_amd import LatticeFlow
    rng = np.random.default_rng(int(1 / precision) % 9973)
    shape = (61, 87)
    lf = LatticeFlow("Full2d_8", "LatticeEnergyMinimizer", shape, numerical_precision=precision)
    n, kp = lf.size, lf.nFillingEdges
    # floating-point energies of a few hundred quanta each, in the reference's
    # (shape + (2,)) and (shape + (k/2, 4)) forms
    un = rng.uniform(0, 300 * precision, shape + (2,))
    b = np.zeros(shape + (kp, 4))
    b[..., 1] = rng.uniform(0, 80 * precision, shape + (kp,))
    b[..., 2] = rng.uniform(0, 80 * precision, shape + (kp,))
    lf.addEnergy(un, b)
    lf.run()
    # the integers the reference's Python layer hands to addEnergyPotentials
    un_i = np.round(un * (1.0 / precision)).astype(np.int64).reshape(n, 2)
    b_i = np.round(b * (1.0 / precision)).astype(np.int64).reshape(n, kp, 4)
    ref = _ref(oracle, shape, lf.directions(), un_i, b_i)
    assert lf.flow() == ref["flow"]
    assert np.array_equal(lf.getCut().reshape(-1), ref["smin"].astype(np.int32))
    assert lf.getCut().shape == shape and lf.getCut().dtype == np.int32


def test_add_energy_twice_accumulates(oracle):
    """addEnergyPotentials adds to what is there (direct_interface.hpp:65-95)."""
    from latticeflow_amd import LatticeFlow
    rng = np.random.default_rng(5)
    shape = (40, 33)
    lf = LatticeFlow("Full2d_4", "LatticeEnergyMinimizer", shape, numerical_precision=1.0)
    n, kp = lf.size, lf.nFillingEdges
    parts = []
    for _ in range(2):
        un = rng.integers(0, 50, (n, 2))
        b = np.zeros((n, kp, 4), np.int64)
        b[:, :, 1] = rng.integers(0, 20, (n, kp))
        b[:, :, 2] = rng.integers(0, 20, (n, kp))
        lf.addEnergy(un, b)
        parts.append((un, b))
    lf.run()
    ref = _ref(oracle, shape, lf.directions(), parts[0][0] + parts[1][0], parts[0][1] + parts[1][1])
    assert lf.flow() == ref["flow"]
    assert np.array_equal(lf.getCut().reshape(-1), ref["smin"].astype(np.int32))


def test_device_energies_and_cut_on_a_torch_stream(oracle):
    import torch
    from latticeflow_amd import LatticeFlow, workloads as W
    w = W.seg8(seed=31, shape=(301, 420))
    lf = LatticeFlow(w["kernel"], "LatticeEnergyMinimizer", w["shape"], numerical_precision=1.0, cap_bytes=4)
    n = lf.size
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        # produced on the side stream: a long kernel queue ahead of the copy so
        # that a reader that does not wait would see the zeros
        du = torch.zeros((n, 2), dtype=torch.int64, device="cuda")
        db = torch.zeros(w["pairwise"].shape, dtype=torch.int64, device="cuda")
        big = torch.ones((4096, 4096), device="cuda")
        for _ in range(20):
            big = big @ big * 1e-4
        du.copy_(torch.from_numpy(np.ascontiguousarray(w["unary"]).reshape(n, 2)).cuda(non_blocking=True))
        db.copy_(torch.from_numpy(np.ascontiguousarray(w["pairwise"])).cuda(non_blocking=True))
        lf.addEnergyDevice(du.data_ptr(), db.data_ptr())
        lf.run()
        out = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        lf.getCutDevice(out.data_ptr(), "smin")
        out2 = out.clone()  # queued on the side stream after the cut
    torch.cuda.synchronize()
    ref = _ref(oracle, w["shape"], lf.directions(), w["unary"], w["pairwise"])
    assert lf.flow() == ref["flow"]
    assert np.array_equal(out2.cpu().numpy(), ref["smin"].astype(np.int32))
    # the default (null) stream as producer and consumer
    lf2 = LatticeFlow(w["kernel"], "LatticeEnergyMinimizer", w["shape"], numerical_precision=1.0, cap_bytes=4)
    lf2.addEnergyDevice(du.data_ptr(), db.data_ptr(), stream=0)
    lf2.run()
    out3 = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    lf2.getCutDevice(out3.data_ptr(), "smax", stream=0)
    assert np.array_equal(out3.cpu().numpy(), ref["smax"].astype(np.int32))


def test_batch_device_energies(oracle):
    import torch
    from latticeflow_amd import LatticeFlowBatch, workloads as W
    inst = W.batch_instances(seed=41, count=6, lo=40, hi=200)
    shapes = [s for s, _ in inst]
    bt = LatticeFlowBatch("Full2d_8", shapes, numerical_precision=1.0, cap_bytes=4)
    keep = []
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        for i, (s, seed) in enumerate(inst):
            w = W.seg8(seed=seed, shape=s)
            du = torch.from_numpy(np.ascontiguousarray(w["unary"]).reshape(-1, 2)).cuda(non_blocking=True)
            db = torch.from_numpy(np.ascontiguousarray(w["pairwise"])).cuda(non_blocking=True)
            bt.addEnergyDevice(i, du.data_ptr(), db.data_ptr())
            keep.append((w, du, db))
    bt.run()
    flows = bt.flows()
    cuts = bt.getCuts()
    for i, (w, _, _) in enumerate(keep):
        ref = _ref(oracle, w["shape"], oracle.edge_order("Full2d_8", w["shape"])[0], w["unary"], w["pairwise"])
        assert flows[i] == ref["flow"]
        assert np.array_equal(cuts[i].reshape(-1), ref["smin"].astype(np.int32))


def test_level_reductions_device_cut(oracle):
    """LatticeLevelReductions: getCutDevice = {r < 0} of the levels."""
    import torch
    from latticeflow_amd import LatticeFlow
    rng = np.random.default_rng(8)
    shape = (33, 29)
    lf = LatticeFlow("Full2d_8", "LatticeLevelReductions", shape, numerical_precision=1.0, cap_bytes=8)
    n, kp = lf.size, lf.nFillingEdges
    un = rng.integers(0, 10 ** 5, (n, 2))
    b = np.zeros((n, kp, 4), np.int64)
    b[:, :, 1] = rng.integers(0, 10 ** 4, (n, kp))
    b[:, :, 2] = b[:, :, 1]
    lf.addEnergy(un, b)
    lf.run()
    r = lf.levels().reshape(-1)
    out = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    lf.getCutDevice(out.data_ptr())
    assert np.array_equal(out.cpu().numpy(), (r < 0).astype(np.int32))
    red, alpha = oracle.reparam(shape, lf.directions(), un, b)
    ref = oracle.pf_solve(shape, lf.directions(), red, alpha)
    assert r.sum() == ref["r"].sum()
    assert np.abs(r - ref["r"]).max() <= 2


def test_run_benchmark_device_vs_bk(tmp_path, capsys):
    """The reference's image benchmark (pyx:305-435) on the device solver with
    check_result: the CPU BK solver must give exactly the same cut (2-D and the
    benchmark's 3-D volume form)."""
    from matplotlib.image import imsave
    from latticeflow_amd.pylatticeflow import _runBenchmark
    rng = np.random.default_rng(2)
    for kernel, (H, W) in (("Full2d_8", (120, 160)), ("Star2d_16", (64, 90)), ("Full3d_6", (24, 20))):
        X = rng.random((H, W, 3))
        X[: H // 3, : W // 2] = 0.0
        X[H // 2:, W // 2:] = 1.0
        path = str(tmp_path / ("img_%s.png" % kernel))
        imsave(path, X)
        print("case", kernel, H, W, flush=True)
        _runBenchmark(path, kernel, "LatticeEnergyMinimizer", check_result=True)
        out = capsys.readouterr().out
        assert "Solutions exactly match." in out, out


def test_bk_equals_device_on_segmentation(oracle):
    """BKGCEnergyMinimizer (CPU) and the device solver on the same 512x384
    segmentation lattice: same flow, S_min and S_max."""
    from latticeflow_amd import LatticeFlow, workloads as W
    w = W.seg8(seed=12, shape=(512, 384))
    out = {}
    for solver in ("LatticeEnergyMinimizer", "BKGCEnergyMinimizer"):
        lf = LatticeFlow(w["kernel"], solver, w["shape"], numerical_precision=1.0)
        lf.addEnergyInt(w["unary"], w["pairwise"])
        lf.run()
        out[solver] = (lf.flow(), lf.getCut("smin"), lf.getCut("smax"))
    a, b = out["LatticeEnergyMinimizer"], out["BKGCEnergyMinimizer"]
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
