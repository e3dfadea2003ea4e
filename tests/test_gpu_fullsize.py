"""Parity at the measured workloads' full sizes, on the final kernels.

* seg8 2048x2048 Full2d_8 int32 (the headline lattice): flow, S_min, S_max;
* blobs 256^3 Star3d_26 int64: flow, S_min, S_max; the unthresholded (harder)
  volume at 96^3 (the oracle needs ~12 minutes for it at 256^3);
* a 128-lattice sample of the 512-lattice batch (sides 64..1024, the batch's
  own seeds and shapes), solved as one batch: every lattice's flow and S_min;
* the 1024x1024 Voronoi TV path at three of its 32 lambdas (the smallest, a
  middle one, the largest): within 1e-9 of the oracle relative to max |R|,
  and the learned (guided) path equal to the plain one.

The oracle is the single-thread PRFlow / divide-and-conquer restatement; each
case prints its name before the device call."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mincut_check(oracle, w, cap_bytes):
    from latticeflow_amd import LatticeFlow
    print("case", w["kernel"], w["shape"], cap_bytes, flush=True)
    lf = LatticeFlow(w["kernel"], "LatticeEnergyMinimizer", w["shape"], numerical_precision=1.0, cap_bytes=cap_bytes)
    lf.addEnergyInt(w["unary"], w["pairwise"])
    lf.run()
    flow, smin, smax = lf.flow(), lf.getCut("smin").reshape(-1), lf.getCut("smax").reshape(-1)
    offs = lf.directions()
    del lf
    red, alpha = oracle.reparam(w["shape"], offs, w["unary"], w["pairwise"])
    ref = oracle.prflow(w["shape"], offs, red, alpha)
    print("flow", flow, "oracle", ref["flow"], "|S_min|", int(ref["smin"].sum()), flush=True)
    assert flow == ref["flow"]
    assert np.array_equal(smin, ref["smin"].astype(np.int32))
    assert np.array_equal(smax, ref["smax"].astype(np.int32))


def test_seg8_2048_full():
    from oracle import oracle as O
    from latticeflow_amd import workloads as W
    O.build()
    _mincut_check(O, W.seg8(seed=2), 4)


@pytest.mark.timeout(900)
def test_blobs26_256_full():
    from oracle import oracle as O
    from latticeflow_amd import workloads as W
    O.build()
    _mincut_check(O, W.blobs26(seed=3), 8)


@pytest.mark.timeout(600)
def test_noise26_96():
    from oracle import oracle as O
    from latticeflow_amd import workloads as W
    O.build()
    _mincut_check(O, W.noise26(seed=3, shape=(96, 96, 96)), 8)


@pytest.mark.timeout(600)
def test_batch512_sample_full():
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    from latticeflow_amd import LatticeFlowBatch, workloads as W
    O.build()
    inst = W.batch_instances()[:128]
    shapes = [s for s, _ in inst]
    print("case batch512[:128]", sum(a * b for a, b in shapes), "nodes", flush=True)
    ws = [W.seg8(seed=seed, shape=s) for s, seed in inst]
    bt = LatticeFlowBatch("Full2d_8", shapes, numerical_precision=1.0, cap_bytes=4)
    for i, w in enumerate(ws):
        bt.addEnergyInt(i, w["unary"], w["pairwise"])
    bt.run()
    flows = bt.flows()
    cuts = bt.getCutsFlat("smin")
    offsets = np.concatenate([[0], np.cumsum([a * b for a, b in shapes])])

    def ref(w):
        offs = O.edge_order("Full2d_8", w["shape"])[0]
        red, alpha = O.reparam(w["shape"], offs, w["unary"], w["pairwise"])
        r = O.prflow(w["shape"], offs, red, alpha)
        return r["flow"], r["smin"]

    with ThreadPoolExecutor(8) as ex:
        refs = list(ex.map(ref, ws))
    for i, (f, smin) in enumerate(refs):
        assert flows[i] == f, (i, shapes[i])
        assert np.array_equal(cuts[offsets[i]:offsets[i + 1]], smin.astype(np.int32)), (i, shapes[i])


@pytest.mark.timeout(600)
def test_tv_voronoi_1024_lambdas():
    from oracle import oracle as O
    from latticeflow_amd import workloads as W
    from latticeflow_amd.pylatticeflow import _tv_path
    from latticeflow_amd.tvnet import TVNet, tv_path_learned
    O.build()
    f = W.tv_voronoi()
    lams_all = W.tv_lambdas()
    lams = [lams_all[0], lams_all[15], lams_all[-1]]
    print("case tv 1024x1024", lams, flush=True)
    R, _ = _tv_path(f, lams, "Full2d_8")
    Rl, _, gs = tv_path_learned(f, lams, "Full2d_8", net=TVNet.load())
    # one oracle solve per lambda, in parallel threads (the C oracle releases
    # the interpreter lock); slice l of the path scales with lambda_l / max
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(3) as ex:
        parts = list(ex.map(lambda lm: O.tv_path(f, [lm], "Full2d_8")[0], lams))
    M = f.mean()
    Rr = np.stack([M + (p - M) * (lm / max(lams)) for p, lm in zip(parts, lams)])
    rel = np.abs(R - Rr).max() / np.abs(Rr).max()
    print("rel", rel, "learned vs plain", np.abs(Rl - R).max() / np.abs(R).max(), gs, flush=True)
    assert rel <= 1e-9
    assert np.abs(Rl - R).max() <= 1e-11 * np.abs(R).max()
