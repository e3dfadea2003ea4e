"""The learned TV predictor on the device against its CPU restatement.

Forward pass and gradients of the BF16 matrix-core kernels are compared with
oracle/tvnet.py (float64), once with the device's bfloat16 roundings restated
(only the summation order differs: tight bound) and once in plain float64
(bound set by bfloat16 precision).  A short training run must lower the loss,
and the guided TV path seeded by the net must give the plain path's output.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(seed, n=2, H=13, W=70):
    rng = np.random.default_rng(seed)
    f = rng.uniform(-0.5, 0.5, (n, H, W)).astype(np.float32)
    lam = rng.uniform(-3, 1, n).astype(np.float32)
    target = (0.5 * f + rng.normal(0, 0.05, f.shape)).astype(np.float32)
    return f, lam, target


def _params(seed):
    from oracle import tvnet as T
    p = T.init_params(seed=seed)
    p[2 * T.LAYER: 2 * T.LAYER + T.WSZ] *= 20  # a non-trivial last layer
    return p.astype(np.float32)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("seed", [0, 1])
def test_forward_matches_oracle(seed):
    import torch
    from oracle import tvnet as T
    from latticeflow_amd.tvnet import TVNet
    f, lam, _ = _batch(seed)
    p = _params(seed)
    net = TVNet(p)
    pred = net.forward(torch.from_numpy(f).cuda(), torch.from_numpy(lam).cuda()).cpu().numpy()
    ref_b = T.forward(p.astype(np.float64), f, lam, bf16_round=True)
    ref = T.forward(p.astype(np.float64), f, lam)
    # the conv part (pred - f) carries all the error
    print("fwd rel bf16-restated %.2e float64 %.2e" % (_rel(pred - f, ref_b - f), _rel(pred - f, ref - f)), flush=True)
    assert _rel(pred - f, ref_b - f) < 2e-3
    assert _rel(pred - f, ref - f) < 3e-2


@pytest.mark.parametrize("seed", [0, 1])
def test_gradients_match_oracle(seed):
    import torch
    from oracle import tvnet as T
    from latticeflow_amd.tvnet import TVNet
    f, lam, target = _batch(seed + 10)
    p = _params(seed + 10)
    net = TVNet(p)
    loss, grad = net.gradients(torch.from_numpy(f).cuda(), torch.from_numpy(lam).cuda(),
                               torch.from_numpy(target).cuda())
    grad = grad.cpu().numpy().astype(np.float64)
    lb, gb = T.loss_and_grad(p.astype(np.float64), f, lam, target, bf16_round=True)
    l64, g64 = T.loss_and_grad(p.astype(np.float64), f, lam, target)
    assert abs(float(loss.item()) - lb) <= 1e-3 * lb
    for layer in range(3):
        sl = slice(layer * T.LAYER, (layer + 1) * T.LAYER)
        rb, r64 = _rel(grad[sl], gb[sl]), _rel(grad[sl], g64[sl])
        print("layer %d grad rel bf16-restated %.2e float64 %.2e" % (layer + 1, rb, r64), flush=True)
        assert rb < 1e-2
        assert r64 < 5e-2
    # padded weights get no gradient
    pad = np.zeros(T.NPARAM, bool)
    pad[:T.WSZ] = (np.arange(T.WSZ) % T.C) >= 2
    w3 = np.zeros((T.C, T.TAPS, T.C), bool)
    w3[1:] = True
    pad[2 * T.LAYER: 2 * T.LAYER + T.WSZ] = w3.reshape(-1)
    pad[2 * T.LAYER + T.WSZ + 1: 3 * T.LAYER] = True
    assert np.all(grad[pad] == 0)


def test_training_lowers_loss():
    import torch
    from latticeflow_amd.tvnet import TVNet, train
    rng = np.random.default_rng(3)
    f = rng.uniform(-0.5, 0.5, (8, 32, 40)).astype(np.float32)
    from scipy import ndimage
    target = np.stack([ndimage.uniform_filter(x, 3) for x in f]).astype(np.float32)
    lam = np.zeros(8, np.float32)
    net = TVNet()
    tf, tl, tt = (torch.from_numpy(a).cuda() for a in (f, lam, target))
    l0 = float(net.gradients(tf, tl, tt)[0].item())
    losses = train(net, (f, lam, target), steps=150, batch=4, lr=3e-3)
    l1 = float(net.gradients(tf, tl, tt)[0].item())
    print("loss %.3e -> %.3e" % (l0, l1), losses[-1], flush=True)
    assert l1 < 0.5 * l0


def test_learned_path_equals_plain():
    from latticeflow_amd import workloads as W
    from latticeflow_amd.pylatticeflow import _tv_path
    from latticeflow_amd.tvnet import TVNet, tv_path_learned
    f = W.tv_voronoi(seed=21, shape=(64, 72), cells=8)
    lams = W.tv_lambdas(6, 0.05, 2)
    R, st = _tv_path(f, lams, "Full2d_8")
    net = TVNet()
    for delta in (0.0, 0.02, 0.2, float("inf")):
        Rl, stl, gs = tv_path_learned(f, lams, "Full2d_8", net=net, delta=delta)
        rel = np.abs(Rl - R).max() / np.abs(R).max()
        print("delta", delta, "rel", rel, gs, flush=True)
        assert rel <= 1e-11


def test_trained_net_gain():
    """What the shipped network gains over the method without it: on a held-out
    image its predictions are closer to the exact TV output than the input
    image is (the plain method's only information) at every lambda, and at
    least twice as close on average;
    the guided path with them gives the plain path's output.  Path times of
    both are printed (on this workload the learned thresholds do not shorten
    the divide and conquer)."""
    import time
    from latticeflow_amd import workloads as W
    from latticeflow_amd.pylatticeflow import _tv_path
    from latticeflow_amd.tvnet import TVNet, normalize, path_targets, tv_path_learned
    f = W.tv_voronoi(seed=99, shape=(256, 256), cells=16)
    lams = W.tv_lambdas()
    net = TVNet.load()
    X = path_targets(f, lams)
    fn, M, S = normalize(f)
    pred = (net.predict_path(f, lams).cpu().numpy() - M) / S
    rmse_net = np.sqrt(((pred - X) ** 2).mean(axis=(1, 2)))
    rmse_in = np.sqrt(((fn[None] - X) ** 2).mean(axis=(1, 2)))
    print("rmse net", np.round(rmse_net, 4).tolist(), flush=True)
    print("rmse input", np.round(rmse_in, 4).tolist(), flush=True)
    assert np.all(rmse_net < 0.75 * rmse_in)
    assert rmse_net.mean() < 0.5 * rmse_in.mean()
    t0 = time.perf_counter()
    R, _ = _tv_path(f, lams, "Full2d_8")
    tp = time.perf_counter() - t0
    t0 = time.perf_counter()
    Rl, _, gs = tv_path_learned(f, lams, "Full2d_8", net=net)
    tl = time.perf_counter() - t0
    print("plain %.3fs learned %.3fs" % (tp, tl), gs, flush=True)
    assert np.abs(Rl - R).max() <= 1e-11 * np.abs(R).max()
