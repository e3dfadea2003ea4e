"""Learned TV predictor for the guided TV path (host side).

The net (liblatticeflow_amd.so, lf_tvnet_*) is a 3-layer 3x3 convolutional
net with 16 channels, BF16 matrix-core operands and FP32 accumulation.  It is
trained here on (image, lambda, TV solution) samples that the library's own
exact TV path produces, with the library's forward, gradient and Adam kernels.
Its predictions only seed the initial partition and thresholds of
``lf_tv_path_guided``; the result is the exact TV path for any prediction.

Units: for an image f with mean M and scale S = 2 max |f - M| (the TV
setup's, tv_solver.hpp:71-112), the net sees f_norm = (f - M) / S and the
feature log(S^2 lambda), and predicts x_norm = (x - M) / S of the TV output x
at lambda.

Training workload: 256x256 Voronoi images (16 cells, values U[0,1], plus
N(0, 0.1)), the 32 lambdas log-spaced in [0.01, 2], 63 seeded images = 2016
samples.  ``python -m latticeflow_amd.tvnet train OUT.npy`` trains on the GPU
and writes the parameters; the shipped ones are ``data/tvnet_params.npy``.
"""
from __future__ import annotations

import ctypes
import math
import os
import sys
import time

import numpy as np

from ._lib import check, lib

PARAMS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tvnet_params.npy")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the learned TV predictor runs on the GPU only")
    return torch


def normalize(f):
    """(f_norm, M, S) with the TV setup's mean and scale."""
    f = np.asarray(f, np.float64)
    M = float(f.mean())
    S = 2 * float(np.max(np.abs(f - M)))
    if S == 0:
        S = 1.0
    return (f - M) / S, M, S


def lam_feature(S, lam):
    return math.log(max(S * S * float(lam), 1e-30))


def init_params(seed=0):
    """He-style initialisation of the live weights (padded ones stay zero)."""
    C, TAPS = 16, 9
    wsz = C * TAPS * C
    layer = wsz + C
    rng = np.random.default_rng(seed)
    p = np.zeros(3 * layer, np.float32)
    w1 = np.zeros((C, TAPS, C))
    w1[:, :, :2] = rng.normal(0, math.sqrt(2.0 / (TAPS * 2)), (C, TAPS, 2))
    w2 = rng.normal(0, math.sqrt(2.0 / (TAPS * C)), (C, TAPS, C))
    w3 = np.zeros((C, TAPS, C))
    w3[0] = rng.normal(0, 0.1 * math.sqrt(1.0 / (TAPS * C)), (TAPS, C))
    p[:wsz] = w1.reshape(-1)
    p[layer: layer + wsz] = w2.reshape(-1)
    p[2 * layer: 2 * layer + wsz] = w3.reshape(-1)
    p[wsz: layer] = 0.01
    p[layer + wsz: 2 * layer] = 0.01
    return p


class TVNet(object):
    """Parameters on the device plus the library's forward / training calls."""

    def __init__(self, params=None):
        torch = _torch()
        n = lib().lf_tvnet_param_count()
        p = init_params() if params is None else np.asarray(params, np.float32).reshape(-1)
        if p.size != n:
            raise ValueError("expected %d parameters, got %d" % (n, p.size))
        self.params = torch.from_numpy(p.copy()).cuda()
        self.grad = torch.zeros(n, dtype=torch.float32, device="cuda")
        self.m = torch.zeros_like(self.grad)
        self.v = torch.zeros_like(self.grad)
        self.step = 0
        self._work = None

    @classmethod
    def load(cls, path=PARAMS_PATH):
        return cls(np.load(path, allow_pickle=False))

    def save(self, path):
        np.save(path, self.params.cpu().numpy())

    def _workspace(self, n, H, W):
        torch = _torch()
        need = lib().lf_tvnet_work_bytes(int(n), int(H), int(W))
        if self._work is None or self._work.numel() < need:
            self._work = torch.empty(need, dtype=torch.uint8, device="cuda")
        return self._work

    def _stream(self):
        return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)

    def forward(self, f, lam):
        """f: (n, H, W) float32 device tensor (normalised), lam: (n,) features."""
        torch = _torch()
        n, H, W = f.shape
        out = torch.empty_like(f)
        wk = self._workspace(n, H, W)
        check(lib().lf_tvnet_forward(ctypes.c_void_p(self.params.data_ptr()), n, H, W,
                                     ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(lam.data_ptr()),
                                     ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(wk.data_ptr()), wk.numel(),
                                     self._stream()))
        return out

    def gradients(self, f, lam, target):
        """(loss, grad) of the mean squared error; grad is the device buffer."""
        torch = _torch()
        n, H, W = f.shape
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        wk = self._workspace(n, H, W)
        check(lib().lf_tvnet_gradients(ctypes.c_void_p(self.params.data_ptr()), n, H, W,
                                       ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(lam.data_ptr()),
                                       ctypes.c_void_p(target.data_ptr()), ctypes.c_void_p(self.grad.data_ptr()),
                                       ctypes.c_void_p(loss.data_ptr()), ctypes.c_void_p(wk.data_ptr()), wk.numel(),
                                       self._stream()))
        return loss, self.grad

    def train_step(self, f, lam, target, lr):
        loss, _ = self.gradients(f, lam, target)
        self.step += 1
        check(lib().lf_tvnet_adam(ctypes.c_void_p(self.params.data_ptr()), ctypes.c_void_p(self.grad.data_ptr()),
                                  ctypes.c_void_p(self.m.data_ptr()), ctypes.c_void_p(self.v.data_ptr()),
                                  self.step, ctypes.c_float(lr), self._stream()))
        return loss

    def predict_path(self, f, lambdas, chunk=8):
        """Predicted TV outputs (image units) at every lambda: float32 device
        tensor (len(lambdas), H, W), the d_pred of lf_tv_path_guided."""
        torch = _torch()
        fn, M, S = normalize(f)
        H, W = fn.shape
        lam = [lam_feature(S, l) for l in lambdas]
        out = torch.empty((len(lam), H, W), dtype=torch.float32, device="cuda")
        fd = torch.from_numpy(fn.astype(np.float32)).cuda()
        for i in range(0, len(lam), chunk):
            k = min(chunk, len(lam) - i)
            fb = fd.expand(k, H, W).contiguous()
            lb = torch.tensor(lam[i:i + k], dtype=torch.float32, device="cuda")
            out[i:i + k] = self.forward(fb, lb) * S + M
        return out


# ---- training data from the solver ---------------------------------------------------

def path_targets(f, lambdas, kernel="Full2d_8"):
    """x_norm = (x - M) / S of the exact TV output at every lambda, from the
    library's TV path: x - M = (R - M) / (2 S lambda / lambda_max)."""
    from .pylatticeflow import _tv_path
    lam = np.asarray(lambdas, np.float64)
    R, _ = _tv_path(f, lam, kernel)
    _, M, S = normalize(f)
    sc = lam / lam.max()
    return (R - M) / (2 * S * S * sc[:, None, None])


def training_set(n_images=63, size=256, seed=1000, kernel="Full2d_8", log=None):
    """(f_norm[n], lam_feature[n], target[n]) for n_images x 32 lambdas."""
    from . import workloads as W
    lams = W.tv_lambdas()
    cells = max(4, int(round(256 * size * size / (1024 * 1024))))
    fs, ls, ts = [], [], []
    for i in range(n_images):
        f = W.tv_voronoi(seed=seed + i, shape=(size, size), cells=cells)
        fn, M, S = normalize(f)
        X = path_targets(f, lams, kernel)
        for j, lam in enumerate(lams):
            fs.append(fn.astype(np.float32))
            ls.append(lam_feature(S, lam))
            ts.append(X[j].astype(np.float32))
        if log:
            log("training images %d/%d" % (i + 1, n_images))
    return np.stack(fs), np.asarray(ls, np.float32), np.stack(ts)


def train(net, data, steps=1500, batch=16, lr=3e-3, seed=0, log=None):
    torch = _torch()
    f, lam, tgt = (torch.from_numpy(a).cuda() for a in data)
    rng = np.random.default_rng(seed)
    n = f.shape[0]
    losses = []
    for s in range(steps):
        idx = torch.from_numpy(rng.choice(n, batch, replace=False)).cuda()
        cur = lr * 0.5 * (1 + math.cos(math.pi * s / steps))
        loss = net.train_step(f[idx].contiguous(), lam[idx].contiguous(), tgt[idx].contiguous(), cur)
        if s % 100 == 0 or s == steps - 1:
            losses.append((s, float(loss.item())))
            if log:
                log("step %d loss %.3e" % losses[-1])
    return losses


def tv_path_learned(f, lambdas, kernel="Full2d_8", net=None, delta=float("inf"), max_repairs=8):
    """Exact TV path with thresholds (and, for a finite delta, the initial
    partition) from the net's predictions; same output as the plain path.
    The default delta = inf uses no partition: on the 1024^2 Voronoi path the
    trained net's partitions cost more in repairs than they save.  Returns
    (R, stats, guided_stats)."""
    from .pylatticeflow import _tv_path_guided
    net = net or TVNet.load()
    pred = net.predict_path(f, lambdas)
    return _tv_path_guided(f, lambdas, kernel, pred, delta, max_repairs)


def main(argv):
    if len(argv) >= 2 and argv[0] == "train":
        out = argv[1]
        t0 = time.time()
        log = lambda m: print("[tvnet %.0fs] %s" % (time.time() - t0, m), flush=True)  # noqa: E731
        _torch().zeros(1, device="cuda")  # device runtime up before the solver's first call
        data = training_set(log=log)
        log("samples %d" % data[0].shape[0])
        net = TVNet()
        train(net, data, log=log)
        net.save(out)
        log("saved " + out)
        return 0
    print("usage: python -m latticeflow_amd.tvnet train OUT.npy", file=sys.stderr)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
