"""Test-only CPU restatement of the learned TV predictor (float64 numpy).

Used by the tests to check the device forward pass and the device gradients
of the 3-layer 3x3 convolutional net in latticeflow_amd/csrc/lf_tvnet.hip.
The net has no counterpart in the reference; this module restates the
device's definition:

* parameters: [W1 | b1 | W2 | b2 | W3 | b3], W[oc][tap][ic] (16 x 9 x 16),
  tap = (dy + 1) * 3 + (dx + 1), biases of 16; padded entries are zero;
* input channels (f, lambda feature), zero padding at the border;
* y1 = relu(conv(x1, W1) + b1), y2 = relu(conv(y1, W2) + b2),
  pred = f + conv(y2, W3)[channel 0] + b3[0];
* loss = mean((pred - target)^2).

With ``bf16=True`` every value the device stores in bfloat16 (the weights,
the input and hidden activations, the hidden gradients) is rounded the same
way (round to nearest even), so only the summation order differs.

The gradients of ``loss_and_grad`` are checked against central finite
differences in tests/test_oracle_tvnet.py.
"""
from __future__ import annotations

import numpy as np

C = 16
TAPS = 9
WSZ = C * TAPS * C
LAYER = WSZ + C
NPARAM = 3 * LAYER


def bf16(a):
    """Round to bfloat16 (nearest even) and back, as float64."""
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, np.float64)
    out = []
    for l in range(3):
        w = p[l * LAYER: l * LAYER + WSZ].reshape(C, TAPS, C)
        b = p[l * LAYER + WSZ: (l + 1) * LAYER]
        out.append((w, b))
    return out


def _shift(x, dy, dx):
    """x[b, y + dy, x + dx, :] with zeros outside (x: n, H, W, c)."""
    n, H, W, c = x.shape
    out = np.zeros_like(x)
    ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
    xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
    out[:, yd, xd, :] = x[:, ys, xs, :]
    return out


def conv(x, w):
    """out[p][oc] = sum_t sum_ic w[oc][t][ic] x[p + d_t][ic]."""
    out = 0
    for t in range(TAPS):
        dy, dx = t // 3 - 1, t % 3 - 1
        out = out + _shift(x, dy, dx) @ w[:, t, :].T
    return out


def conv_t(g, w):
    """Adjoint of conv: gin[p][ic] = sum_t sum_oc w[oc][t][ic] g[p - d_t][oc]."""
    out = 0
    for t in range(TAPS):
        dy, dx = t // 3 - 1, t % 3 - 1
        out = out + _shift(g, -dy, -dx) @ w[:, t, :]
    return out


def conv_dw(g, x):
    """dW[oc][t][ic] = sum_p g[p][oc] x[p + d_t][ic]."""
    dw = np.zeros((C, TAPS, C))
    gf = g.reshape(-1, C)
    for t in range(TAPS):
        dy, dx = t // 3 - 1, t % 3 - 1
        dw[:, t, :] = gf.T @ _shift(x, dy, dx).reshape(-1, C)
    return dw


def forward(params, f, lam, bf16_round=False, keep=False):
    rnd = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    (w1, b1), (w2, b2), (w3, b3) = unpack(params)
    w1, w2, w3 = rnd(w1), rnd(w2), rnd(w3)
    f = np.asarray(f, np.float64)
    n, H, W = f.shape
    x1 = np.zeros((n, H, W, C))
    x1[..., 0] = f
    x1[..., 1] = np.asarray(lam, np.float64)[:, None, None]
    x1 = rnd(x1)
    y1 = rnd(np.maximum(conv(x1, w1) + b1, 0))
    y2 = rnd(np.maximum(conv(y1, w2) + b2, 0))
    pred = f + conv(y2, w3)[..., 0] + b3[0]
    if keep:
        return pred, (x1, y1, y2, w1, w2, w3)
    return pred


def loss_and_grad(params, f, lam, target, bf16_round=False):
    rnd = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    pred, (x1, y1, y2, w1, w2, w3) = forward(params, f, lam, bf16_round, keep=True)
    d = pred - np.asarray(target, np.float64)
    P = d.size
    loss = float(np.mean(d * d))
    g3 = np.zeros(y2.shape)
    g3[..., 0] = 2 * d / P
    g3 = rnd(g3)
    grad = np.zeros(NPARAM)
    grad[2 * LAYER: 2 * LAYER + WSZ] = conv_dw(g3, y2).reshape(-1)
    grad[2 * LAYER + WSZ: 3 * LAYER] = g3.reshape(-1, C).sum(0)
    g2 = rnd(conv_t(g3, w3) * (y2 > 0))
    grad[LAYER: LAYER + WSZ] = conv_dw(g2, y1).reshape(-1)
    grad[LAYER + WSZ: 2 * LAYER] = g2.reshape(-1, C).sum(0)
    g1 = rnd(conv_t(g2, w2) * (y1 > 0))
    grad[:WSZ] = conv_dw(g1, x1).reshape(-1)
    grad[WSZ: LAYER] = g1.reshape(-1, C).sum(0)
    return loss, grad


def init_params(seed=0, in_ch=2):
    """He-style initialisation of the live entries; padded entries stay zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    w1 = np.zeros((C, TAPS, C))
    w1[:, :, :in_ch] = rng.normal(0, np.sqrt(2.0 / (TAPS * in_ch)), (C, TAPS, in_ch))
    w2 = rng.normal(0, np.sqrt(2.0 / (TAPS * C)), (C, TAPS, C))
    w3 = np.zeros((C, TAPS, C))
    w3[0] = rng.normal(0, 0.1 * np.sqrt(1.0 / (TAPS * C)), (TAPS, C))
    p[:WSZ] = w1.reshape(-1)
    p[LAYER: LAYER + WSZ] = w2.reshape(-1)
    p[2 * LAYER: 2 * LAYER + WSZ] = w3.reshape(-1)
    p[WSZ: LAYER] = 0.01
    p[LAYER + WSZ: 2 * LAYER] = 0.01
    return p
