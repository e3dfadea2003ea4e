"""Pins the CPU restatement of the learned TV predictor (oracle/tvnet.py).

The restated gradients must equal central finite differences of the restated
loss (float64, no bfloat16 rounding); the adjoint convolution must be the
transpose of the convolution (<conv(x), g> == <x, conv_t(g)>); and planted
mistakes (a transposed tap, a dropped ReLU mask) must fail those checks.
"""
import numpy as np

from oracle import tvnet as T


def _data(seed=0, n=2, H=7, W=9):
    rng = np.random.default_rng(seed)
    f = rng.uniform(-0.5, 0.5, (n, H, W))
    lam = rng.uniform(-3, 1, n)
    target = f * 0.5 + rng.normal(0, 0.05, f.shape)
    return f, lam, target


def test_adjoint_is_transpose():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 5, 6, T.C))
    g = rng.normal(size=(2, 5, 6, T.C))
    w = rng.normal(size=(T.C, T.TAPS, T.C))
    assert np.isclose(np.sum(T.conv(x, w) * g), np.sum(x * T.conv_t(g, w)), rtol=1e-12)
    # dW is the derivative of <conv(x, w), g> in w
    dw = T.conv_dw(g, x)
    w2 = w.copy()
    w2[3, 4, 5] += 1.0
    assert np.isclose(np.sum(T.conv(x, w2) * g) - np.sum(T.conv(x, w) * g), dw[3, 4, 5], rtol=1e-10)


def test_gradient_matches_finite_differences():
    f, lam, target = _data()
    p = T.init_params(seed=3)
    p[2 * T.LAYER: 2 * T.LAYER + T.WSZ] *= 20  # a non-trivial last layer
    loss, grad = T.loss_and_grad(p, f, lam, target)
    rng = np.random.default_rng(4)
    live = np.flatnonzero(p != 0)
    idx = np.concatenate([rng.choice(live, 40, replace=False), [T.WSZ, T.LAYER + T.WSZ, 2 * T.LAYER + T.WSZ]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (T.loss_and_grad(pp, f, lam, target)[0] - T.loss_and_grad(pm, f, lam, target)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, grad[i])


def test_planted_mistakes_are_caught():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1, 5, 6, T.C))
    g = rng.normal(size=(1, 5, 6, T.C))
    w = rng.normal(size=(T.C, T.TAPS, T.C))

    def conv_t_wrong(g, w):  # taps not mirrored
        out = 0
        for t in range(T.TAPS):
            dy, dx = t // 3 - 1, t % 3 - 1
            out = out + T._shift(g, dy, dx) @ w[:, t, :]
        return out

    assert not np.isclose(np.sum(T.conv(x, w) * g), np.sum(x * conv_t_wrong(g, w)), rtol=1e-6)
    # dropping the ReLU mask changes the gradient of the first layer
    f, lam, target = _data(5)
    p = T.init_params(seed=6)
    loss, grad = T.loss_and_grad(p, f, lam, target)
    pred, (x1, y1, y2, w1, w2, w3) = T.forward(p, f, lam, keep=True)
    d = pred - target
    g3 = np.zeros(y2.shape)
    g3[..., 0] = 2 * d / d.size
    g2 = T.conv_t(g3, w3)  # mask missing
    g1 = T.conv_t(g2, w2) * (y1 > 0)
    wrong = T.conv_dw(g1, x1).reshape(-1)
    assert not np.allclose(wrong, grad[:T.WSZ], rtol=1e-3, atol=1e-12)


def test_bf16_rounding():
    assert T.bf16(1.0) == 1.0
    assert T.bf16(1.0 + 2 ** -9) == 1.0              # below half an ulp (2^-7)
    assert T.bf16(1.0 + 3 * 2 ** -9) == 1.0 + 2 ** -7
    assert T.bf16(1.0 + 2 ** -8) == 1.0              # halfway: to even
    assert T.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    assert T.bf16(-2.5) == -2.5
