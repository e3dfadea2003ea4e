"""CPU oracle for the lattice min-cut path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this module, and only as the checker (or the timed CPU baseline). The
shipped library never imports it.

Contents (reference paths relative to the hoytak/latticeflow root):

* ``kernel_offsets`` / ``geocut_weights`` / ``kernel_family`` -- restate the
  kernel generator ``install/code_setup.py:161-263`` (Full / Star kernels,
  Cauchy-Crofton geocut weights) and the radius list ``:71-155``.
* ``edge_order`` -- restates the jump sort of ``KernelLattice``
  (``solver/lattices/kernellattice.hpp:61-96``) over the padded strides of
  ``LatticeArray`` (``solver/lattices/lattice.hpp:102-111``).
* ``reparam`` / ``prflow`` / ``dinic`` / ``cuts`` -- ctypes wrappers around
  ``lf_oracle.c`` (see its header for the line-by-line citations).
* ``energy`` / ``brute_force`` -- exact int64 evaluation of the binary energy
  and exhaustive minimisation for tiny lattices.
"""
from __future__ import annotations

import ctypes
import itertools
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


# ----------------------------------------------------------------------------
# kernels (install/code_setup.py)

def _radius_intervals(dimension, max_radius):
    """install/code_setup.py:75-78."""
    vals = set(math.sqrt(sum(x * x for x in loc))
               for loc in itertools.product(*([range(0, max_radius + 1)] * dimension)))
    return sorted(d for d in vals if 0 < d <= max_radius + 1e-8)


def _full(dimension, radius):
    """Full kernel: every integer offset x with 0 < |x|^2 <= r^2, sorted (code_setup.py:199-222)."""
    h = radius
    rng = range(int(math.floor(-h)), int(math.ceil(h)) + 1)
    lr = [x for x in itertools.product(*([rng] * dimension))
          if 0 < sum(e * e for e in x) <= radius ** 2]
    return sorted(lr)


def _star(dimension, radius):
    """Star kernel: Full minus offsets collinear with a shorter one (code_setup.py:224-256)."""
    lr = set(_full(dimension, radius))
    lr_s = sorted(lr, key=lambda e: sum(u * u for u in e))
    seen = set()
    for e in lr_s:
        ang = tuple(math.atan2(u, e[0]) for u in e[1:])
        if ang in seen:
            lr.remove(e)
        else:
            seen.add(ang)
    return sorted(lr)


def geocut_weights(edges):
    """Cauchy-Crofton weights w_e = (angular width)/(2 |e|), 2D only (code_setup.py:163-197).

    Returns (applicable, weights) in the order of ``edges``.
    """
    if len(edges) == 0 or len(edges[0]) != 2 or len(edges) < 4:
        return False, [1.0] * len(edges)
    angles = [math.atan2(e[1], e[0]) % (2 * math.pi) for e in edges]
    if len(set(angles)) != len(edges):
        return False, [1.0] * len(edges)
    info = sorted(zip(angles, edges))
    angs = [a for a, _ in info]
    n = len(angs)
    widths = []
    for i in range(n):
        a1, a2 = angs[(i - 1) % n], angs[(i + 1) % n]
        if a2 < a1:
            a2 += 2 * math.pi
        widths.append((a2 - a1) / 2)
    wmap = {e: w / (2.0 * math.sqrt(e[0] * e[0] + e[1] * e[1])) for w, (_, e) in zip(widths, info)}
    return True, [wmap[e] for e in edges]


def kernel_family():
    """All kernels the reference generates by default (code_setup.py:12,71-155):
    2D Full/Star up to radius 4, 3D Full/Star up to radius 2, the 4D Full/Star
    of radius 1 that its --4d option adds (max-4d-radius 1, wscript:64-67), plus
    the 1D ones this build also registers (Full1d_2 .. Full1d_8, radius 1..4).

    Returns {name: (ndim, offsets list in generator order)}.
    """
    out = {}
    for r in range(1, 5):
        e = _full(1, r)
        out["Full1d_%d" % len(e)] = (1, e)
    for dim, maxr in ((2, 4), (3, 2), (4, 1)):
        for r in _radius_intervals(dim, maxr):
            for typ, fn in (("Full", _full), ("Star", _star)):
                e = fn(dim, r)
                name = "%s%dd_%d" % (typ, dim, len(e))
                out.setdefault(name, (dim, e))
    return out


def kernel_offsets(name):
    return kernel_family()[name][1]


def edge_order(name, shape):
    """Kernel directions in the reference's lattice order (kernellattice.hpp:61-96).

    Jumps are taken over the padded strides (lattice.hpp:102-111); directions are
    sorted by descending jump and the first half is reversed, so the first k/2
    are the positive jumps in ascending order and entry i + k/2 is -entry i.
    Returns (offsets [k][nd] int32, geocut weights [k] float64).
    """
    nd, edges = kernel_family()[name]
    shape = tuple(int(s) for s in shape)
    assert len(shape) == nd
    eb = [max(abs(e[d]) for e in edges) for d in range(nd)]
    pdims = [shape[0] + 4 * eb[0]] + [shape[d] + eb[d] for d in range(1, nd)]
    stride = [1] * nd
    for d in range(nd - 2, -1, -1):
        stride[d] = stride[d + 1] * pdims[d + 1]
    jumps = [sum(e[d] * stride[d] for d in range(nd)) for e in edges]
    _, gw = geocut_weights(edges)
    idx = sorted(range(len(edges)), key=lambda i: -jumps[i])
    k = len(edges)
    idx[: k // 2] = idx[: k // 2][::-1]
    offs = np.array([edges[i] for i in idx], dtype=np.int32).reshape(k, nd)
    w = np.array([gw[i] for i in idx], dtype=np.float64)
    return offs, w


# ----------------------------------------------------------------------------
# the C oracle

def build():
    """Compile lf_oracle.c into liblf_oracle.so (gcc, via the Makefile here)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblf_oracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        i64p = ctypes.POINTER(ctypes.c_int64)
        i32p = ctypes.POINTER(ctypes.c_int32)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        ci = ctypes.c_int
        L.orc_reparam.argtypes = [ci, i64p, ci, i32p, i64p, i64p, i64p, i64p]
        L.orc_prflow_solve.argtypes = [ci, i64p, ci, i32p, i64p, i64p, i64p, i64p, u8p, i64p]
        L.orc_dinic.argtypes = [ci, i64p, ci, i32p, i64p, i64p, i64p, i64p, i64p]
        L.orc_cuts.argtypes = [ci, i64p, ci, i32p, i64p, i64p, i64p, i64p, u8p, u8p]
        L.orc_pf_solve.argtypes = [ci, i64p, ci, i32p, i64p, i64p, i64p, i64p, i64p]
        L.orc_pf_solve_mode.argtypes = [ci, i64p, ci, i32p, i64p, i64p, i64p, i64p, i64p, ci]
        f64p = ctypes.POINTER(ctypes.c_double)
        L.orc_tv_single.argtypes = [ctypes.c_int64, ctypes.c_int64, ci, i32p, f64p, f64p,
                                    ctypes.c_double, f64p, i64p]
        L.orc_tv_path.argtypes = [ctypes.c_int64, ctypes.c_int64, ci, i32p, f64p, f64p, ctypes.c_int64,
                                  f64p, f64p]
        for f in (L.orc_reparam, L.orc_prflow_solve, L.orc_dinic, L.orc_cuts, L.orc_pf_solve, L.orc_pf_solve_mode,
                  L.orc_tv_single, L.orc_tv_path):
            f.restype = ci
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _lat(shape, offs):
    shape = np.ascontiguousarray(shape, dtype=np.int64)
    offs = np.ascontiguousarray(offs, dtype=np.int32)
    return shape, offs, int(offs.shape[0]), int(offs.shape[1])


def reparam(shape, offs, unary, binary):
    """Energy arrays (reference layout) -> (red[N], alpha[N, k]) in raw units."""
    shape, offs, k, nd = _lat(shape, offs)
    n = int(np.prod(shape))
    unary = np.ascontiguousarray(unary, dtype=np.int64).reshape(-1)
    binary = np.ascontiguousarray(binary, dtype=np.int64).reshape(-1)
    assert unary.size == 2 * n and binary.size == n * (k // 2) * 4
    red = np.zeros(n, np.int64)
    alpha = np.zeros((n, k), np.int64)
    rc = lib().orc_reparam(nd, _p(shape, ctypes.c_int64), k, _p(offs, ctypes.c_int32),
                           _p(unary, ctypes.c_int64), _p(binary, ctypes.c_int64),
                           _p(red, ctypes.c_int64), _p(alpha, ctypes.c_int64))
    if rc == -2:
        raise ValueError("non-submodular pairwise term")
    assert rc == 0
    return red, alpha


def cuts(shape, offs, red0, redF, res):
    shape, offs, k, nd = _lat(shape, offs)
    n = int(np.prod(shape))
    red0 = np.ascontiguousarray(red0, np.int64)
    redF = np.ascontiguousarray(redF, np.int64)
    res = np.ascontiguousarray(res, np.int64)
    flow = np.zeros(1, np.int64)
    smin = np.zeros(n, np.uint8)
    smax = np.zeros(n, np.uint8)
    rc = lib().orc_cuts(nd, _p(shape, ctypes.c_int64), k, _p(offs, ctypes.c_int32),
                        _p(red0, ctypes.c_int64), _p(redF, ctypes.c_int64), _p(res, ctypes.c_int64),
                        _p(flow, ctypes.c_int64), _p(smin, ctypes.c_uint8), _p(smax, ctypes.c_uint8))
    assert rc == 0
    return int(flow[0]), smin, smax


def prflow(shape, offs, red, alpha):
    """Faithful PRFlow restatement. Returns dict(flow, smin, smax, flipped, red, res, stats)."""
    shape, offs, k, nd = _lat(shape, offs)
    n = int(np.prod(shape))
    red = np.ascontiguousarray(red, np.int64)
    alpha = np.ascontiguousarray(alpha, np.int64)
    red_o = np.zeros(n, np.int64)
    res_o = np.zeros((n, k), np.int64)
    flipped = np.zeros(n, np.uint8)
    stats = np.zeros(4, np.int64)
    rc = lib().orc_prflow_solve(nd, _p(shape, ctypes.c_int64), k, _p(offs, ctypes.c_int32),
                                _p(red, ctypes.c_int64), _p(alpha, ctypes.c_int64),
                                _p(red_o, ctypes.c_int64), _p(res_o, ctypes.c_int64),
                                _p(flipped, ctypes.c_uint8), _p(stats, ctypes.c_int64))
    assert rc == 0
    flow, smin, smax = cuts(shape, offs, red, red_o, res_o)
    return dict(flow=flow, smin=smin, smax=smax, flipped=flipped, red=red_o, res=res_o,
                stats=dict(discharges=int(stats[0]), pushes=int(stats[1]),
                           gaps=int(stats[2]), flips=int(stats[3])))


def dinic(shape, offs, red, alpha):
    """Independent max-flow. Returns dict(flow, smin, smax, red, res)."""
    shape, offs, k, nd = _lat(shape, offs)
    n = int(np.prod(shape))
    red = np.ascontiguousarray(red, np.int64)
    alpha = np.ascontiguousarray(alpha, np.int64)
    red_o = np.zeros(n, np.int64)
    res_o = np.zeros((n, k), np.int64)
    flow = np.zeros(1, np.int64)
    rc = lib().orc_dinic(nd, _p(shape, ctypes.c_int64), k, _p(offs, ctypes.c_int32),
                         _p(red, ctypes.c_int64), _p(alpha, ctypes.c_int64),
                         _p(flow, ctypes.c_int64), _p(red_o, ctypes.c_int64), _p(res_o, ctypes.c_int64))
    assert rc == 0
    f2, smin, smax = cuts(shape, offs, red, red_o, res_o)
    assert f2 == int(flow[0])
    return dict(flow=int(flow[0]), smin=smin, smax=smax, red=red_o, res=res_o)


def pf_solve(shape, offs, red, alpha, smin_split=False):
    """ParametricFlowSolver restatement. Returns dict(r, level, sections, n_levels).
    smin_split: each section splits by its S_min (the device's rule) instead of
    by the flipped set the run leaves."""
    shape, offs, k, nd = _lat(shape, offs)
    n = int(np.prod(shape))
    red = np.ascontiguousarray(red, np.int64)
    alpha = np.ascontiguousarray(alpha, np.int64)
    r = np.zeros(n, np.int64)
    lev = np.zeros(n, np.int64)
    st = np.zeros(2, np.int64)
    rc = lib().orc_pf_solve_mode(nd, _p(shape, ctypes.c_int64), k, _p(offs, ctypes.c_int32),
                                 _p(red, ctypes.c_int64), _p(alpha, ctypes.c_int64),
                                 _p(r, ctypes.c_int64), _p(lev, ctypes.c_int64), _p(st, ctypes.c_int64),
                                 1 if smin_split else 0)
    assert rc == 0, rc
    return dict(r=r, level=lev, sections=int(st[0]), n_levels=int(st[1]))


def tv_single(f, lam, kernel="Full2d_8"):
    """TVSolver::runSingleLambda restatement (before the Python layer's factor lambda)."""
    f = np.ascontiguousarray(f, np.float64)
    nx, ny = f.shape
    offs, w = edge_order(kernel, (nx, ny))
    offs = np.ascontiguousarray(offs, np.int32)
    w = np.ascontiguousarray(w, np.float64)
    R = np.zeros(f.size, np.float64)
    st = np.zeros(2, np.int64)
    rc = lib().orc_tv_single(nx, ny, int(offs.shape[0]), _p(offs, ctypes.c_int32),
                             _p(w, ctypes.c_double), _p(f, ctypes.c_double), float(lam),
                             _p(R, ctypes.c_double), _p(st, ctypes.c_int64))
    assert rc == 0, rc
    return R.reshape(nx, ny), dict(sections=int(st[0]), n_levels=int(st[1]))


def tv_path(f, lambdas, kernel="Full2d_8"):
    """Regularization-path output (TVSolver::getRegularizationPath conventions)."""
    f = np.ascontiguousarray(f, np.float64)
    nx, ny = f.shape
    offs, w = edge_order(kernel, (nx, ny))
    offs = np.ascontiguousarray(offs, np.int32)
    w = np.ascontiguousarray(w, np.float64)
    lam = np.ascontiguousarray(lambdas, np.float64).reshape(-1)
    R = np.zeros((lam.size, nx, ny), np.float64)
    rc = lib().orc_tv_path(nx, ny, int(offs.shape[0]), _p(offs, ctypes.c_int32), _p(w, ctypes.c_double),
                           _p(f, ctypes.c_double), lam.size, _p(lam, ctypes.c_double), _p(R, ctypes.c_double))
    assert rc == 0, rc
    return R


def tv1d_condat(y, lam):
    """Exact 1-D TV denoising argmin 1/2 sum (x - y)^2 + lam sum |x_{i+1} - x_i|
    (L. Condat, "A direct algorithm for 1D total variation denoising", 2013)."""
    y = np.asarray(y, np.float64)
    N = len(y)
    x = np.empty(N)
    if N == 0:
        return x
    k = k0 = kmin = kplus = 0
    vmin = y[0] - lam
    vmax = y[0] + lam
    umin = lam
    umax = -lam
    while True:
        while k == N - 1:
            if umin < 0.0:
                while True:
                    x[k0] = vmin
                    k0 += 1
                    if k0 > kmin:
                        break
                k = kmin = k0
                vmin = y[k]
                umin = lam
                umax = vmin + umin - vmax
            elif umax > 0.0:
                while True:
                    x[k0] = vmax
                    k0 += 1
                    if k0 > kplus:
                        break
                k = kplus = k0
                vmax = y[k]
                umax = -lam
                umin = vmax + umax - vmin
            else:
                vmin += umin / (k - k0 + 1)
                while True:
                    x[k0] = vmin
                    k0 += 1
                    if k0 > k:
                        return x
        umin += y[k + 1] - vmin
        umax += y[k + 1] - vmax
        if umin < -lam:
            while True:
                x[k0] = vmin
                k0 += 1
                if k0 > kmin:
                    break
            k = kmin = kplus = k0
            vmin = y[k]
            vmax = y[k] + 2 * lam
            umin = lam
            umax = -lam
        elif umax > lam:
            while True:
                x[k0] = vmax
                k0 += 1
                if k0 > kplus:
                    break
            k = kmin = kplus = k0
            vmax = y[k]
            vmin = y[k] - 2 * lam
            umin = lam
            umax = -lam
        else:
            k += 1
            if umin >= lam:
                vmin += (umin - lam) / (k - k0 + 1)
                umin = lam
                kmin = k
            if umax <= -lam:
                vmax += (umax + lam) / (k - k0 + 1)
                umax = -lam
                kplus = k


def rof_dual(y, shape, offs, c):
    """argmin_x 1/2 |x - y|^2 + sum_e c_e |x_i - x_j| over the undirected lattice
    edges (first k/2 directions), via its box-constrained dual solved with
    L-BFGS-B: x = y - D^T u, |u_e| <= c_e.  For small lattices."""
    from scipy.optimize import minimize
    y = np.asarray(y, np.float64).reshape(-1)
    pairs = _pairs(shape, np.asarray(offs))
    rows_i, rows_j, cs = [], [], []
    for e, v, w in pairs:
        rows_i.append(v); rows_j.append(w); cs.append(np.full(len(v), c[e] if np.ndim(c) else c))
    I = np.concatenate(rows_i); J = np.concatenate(rows_j); C = np.concatenate(cs)
    m = len(I)

    def xof(u):
        x = y.copy()
        np.subtract.at(x, I, u)
        np.add.at(x, J, u)
        return x

    def fg(u):
        x = xof(u)
        g = -(x[I] - x[J])
        return 0.5 * float(x @ x), g
    res = minimize(fg, np.zeros(m), jac=True, method="L-BFGS-B", bounds=list(zip(-C, C)),
                   options=dict(ftol=1e-16, gtol=1e-13, maxiter=200000, maxcor=50))
    return xof(res.x)


# ----------------------------------------------------------------------------
# exact energies

def _pairs(shape, offs):
    """(node, edge, neighbour) for every in-bounds positive edge, C order."""
    shape = tuple(int(s) for s in shape)
    n = int(np.prod(shape))
    coords = np.stack(np.unravel_index(np.arange(n), shape), axis=1)
    out = []
    k = offs.shape[0]
    for e in range(k // 2):
        c2 = coords + offs[e][None, :]
        ok = np.all((c2 >= 0) & (c2 < np.array(shape)[None, :]), axis=1)
        v = np.nonzero(ok)[0]
        w = np.ravel_multi_index(tuple(c2[ok].T), shape)
        out.append((e, v, w))
    return out


def energy(shape, offs, unary, binary, labels):
    """E(x) = sum_i E_{x_i} + sum_(i,j) E_{x_i x_j} (interface.hpp:46-61), exact int."""
    offs = np.asarray(offs)
    x = np.asarray(labels).reshape(-1).astype(np.int64)
    unary = np.asarray(unary, np.int64).reshape(-1, 2)
    kp = offs.shape[0] // 2
    binary = np.asarray(binary, np.int64).reshape(-1, kp, 4)
    E = int(np.sum(np.where(x == 1, unary[:, 1], unary[:, 0])))
    for e, v, w in _pairs(shape, offs):
        idx = 2 * x[v] + x[w]
        E += int(np.sum(binary[v, e, idx]))
    return E


def min_energy_from_flow(unary, binary, red0, flow, shape, offs):
    """2*E_min = 2*sum(E0) + 2*sum(E00 over in-bounds edges) - sum(max(-red0,0)) + flow."""
    offs = np.asarray(offs)
    unary = np.asarray(unary, np.int64).reshape(-1, 2)
    kp = offs.shape[0] // 2
    binary = np.asarray(binary, np.int64).reshape(-1, kp, 4)
    k0 = 2 * int(unary[:, 0].sum())
    for e, v, w in _pairs(shape, offs):
        k0 += 2 * int(binary[v, e, 0].sum())
    nsrc = int(np.maximum(-np.asarray(red0, np.int64), 0).sum())
    return k0 - nsrc + flow  # == 2 * E_min


def brute_force(shape, offs, unary, binary):
    """Exhaustive minimisation (n <= 20). Returns (E_min, S_min = AND, S_max = OR of minimisers)."""
    n = int(np.prod(shape))
    assert n <= 20
    offs = np.asarray(offs)
    unary = np.asarray(unary, np.int64).reshape(-1, 2)
    kp = offs.shape[0] // 2
    binary = np.asarray(binary, np.int64).reshape(-1, kp, 4)
    pairs = _pairs(shape, offs)
    X = ((np.arange(1 << n)[:, None] >> np.arange(n)[None, :]) & 1).astype(np.int64)
    E = np.where(X == 1, unary[None, :, 1], unary[None, :, 0]).sum(axis=1)
    for e, v, w in pairs:
        idx = 2 * X[:, v] + X[:, w]
        E = E + np.take_along_axis(binary[v, e, :][None, :, :].repeat(X.shape[0], 0),
                                   idx[:, :, None], axis=2)[:, :, 0].sum(axis=1)
    emin = int(E.min())
    opt = X[E == emin]
    return emin, opt.min(axis=0).astype(np.uint8), opt.max(axis=0).astype(np.uint8)
