"""Seeded synthetic workloads for the lattice min-cut benchmarks.

Every generator returns integer energies in the reference layout
(``interface.hpp:46-61``): ``unary[n, 2] = (E0, E1)`` and
``pairwise[n, k/2, 4] = (E00, E01, E10, E11)`` over the edges in lattice order
(``edge_order``); pairs that leave the lattice are zero and ignored.  The
capacity -> energy mapping is E0 = c_s, E1 = c_t, E01 = E10 = w, E00 = E11 = 0.

* ``grid4_256``      256x256 Full2d_4: smoothed Gaussian noise image,
                     c_s/c_t = round(1000 max(+-(I - 0.5), 0)), w ~ U{0..200}.
* ``seg8``           HxW Full2d_8 contrast-sensitive segmentation: 64 random
                     rectangles with intensities U[0,1] plus N(0, 0.05) noise,
                     w = round(500 exp(-(Ip-Iq)^2 / (2 0.1^2)) / |e|),
                     E0 = round(1000 |I - 0.3|), E1 = round(1000 |I - 0.7|).
* ``blobs26``        DxHxW 26-connected volume: 4-octave value noise
                     thresholded into blobs plus N(0, 0.1),
                     w = round(300 exp(-diff^2 / 0.02) / |e|),
                     c_s = round(800 max(v - 0.5, 0)), c_t = round(800 max(0.5 - v, 0)).
* ``noise26``        the same volume without the threshold (soft boundaries).
* ``batch_sides``    side lengths of the batched workload (log-uniform 64..1024,
                     aspect ratio in [0.5, 2]).
* ``batch_instances`` (shape, seed) of every lattice of the batched workload;
                     each lattice is ``seg8`` with its own seed.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import ndimage


def edge_order(kernel, shape):
    """Lattice-order directions of a kernel (reference KernelLattice order), from the library."""
    from ._lib import lib, check
    nd = len(shape)
    k = lib().lf_kernel_size(kernel.encode())
    check(k)
    # the order depends only on the padded strides; compute it like the library does
    offs = np.empty((k, nd), np.int32)
    w = np.empty(k, np.float64)
    check(lib().lf_kernel_offsets(kernel.encode(), offs.ctypes.data, w.ctypes.data))
    edges = [tuple(int(x) for x in r) for r in offs]
    eb = [max(abs(e[d]) for e in edges) for d in range(nd)]
    pd = [shape[0] + 4 * eb[0]] + [shape[d] + eb[d] for d in range(1, nd)]
    st = [1] * nd
    for d in range(nd - 2, -1, -1):
        st[d] = st[d + 1] * pd[d + 1]
    jumps = [sum(e[d] * st[d] for d in range(nd)) for e in edges]
    idx = sorted(range(k), key=lambda i: -jumps[i])
    idx[: k // 2] = idx[: k // 2][::-1]
    return np.array([edges[i] for i in idx], np.int32), np.array([w[i] for i in idx])


def _pair_fields(img, offs):
    """For each positive direction e: neighbour values (nan outside) of img."""
    out = []
    for e in offs[: len(offs) // 2]:
        sh = np.full(img.shape, np.nan)
        src = tuple(slice(max(0, int(d)), img.shape[a] + min(0, int(d))) for a, d in enumerate(e))
        dst = tuple(slice(max(0, -int(d)), img.shape[a] + min(0, -int(d))) for a, d in enumerate(e))
        sh[dst] = img[src]
        out.append(sh)
    return out


def _energies(unary_cs, unary_ct, w_list):
    n = unary_cs.size
    kp = len(w_list)
    unary = np.empty((n, 2), np.int64)
    unary[:, 0] = unary_cs.reshape(-1)
    unary[:, 1] = unary_ct.reshape(-1)
    pw = np.zeros((n, kp, 4), np.int64)
    for e, w in enumerate(w_list):
        w = w.reshape(-1)
        pw[:, e, 1] = w
        pw[:, e, 2] = w
    return unary, pw


def grid4_256(seed=1, shape=(256, 256), kernel="Full2d_4"):
    rng = np.random.default_rng(seed)
    offs, _ = edge_order(kernel, shape)
    I = ndimage.gaussian_filter(rng.standard_normal(shape), 3.0)
    I = (I - I.min()) / max(I.max() - I.min(), 1e-300)
    cs = np.round(1000 * np.maximum(I - 0.5, 0)).astype(np.int64)
    ct = np.round(1000 * np.maximum(0.5 - I, 0)).astype(np.int64)
    w = [rng.integers(0, 201, size=shape).astype(np.int64) for _ in range(len(offs) // 2)]
    unary, pw = _energies(cs, ct, w)
    return dict(kernel=kernel, shape=tuple(shape), unary=unary, pairwise=pw, seed=seed)


def seg_image(seed, shape, n_rect=64):
    rng = np.random.default_rng(seed)
    H, W = shape
    I = np.zeros(shape)
    for _ in range(n_rect):
        y0, y1 = np.sort(rng.integers(0, H + 1, 2))
        x0, x1 = np.sort(rng.integers(0, W + 1, 2))
        I[y0:y1, x0:x1] = rng.uniform(0, 1)
    I += rng.normal(0, 0.05, shape)
    return I


def seg8(seed=2, shape=(2048, 2048), kernel="Full2d_8"):
    offs, _ = edge_order(kernel, shape)
    I = seg_image(seed, shape)
    e0 = np.round(1000 * np.abs(I - 0.3)).astype(np.int64)
    e1 = np.round(1000 * np.abs(I - 0.7)).astype(np.int64)
    w = []
    for e, nb in zip(offs[: len(offs) // 2], _pair_fields(I, offs)):
        dist = math.sqrt(float(np.sum(e.astype(np.float64) ** 2)))
        d = np.where(np.isnan(nb), 0.0, I - np.nan_to_num(nb))
        ww = np.round(500 * np.exp(-(d ** 2) / (2 * 0.1 ** 2)) / dist)
        ww[np.isnan(nb)] = 0
        w.append(ww.astype(np.int64))
    unary, pw = _energies(e0, e1, w)
    return dict(kernel=kernel, shape=tuple(shape), unary=unary, pairwise=pw, seed=seed)


def value_noise(seed, shape, octaves=4, base=4):
    rng = np.random.default_rng(seed)
    out = np.zeros(shape)
    amp, tot = 1.0, 0.0
    for o in range(octaves):
        f = base * (2 ** o)
        grid = rng.random(tuple(f + 1 for _ in shape))
        zoom = [s / (f + 1) for s in shape]
        out += amp * ndimage.zoom(grid, zoom, order=3, mode="nearest")[tuple(slice(0, s) for s in shape)]
        tot += amp
        amp *= 0.5
    out /= tot
    return (out - out.min()) / max(out.max() - out.min(), 1e-300)


def blobs26(seed=3, shape=(256, 256, 256), kernel="Star3d_26"):
    return _volume26(seed, shape, kernel, threshold=True)


def noise26(seed=3, shape=(256, 256, 256), kernel="Star3d_26"):
    """The volume of ``blobs26`` without the threshold into blobs: the value
    noise itself plus N(0, 0.1), so that boundaries are soft and flow travels
    (a harder 26-connected min-cut than the thresholded blobs)."""
    return _volume26(seed, shape, kernel, threshold=False)


def _volume26(seed, shape, kernel, threshold):
    rng = np.random.default_rng(seed + 1000003)
    offs, _ = edge_order(kernel, shape)
    v = value_noise(seed, shape)
    if threshold:
        v = (v > 0.5).astype(np.float64)
    v += rng.normal(0, 0.1, shape)
    cs = np.round(800 * np.maximum(v - 0.5, 0)).astype(np.int64)
    ct = np.round(800 * np.maximum(0.5 - v, 0)).astype(np.int64)
    w = []
    for e, nb in zip(offs[: len(offs) // 2], _pair_fields(v, offs)):
        dist = math.sqrt(float(np.sum(e.astype(np.float64) ** 2)))
        d = np.where(np.isnan(nb), 0.0, v - np.nan_to_num(nb))
        ww = np.round(300 * np.exp(-(d ** 2) / 0.02) / dist)
        ww[np.isnan(nb)] = 0
        w.append(ww.astype(np.int64))
    unary, pw = _energies(cs, ct, w)
    return dict(kernel=kernel, shape=tuple(shape), unary=unary, pairwise=pw, seed=seed)


def batch_sides(seed=4, count=512, lo=64, hi=1024):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        s = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        a = math.exp(rng.uniform(math.log(0.5), math.log(2.0)))
        h = int(round(s * math.sqrt(a)))
        w = int(round(s / math.sqrt(a)))
        out.append((max(2, h), max(2, w)))
    return out


def batch_instances(seed=4, count=512, lo=64, hi=1024):
    """The batched workload: ``count`` lattices with ``batch_sides`` shapes, lattice
    i generated by ``seg8(seed=10000 * seed + i, shape)`` (its own seed)."""
    return [(shape, 10000 * seed + i) for i, shape in enumerate(batch_sides(seed, count, lo, hi))]


def n_edges(kernel_offsets, shape):
    """Undirected in-bound edges of a lattice (the Medges/s numerator)."""
    tot = 0
    for e in kernel_offsets[: len(kernel_offsets) // 2]:
        m = 1
        for a, d in enumerate(e):
            m *= max(0, shape[a] - abs(int(d)))
        tot += m
    return tot


def tv_voronoi(seed=5, shape=(1024, 1024), cells=256, noise=0.1):
    """Piecewise-constant image: ``cells`` random Voronoi cells with values U[0,1],
    plus N(0, noise) (the TV regularization-path workload)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    H, W = shape
    pts = rng.random((cells, 2)) * np.array([H, W])
    vals = rng.random(cells)
    yy, xx = np.mgrid[0:H, 0:W]
    _, idx = cKDTree(pts).query(np.stack([yy.ravel() + 0.5, xx.ravel() + 0.5], axis=1))
    img = vals[idx].reshape(shape)
    return img + rng.normal(0, noise, shape)


def tv_lambdas(count=32, lo=0.01, hi=2.0):
    return np.exp(np.linspace(np.log(lo), np.log(hi), count))
