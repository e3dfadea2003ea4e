"""The library's state checks (lf_verify_state, the checks behind
lf_set_debug_checks) on host arrays, no GPU: states built from the oracle's
reparametrised network pass, and each planted mistake is reported as the
violation it is.  The checks stand in for the reference's debug verification
(solver/network_flow/push_relabel.hpp:132-267)."""
import collections

import numpy as np
import pytest

INF = 1 << 30


def _network(oracle, kernel, shape, seed):
    rng = np.random.default_rng(seed)
    offs, _ = oracle.edge_order(kernel, shape)
    n = int(np.prod(shape))
    k = len(offs)
    un = rng.integers(0, 1000, (n, 2))
    b = np.zeros((n, k // 2, 4), np.int64)
    w = rng.integers(0, 300, (n, k // 2))
    b[:, :, 1] = w
    b[:, :, 2] = w
    red, alpha = oracle.reparam(shape, offs, un, b)
    offs = np.asarray(offs, np.int32).reshape(k, -1)
    dirs = np.zeros((k, 3), np.int32)
    dirs[:, 3 - offs.shape[1]:] = offs
    dhw = (1,) * (3 - len(shape)) + tuple(shape)
    return dhw, dirs, np.asarray(red, np.int64), np.ascontiguousarray(np.asarray(alpha, np.int64).reshape(n, k).T)


def _nbr(dhw, dirs, v, i):
    D, H, W = dhw
    z, y, x = v // (H * W), (v // W) % H, v % W
    nz, ny, nx = z + dirs[i, 0], y + dirs[i, 1], x + dirs[i, 2]
    if 0 <= nz < D and 0 <= ny < H and 0 <= nx < W:
        return (nz * H + ny) * W + nx
    return -1


def _push_some(dhw, dirs, res, ex, mbox, rng, count):
    """Move flow along random residual arcs: half applied at the receiver,
    half left in flight in the mailbox."""
    k, n = res.shape
    moved = 0
    for _ in range(count * 20):
        v, i = int(rng.integers(n)), int(rng.integers(k))
        u = _nbr(dhw, dirs, v, i)
        if u < 0 or res[i, v] <= 0:
            continue
        d = int(rng.integers(1, res[i, v] + 1))
        res[i, v] -= d
        ex[v] -= d
        if rng.random() < 0.5:
            res[(i + k // 2) % k, u] += d
            ex[u] += d
        else:
            mbox[int(rng.integers(2)), i, v] += d
        moved += 1
        if moved == count:
            break
    assert moved == count


def _masks(res):
    m = np.zeros(res.shape[1], np.uint64)
    for i in range(res.shape[0]):
        m |= (res[i] > 0).astype(np.uint64) << np.uint64(i)
    return m


def _exact_labels(dhw, dirs, res, ex):
    k, n = res.shape
    h = np.full(n, INF, np.int64)
    q = collections.deque()
    for v in np.nonzero(ex < 0)[0]:
        h[v] = 0
        q.append(int(v))
    while q:
        w = q.popleft()
        for i in range(k):
            u = _nbr(dhw, dirs, w, i)  # arc u -> w is u's direction rev(i)
            if u < 0 or h[u] != INF or res[(i + k // 2) % k, u] <= 0:
                continue
            h[u] = h[w] + 1
            q.append(u)
    return h.astype(np.int32)


@pytest.mark.parametrize("kernel,shape,dtype", [("Full2d_8", (9, 11), np.int32), ("Full2d_4", (7, 6), np.int64),
                                                ("Star3d_26", (4, 5, 3), np.int64)])
def test_consistent_states_pass(oracle, kernel, shape, dtype):
    from latticeflow_amd.pylatticeflow import verify_state
    dhw, dirs, red0, alpha0 = _network(oracle, kernel, shape, 3)
    res = alpha0.astype(dtype)
    ex = (-red0).astype(dtype)
    assert verify_state(dhw, dirs, red0, alpha0, res, ex, mask=_masks(res).astype(np.uint32)) is None
    mbox = np.zeros((2,) + res.shape, dtype)
    _push_some(dhw, dirs, res, ex, mbox, np.random.default_rng(5), 40)
    assert verify_state(dhw, dirs, red0, alpha0, res, ex, mbox=mbox, mask=_masks(res).astype(np.uint32)) is None
    # with every flow delivered, the exact labels pass too
    k = res.shape[0]
    for p in range(2):
        for i in range(k):
            for v in np.nonzero(mbox[p, i])[0]:
                u = _nbr(dhw, dirs, int(v), i)
                d = mbox[p, i, v]
                res[(i + k // 2) % k, u] += d
                ex[u] += d
                mbox[p, i, v] = 0
    h = _exact_labels(dhw, dirs, res, ex)
    assert verify_state(dhw, dirs, red0, alpha0, res, ex, mbox=mbox, h=h, labels_exact=True) is None


def test_planted_mistakes_are_reported(oracle):
    from latticeflow_amd.pylatticeflow import verify_state
    dhw, dirs, red0, alpha0 = _network(oracle, "Full2d_8", (10, 12), 7)
    rng = np.random.default_rng(1)
    res0 = alpha0.astype(np.int64)
    ex0 = -red0
    mbox0 = np.zeros((2,) + res0.shape, np.int64)
    _push_some(dhw, dirs, res0, ex0, mbox0, rng, 60)
    k, n = res0.shape
    # an interior node and a direction with a residual arc
    v = next(v for v in range(n) if all(_nbr(dhw, dirs, v, i) >= 0 for i in range(k)) and res0[1, v] > 0)

    def run(res=None, ex=None, mbox=None, mask=None, h=None, exact=False):
        r = res0 if res is None else res
        return verify_state(dhw, dirs, red0, alpha0, r, ex0 if ex is None else ex,
                            mbox=mbox0 if mbox is None else mbox, mask=mask, h=h, labels_exact=exact)

    assert run() is None
    r = res0.copy(); r[1, v] -= 1  # flow lost from one arc
    assert "arc pair capacity" in run(res=r)
    e = ex0.copy(); e[v] += 1  # excess made up
    assert "flow conservation" in run(ex=e)
    m = mbox0.copy(); m[1, 1, v] += 1  # flow in flight twice
    assert "arc pair capacity" in run(mbox=m)
    r = res0.copy(); r[1, v] = -r[1, v] - 1
    assert "negative residual" in run(res=r)
    edge = next(v for v in range(n) if _nbr(dhw, dirs, v, 0) < 0)
    r = res0.copy(); r[0, edge] = 5
    assert "leaving the lattice" in run(res=r)
    mk = _masks(res0).astype(np.uint8); mk[v] ^= 2
    assert "arc mask" in run(mask=mk)
    # labels: exact on the delivered state, one wrong label is caught
    res, ex = res0.copy(), ex0.copy()
    for p in range(2):
        for i in range(k):
            for w in np.nonzero(mbox0[p, i])[0]:
                u = _nbr(dhw, dirs, int(w), i)
                res[(i + k // 2) % k, u] += mbox0[p, i, w]
                ex[u] += mbox0[p, i, w]
    zero = np.zeros_like(mbox0)
    h = _exact_labels(dhw, dirs, res, ex)
    assert run(res=res, ex=ex, mbox=zero, h=h, exact=True) is None
    j = int(np.nonzero((h > 0) & (h < INF))[0][0])
    h2 = h.copy(); h2[j] += 1
    assert "exact distance label" in run(res=res, ex=ex, mbox=zero, h=h2, exact=True)
