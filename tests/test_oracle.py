"""Pins the CPU oracle (oracle/) before anything is checked against it.

The reference ships no tests or golden vectors for this path, so the oracle is
pinned by what the mathematics fixes:

* kernel family / geocut weights / edge order: closed forms and hand-derived
  values (install/code_setup.py:161-263, kernellattice.hpp:61-96);
* reparametrisation: the identity 2 E(x) = K0 + sum red x + sum alpha |xi - xj|
  (network_flow_node.hpp:419-485) checked on random labellings;
* flow value and cuts: exhaustive search on tiny lattices (S_min = AND and
  S_max = OR of all minimisers) and an independent Dinic max-flow;
* the reference's own cut (the flipped set F) must lie between S_min and S_max
  and have minimum energy (push_relabel.hpp:394-402, 677-711).

A planted mistake in the oracle (test_planted_mistakes_are_caught) must make
one of these checks fail.
"""
import math

import numpy as np
import pytest


def _rand_energy(rng, shape, offs, cap=20, wcap=10):
    n = int(np.prod(shape))
    kp = len(offs) // 2
    un = rng.integers(0, cap, (n, 2))
    b = np.zeros((n, kp, 4), np.int64)
    b[:, :, 0] = rng.integers(0, 3, (n, kp))
    b[:, :, 3] = rng.integers(0, 3, (n, kp))
    b[:, :, 1] = rng.integers(0, wcap, (n, kp)) + b[:, :, 0]
    b[:, :, 2] = rng.integers(0, wcap, (n, kp)) + b[:, :, 3]
    return un, b


def test_kernel_family_sizes(oracle):
    fam = oracle.kernel_family()
    sizes = {k: len(v[1]) for k, v in fam.items()}
    # hand-enumerated: |x|^2 <= r^2 in IEEE doubles, sqrt(13)^2 < 13 and sqrt(3)^2 < 3
    assert sizes["Full2d_4"] == 4 and sizes["Full2d_8"] == 8 and sizes["Full2d_12"] == 12
    assert set(k for k in fam if k.startswith("Full2d")) == {
        "Full2d_4", "Full2d_8", "Full2d_12", "Full2d_20", "Full2d_24", "Full2d_28", "Full2d_36", "Full2d_48"}
    assert set(k for k in fam if k.startswith("Star2d")) == {
        "Star2d_4", "Star2d_8", "Star2d_16", "Star2d_24", "Star2d_32"}
    assert "Full3d_26" not in fam and "Star3d_26" in fam and sizes["Full3d_32"] == 32
    # the 26-neighbourhood is the 3x3x3 cube minus its centre
    cube = sorted(x for x in ((a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1))
                  if x != (0, 0, 0))
    assert sorted(fam["Star3d_26"][1]) == cube
    # every kernel is symmetric
    for name, (_, e) in fam.items():
        assert set(e) == set(tuple(-x for x in v) for v in e), name


def test_geocut_weights_closed_form(oracle):
    _, w4 = oracle.geocut_weights(oracle.kernel_offsets("Full2d_4"))
    assert np.allclose(w4, math.pi / 4)
    e8 = oracle.kernel_offsets("Full2d_8")
    _, w8 = oracle.geocut_weights(e8)
    for e, w in zip(e8, w8):
        L = math.hypot(*e)
        assert abs(w - (math.pi / 4) / (2 * L)) < 1e-15
    # the angular widths 2 |e| w_e cover the full turn exactly once
    for name in ("Full2d_8", "Star2d_16", "Star2d_32"):
        e = oracle.kernel_offsets(name)
        ok, w = oracle.geocut_weights(e)
        assert ok
        assert abs(sum(wi * 2 * math.hypot(*ei) for wi, ei in zip(w, e)) - 2 * math.pi) < 1e-12
    # collinear offsets (Full kernels past radius sqrt(2)) are not geocut-applicable
    assert oracle.geocut_weights(oracle.kernel_offsets("Full2d_12"))[0] is False


def test_edge_order(oracle):
    offs, _ = oracle.edge_order("Full2d_8", (7, 9))
    assert offs[:4].tolist() == [[0, 1], [1, -1], [1, 0], [1, 1]]
    assert (offs[4:] == -offs[:4]).all()
    offs, _ = oracle.edge_order("Full2d_4", (5, 5))
    assert offs.tolist() == [[0, 1], [1, 0], [0, -1], [-1, 0]]
    offs, _ = oracle.edge_order("Star3d_26", (4, 5, 6))
    k = len(offs)
    assert (offs[k // 2:] == -offs[: k // 2]).all()
    # positive half ascending in linear jump over the padded strides
    st = np.array([(5 + 1) * (6 + 1), 6 + 1, 1])
    j = offs[: k // 2] @ st
    assert (np.diff(j) > 0).all() and j[0] > 0


def test_reparam_identity(oracle):
    rng = np.random.default_rng(11)
    for kernel, shape in [("Full2d_4", (3, 4)), ("Full2d_8", (4, 3)), ("Star3d_26", (2, 3, 2))]:
        offs, _ = oracle.edge_order(kernel, shape)
        un, b = _rand_energy(rng, shape, offs)
        red, alpha = oracle.reparam(shape, offs, un, b)
        n = int(np.prod(shape))
        pairs = oracle._pairs(shape, offs)
        k0 = 2 * int(un[:, 0].sum()) + sum(2 * int(b[v, e, 0].sum()) for e, v, w in pairs)
        for _ in range(30):
            x = rng.integers(0, 2, n)
            rhs = k0 + int((red * x).sum())
            for e, v, w in pairs:
                rhs += int((alpha[v, e] * np.abs(x[v] - x[w])).sum())
            assert 2 * oracle.energy(shape, offs, un, b, x) == rhs


@pytest.mark.parametrize("kernel", ["Full2d_4", "Full2d_8", "Star2d_16", "Full3d_6", "Star3d_26", "Full1d_2"])
def test_prflow_vs_bruteforce_and_dinic(oracle, kernel):
    rng = np.random.default_rng(len(kernel) * 7 + ord(kernel[-1]))
    nd = oracle.kernel_family()[kernel][0]
    for _ in range(40):
        if nd == 1:
            shape = (int(rng.integers(1, 17)),)
        elif nd == 2:
            shape = tuple(int(x) for x in rng.integers(1, 5, 2))
        else:
            shape = tuple(int(x) for x in rng.integers(1, 4, 3))
        offs, _ = oracle.edge_order(kernel, shape)
        un, b = _rand_energy(rng, shape, offs)
        red, alpha = oracle.reparam(shape, offs, un, b)
        pr = oracle.prflow(shape, offs, red, alpha)
        dn = oracle.dinic(shape, offs, red, alpha)
        assert pr["flow"] == dn["flow"]
        assert (pr["smin"] == dn["smin"]).all() and (pr["smax"] == dn["smax"]).all()
        assert ((pr["smin"] <= pr["flipped"]) & (pr["flipped"] <= pr["smax"])).all()
        if int(np.prod(shape)) <= 16:
            emin, smin, smax = oracle.brute_force(shape, offs, un, b)
            assert oracle.min_energy_from_flow(un, b, red, pr["flow"], shape, offs) == 2 * emin
            assert (smin == pr["smin"]).all() and (smax == pr["smax"]).all()
            assert oracle.energy(shape, offs, un, b, pr["flipped"]) == emin


def test_prflow_medium_vs_dinic(oracle):
    rng = np.random.default_rng(3)
    for kernel, shape in [("Full2d_4", (120, 90)), ("Full2d_8", (64, 200)), ("Star3d_26", (12, 14, 16))]:
        offs, _ = oracle.edge_order(kernel, shape)
        un, b = _rand_energy(rng, shape, offs, cap=1000, wcap=200)
        red, alpha = oracle.reparam(shape, offs, un, b)
        pr = oracle.prflow(shape, offs, red, alpha)
        dn = oracle.dinic(shape, offs, red, alpha)
        assert pr["flow"] == dn["flow"] and (pr["smin"] == dn["smin"]).all()
        assert (pr["smax"] == dn["smax"]).all()


def test_edge_cases(oracle):
    # single node, isolated zero-capacity node, all-source, all-sink
    offs, _ = oracle.edge_order("Full2d_4", (1, 1))
    for un in ([[0, 0]], [[5, 0]], [[0, 5]]):
        un = np.array(un)
        b = np.zeros((1, 2, 4), np.int64)
        red, alpha = oracle.reparam((1, 1), offs, un, b)
        pr = oracle.prflow((1, 1), offs, red, alpha)
        assert pr["flow"] == 0
        # E0 > E1 -> node prefers "on": S_min contains it only when it has excess
        assert pr["smin"][0] == (1 if un[0, 0] > un[0, 1] else 0)
        assert pr["smax"][0] == (0 if un[0, 1] > un[0, 0] else 1)
    # the reference flips an isolated zero-capacity node (push_relabel.hpp:394-402)
    assert pr["flipped"][0] in (0, 1)
    un = np.array([[0, 0]])
    red, alpha = oracle.reparam((1, 1), offs, un, np.zeros((1, 2, 4), np.int64))
    assert oracle.prflow((1, 1), offs, red, alpha)["flipped"][0] == 1


def test_nonsubmodular_rejected(oracle):
    offs, _ = oracle.edge_order("Full2d_4", (1, 2))
    un = np.zeros((2, 2), np.int64)
    b = np.zeros((2, 2, 4), np.int64)
    b[0, 0] = [5, 0, 0, 5]
    with pytest.raises(ValueError):
        oracle.reparam((1, 2), offs, un, b)


def test_planted_mistakes_are_caught(oracle):
    """A dropped term, a wrong sign or a transposed operand in the restatement must
    break one of the checks above: simulate them on the oracle's outputs."""
    rng = np.random.default_rng(5)
    shape = (3, 4)
    offs, _ = oracle.edge_order("Full2d_8", shape)
    un, b = _rand_energy(rng, shape, offs)
    red, alpha = oracle.reparam(shape, offs, un, b)
    emin, smin, smax = oracle.brute_force(shape, offs, un, b)
    # dropped q12 term in the node reductions
    kp = len(offs) // 2
    bad_red = red.copy()
    for e, v, w in oracle._pairs(shape, offs):
        q12 = b[v, e, 0] + b[v, e, 3] - b[v, e, 1] - b[v, e, 2]
        bad_red[v] -= q12
    # transposed operand: E01 and E10 swapped
    bsw = b.copy()
    bsw[:, :, [1, 2]] = bsw[:, :, [2, 1]]
    red2, alpha2 = oracle.reparam(shape, offs, un, bsw)
    caught = 0
    for r, a in ((bad_red, alpha), (red2, alpha2), (-red, alpha)):
        pr = oracle.prflow(shape, offs, r, a)
        ok = (oracle.min_energy_from_flow(un, b, red, pr["flow"], shape, offs) == 2 * emin
              and (pr["smin"] == smin).all() and (pr["smax"] == smax).all())
        caught += not ok
    assert caught == 3
    del kp
