"""The event-driven TV regularization path (TVSolver::run, tv_solver.hpp:114-149,
300-642; TVRegPathSegment, tv_regpath_node.hpp:67-1028) against the oracle's
per-lambda restatement of the path output (orc_tv_path, pinned to the FP64 ROF
dual through the path conventions by tests/test_oracle_levels.py).

The reference's own check of its path is agreement with independent
single-lambda solves on its benchmark images (scripts/test_reg_path.py:26-88);
here the path is built once at lambda_max and read at many lambdas, and every
slice must equal the exact per-lambda solution within 1e-9 of the image scale
(the fixed point quantum is 2^-40 of it)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tv_images.npz")


def _err(R, Rr, f):
    S = 2 * np.abs(f - f.mean()).max()
    return float(np.abs(R - Rr).max() / S)


def _check(oracle, f, lams, kernel, tol=1e-9, host_max=0):
    from latticeflow_amd import TVRegPath
    lams = np.asarray(lams, np.float64)
    lmax = float(lams.max())
    print("case", kernel, f.shape, "lambda_max", lmax, "host_max", host_max, flush=True)
    path = TVRegPath(f, lmax, kernel, host_max=host_max)
    R = path(lams)
    Rr = oracle.tv_path(f, lams, kernel)
    st = path.stats()
    err = _err(R, Rr, f)
    print("  err %.3g stats %s" % (err, st), flush=True)
    assert err <= tol, (kernel, f.shape, err)
    return st


@pytest.mark.parametrize("kernel", ["Full2d_4", "Full2d_8", "Star2d_16"])
def test_path_vs_oracle_voronoi(oracle, kernel):
    from latticeflow_amd import workloads as W
    f = W.tv_voronoi(seed=11, shape=(48, 40), cells=12)
    lams = np.concatenate([np.linspace(0, 5.0, 21), [0.013, 0.37, 1.9]])
    _check(oracle, f, lams, kernel)


@pytest.mark.parametrize("kernel", ["Full2d_4", "Full2d_8"])
def test_path_vs_oracle_noise(oracle, kernel):
    # little regularization at lambda_max: nearly every pixel starts as its own
    # level set, so the path is almost all joins
    rng = np.random.default_rng(5)
    f = rng.random((20, 26))
    lams = np.concatenate([np.geomspace(1e-3, 40.0, 30), [0.0]])
    st = _check(oracle, f, lams, kernel)
    assert st["initial_regions"] > 100 and st["joins"] > 100


def test_reference_images(oracle):
    # the images of the reference's own path check, at its lambda grid
    with np.load(GOLDEN) as z:
        imgs = {k: z[k] for k in z.files}
    lams = np.linspace(0.1, 1, 50)
    for name in ("sanity", "truffles_small"):
        _check(oracle, imgs[name], lams, "Star2d_16")


def test_host_and_device_checks_agree(oracle):
    # the device and the host cut a level set by the same rule, so where the
    # checks run does not change the path
    from latticeflow_amd import TVRegPath
    from latticeflow_amd import workloads as W
    f = W.tv_voronoi(seed=3, shape=(30, 34), cells=9)
    lams = np.linspace(0, 3.0, 13)
    outs = []
    for host_max in (-1, 40, 1 << 20):
        p = TVRegPath(f, 3.0, "Full2d_8", host_max=host_max)
        outs.append(p(lams))
        st = p.stats()
        print(host_max, st, flush=True)
        if host_max < 0:
            assert st["host_checks"] == 0 and st["device_calls"] > 0
        if host_max == 1 << 20:
            assert st["device_calls"] == 0
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    assert _err(outs[0], oracle.tv_path(f, lams, "Full2d_8"), f) <= 1e-9


def test_device_checks_vs_oracle(oracle):
    # every check on the device
    from latticeflow_amd import workloads as W
    f = W.tv_voronoi(seed=8, shape=(36, 28), cells=8)
    st = _check(oracle, f, np.linspace(0, 4.0, 17), "Star2d_16", host_max=-1)
    assert st["host_checks"] == 0


def test_splits_random_search(oracle):
    # small random images over a range of lambda_max: every path must match,
    # and the search reports how many split events it exercised
    rng = np.random.default_rng(21)
    total_splits = 0
    for trial in range(24):
        shape = tuple(int(x) for x in rng.integers(4, 14, 2))
        kernel = ["Full2d_4", "Full2d_8", "Star2d_16"][trial % 3]
        f = rng.random(shape)
        if trial % 2:
            f = np.round(f * 4) / 4  # ties in the data
        lmax = float(10 ** rng.uniform(-1.5, 1.5))
        lams = np.concatenate([np.linspace(0, lmax, 17), lmax * rng.random(6)])
        st = _check(oracle, f, lams, kernel)
        total_splits += st["splits"]
    print("splits exercised", total_splits, flush=True)


def test_calculate2dTVPathEvents_matches_layered(oracle):
    from latticeflow_amd import calculate2dTVPathEvents
    from latticeflow_amd.pylatticeflow import _tv_path
    rng = np.random.default_rng(9)
    f = rng.random((17, 23))
    lams = [4.0, 1.0, 2.0, 0.0, -1.0]
    R, st = calculate2dTVPathEvents(f, lams, return_stats=True)
    R2, _ = _tv_path(f, lams, "Star2d_16")
    assert _err(R, R2, f) <= 1e-9
    assert calculate2dTVPathEvents(f, []).shape == (0, 17, 23)


def test_errors():
    from latticeflow_amd import TVRegPath
    f = np.random.default_rng(0).random((6, 7))
    p = TVRegPath(f, 2.0, "Full2d_8")
    with pytest.raises(IndexError):
        p([3.0])  # above lambda_max
    with pytest.raises(ValueError):
        TVRegPath(f, 2.0, "Full3d_6")
    one = TVRegPath(np.array([[0.25]]), 1.0, "Full2d_4")
    assert np.allclose(one([0.0, 0.5, 1.0]), 0.25)
