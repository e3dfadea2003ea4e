"""The event loop of the TV regularization path (lf_tvpath_core.cpp: closed-form
joins, level-set checks by minimum cuts, Dinkelbach split points, connected
parts, per-pixel trace) on the CPU, through tests/cpp/tvpath_cpu: the level
sets at lambda_max come from the oracle's divide and conquer on the library's
own network, every check runs on the host.  The path must equal the oracle's
per-lambda solution (orc_tv_path, pinned to the FP64 ROF dual) within 1e-9
of the image scale at every lambda, as the reference's own path check asks
(scripts/test_reg_path.py:26-88)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "tvpath_cpu")
GOLDEN = os.path.join(ROOT, "tests", "golden", "tv_images.npz")


def _run(*args):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/tvpath_cpu is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr


def cpu_path(oracle, tmp_path, f, lams, kernel):
    f = np.ascontiguousarray(f, np.float64)
    nx, ny = f.shape
    lams = np.ascontiguousarray(lams, np.float64)
    lmax = float(max(lams.max(), 0.0))
    d = str(tmp_path)
    f.tofile(os.path.join(d, "f.bin"))
    _run("net", d, kernel, nx, ny, repr(lmax))
    n = nx * ny
    red = np.fromfile(os.path.join(d, "red.bin"), np.int64)
    offs = np.fromfile(os.path.join(d, "offs.bin"), np.int32).reshape(-1, 2)
    alpha = np.fromfile(os.path.join(d, "alpha.bin"), np.int64).reshape(n, -1)
    ref = oracle.pf_solve((nx, ny), offs, red, alpha)
    ref["level"].astype(np.int64).tofile(os.path.join(d, "reg.bin"))
    ref["r"].astype(np.int64).tofile(os.path.join(d, "r.bin"))
    lams.tofile(os.path.join(d, "lams.bin"))
    _run("path", d, kernel, nx, ny, repr(lmax), lams.size)
    R = np.fromfile(os.path.join(d, "R.bin"), np.float64).reshape(lams.size, nx, ny)
    st = dict(l.split() for l in open(os.path.join(d, "stats.txt")).read().strip().split("\n"))
    return R, {k: float(v) for k, v in st.items()}


def _err(R, Rr, f):
    return float(np.abs(R - Rr).max() / (2 * np.abs(f - f.mean()).max()))


def _check(oracle, tmp_path, f, lams, kernel, tol=1e-9):
    R, st = cpu_path(oracle, tmp_path, f, lams, kernel)
    Rr = oracle.tv_path(f, lams, kernel)
    err = _err(R, Rr, f)
    print(kernel, f.shape, "err %.3g" % err, st, flush=True)
    assert err <= tol, (kernel, f.shape, err, st)
    return st


@pytest.mark.parametrize("kernel", ["Full2d_4", "Full2d_8", "Star2d_16"])
def test_voronoi(oracle, tmp_path, kernel):
    from latticeflow_amd import workloads as W
    f = W.tv_voronoi(seed=11, shape=(40, 36), cells=10)
    lams = np.concatenate([np.linspace(0, 5.0, 21), [0.013, 0.37, 1.9]])
    _check(oracle, tmp_path, f, lams, kernel)


def test_noise_many_joins(oracle, tmp_path):
    rng = np.random.default_rng(5)
    f = rng.random((18, 22))
    st = _check(oracle, tmp_path, f, np.concatenate([np.geomspace(1e-3, 40.0, 30), [0.0]]), "Full2d_8")
    assert st["initial_regions"] > 100 and st["joins"] > 100


def test_splits_occur_and_match(oracle, tmp_path):
    # random small images over a range of lambda_max; the search must exercise
    # split events, and every path must match
    rng = np.random.default_rng(21)
    total = 0
    for trial in range(16):
        shape = tuple(int(x) for x in rng.integers(4, 13, 2))
        kernel = ["Full2d_4", "Full2d_8", "Star2d_16"][trial % 3]
        f = rng.random(shape)
        if trial % 2:
            f = np.round(f * 4) / 4  # ties in the data
        lmax = float(10 ** rng.uniform(-1.5, 1.5))
        lams = np.concatenate([np.linspace(0, lmax, 17), lmax * rng.random(6)])
        st = _check(oracle, tmp_path, f, lams, kernel)
        total += st["splits"]
    assert total > 0


def test_reference_images(oracle, tmp_path):
    with np.load(GOLDEN) as z:
        imgs = {k: z[k] for k in z.files}
    lams = np.linspace(0.1, 1, 50)  # the reference check's grid (scripts/test_reg_path.py:34)
    for name in ("sanity", "truffles_small"):
        _check(oracle, tmp_path, imgs[name], lams, "Star2d_16")


def test_planted_mistake_is_caught(oracle, tmp_path):
    # the same comparison fails for a path read at the wrong point
    from latticeflow_amd import workloads as W
    f = W.tv_voronoi(seed=3, shape=(20, 24), cells=6)
    lams = np.array([0.3, 1.0, 2.0])
    R, _ = cpu_path(oracle, tmp_path, f, lams, "Full2d_8")
    wrong = oracle.tv_path(f, lams * 0.9, "Full2d_8")
    assert _err(R[:2], wrong[:2], f) > 1e-4
