"""The C++ interface headers (include/latticeflow_amd/interface.hpp and
solvers.hpp: the reference's solver/interface.hpp:15-191 and the
calculate2dTV templates of tv_solver.hpp:648-681, over the C ABI), driven by
the compiled program tests/cpp/test_interface (built by build()) the way a
reference caller uses them: getLatticeFlowInterface, the index helpers,
addEnergyPotentials, run, getCut, calculate2dTV<Kernel, int64_t>.

Names and the CPU solver run without a GPU; the device solver and TV runs are
marked gpu.  Every output is compared with the oracle."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_interface")


def _run(*args):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_interface is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_names_and_bad_name():
    out = _run("names")
    assert "solver LatticeEnergyMinimizer" in out and "solver BKGCEnergyMinimizer" in out
    assert "dim Full2d_8 2 Star3d_26 3 Full4d_8 4" in out
    assert "valid 110" in out
    assert "exception Bad Kernel or Solver Name: NoSuchSolver" in out


def _solve_case(oracle, tmp_path, solver, kernel="Full2d_8", shape=(37, 53), seed=0):
    rng = np.random.default_rng(seed)
    offs, _ = oracle.edge_order(kernel, shape)
    n, kp = shape[0] * shape[1], len(offs) // 2
    un = rng.integers(0, 60, (n, 2)).astype(np.int64)
    b = np.zeros((n, kp, 4), np.int64)
    b[:, :, 1] = rng.integers(0, 25, (n, kp))
    b[:, :, 2] = rng.integers(0, 25, (n, kp))
    un.tofile(str(tmp_path / "unary.bin"))
    b.tofile(str(tmp_path / "binary.bin"))
    _run("solve", tmp_path, kernel, solver, shape[0], shape[1])
    cut = np.fromfile(str(tmp_path / ("cut_%s.bin" % solver)), np.int32)
    meta = open(str(tmp_path / ("meta_%s.txt" % solver))).read().split("\n")[0].split()
    assert [int(x) for x in meta[:3]] == [2, n, kp]
    assert np.array_equal(np.array([int(x) for x in meta[3:]]).reshape(kp, 2), offs[:kp])
    red, alpha = oracle.reparam(shape, offs, un, b)
    ref = oracle.prflow(shape, offs, red, alpha)
    assert np.array_equal(cut, ref["smin"].astype(np.int32))


def test_cpu_solver_through_interface(oracle, tmp_path):
    _solve_case(oracle, tmp_path, "BKGCEnergyMinimizer")


@pytest.mark.gpu
def test_device_solver_through_interface(oracle, tmp_path):
    for seed, kernel, shape in ((1, "Full2d_8", (37, 53)), (2, "Full2d_4", (128, 97)), (3, "Star2d_16", (40, 33))):
        _solve_case(oracle, tmp_path, "LatticeEnergyMinimizer", kernel, shape, seed)


@pytest.mark.gpu
def test_calculate2dTV_templates(oracle, tmp_path):
    from latticeflow_amd import workloads as W
    f = W.tv_voronoi(seed=3, shape=(30, 26), cells=9)
    f.astype(np.float64).tofile(str(tmp_path / "f.bin"))
    lams = [0.7, 0.05, 2.0, -1.0]
    for kernel in ("Full2d_8", "Star2d_16"):
        _run("tv", tmp_path, kernel, 30, 26, *lams)
        R = np.fromfile(str(tmp_path / "tv.bin"), np.float64).reshape(30, 26)
        Rr, _ = oracle.tv_single(f, lams[0], kernel)
        assert np.abs(R - Rr).max() <= 1e-9 * np.abs(Rr).max()
        P = np.fromfile(str(tmp_path / "path.bin"), np.float64).reshape(len(lams), 30, 26)
        Pr = oracle.tv_path(f, lams, kernel)
        assert np.abs(P - Pr).max() <= 1e-9 * np.abs(Pr).max()
