"""The reference's image benchmark (pylatticeflow.pyx:305-435) on the CPU:
its energies, built here with array operations, against a literal loop
restatement of the pyx's loops on small images (2-D and 3-D kernels), and the
whole benchmark driven through the CPU solver on a PNG written by the test
(the reference's own PNGs stay with the reference)."""
import math

import numpy as np


def _loop_energies(X, edges, mode_3d):
    H, W = X.shape[0], X.shape[1]
    n = H * W * W if mode_3d else H * W
    scale = [1.0 / sum(int(d) * int(d) for d in e) for e in edges]
    un = np.zeros((n, 2))
    b = np.zeros((n, len(edges), 4))
    for xi in range(H):
        for yi in range(W):
            for zi in range(W if mode_3d else 1):
                idx = (xi * W * W + yi * W + zi) if mode_3d else (xi * W + yi)
                if not mode_3d:
                    x1, y1, z1 = X[xi, yi, 0], X[xi, yi, 1], X[xi, yi, 2]
                else:
                    x1, y1, z1 = [(X[xi, yi, c] + X[xi, zi, c]) / 2 for c in range(3)]
                if x1 == y1 == z1 == 0:
                    un[idx, 0] = 50000
                elif x1 == y1 == z1 == 1:
                    un[idx, 1] = 50000
                for ei, e in enumerate(edges):
                    xj, yj = xi + int(e[0]), yi + int(e[1])
                    zj = zi + (int(e[2]) if mode_3d else 0)
                    if not (0 <= xj < H and 0 <= yj < W and (not mode_3d or 0 <= zj < W)):
                        continue
                    if not mode_3d:
                        x2, y2, z2 = X[xj, yj, 0], X[xj, yj, 1], X[xj, yj, 2]
                    else:
                        x2, y2, z2 = [(X[xj, yj, c] + X[xj, zj, c]) / 2 for c in range(3)]
                    d = math.exp(-((x1 - x2) ** 2 + (y1 - y2) ** 2 + (z1 - z2) ** 2)) * scale[ei]
                    b[idx, ei, 1] = d
                    b[idx, ei, 2] = d
    b[:, :, 1:2] -= b[:, :, 1].min()
    return un, b


def _image(rng, H, W):
    X = rng.random((H, W, 3))
    X[: H // 3, : W // 2] = 0.0
    X[H // 2:, W // 2:] = 1.0
    return X


def test_benchmark_energies_match_loops(oracle):
    from latticeflow_amd.pylatticeflow import _benchmark_energies
    rng = np.random.default_rng(0)
    for kernel, mode_3d in (("Full2d_8", False), ("Star2d_16", False), ("Full3d_6", True), ("Star3d_26", True)):
        X = _image(rng, 7, 6)
        shape = (7, 6, 6) if mode_3d else (7, 6)
        offs, _ = oracle.edge_order(kernel, shape)
        edges = offs[: len(offs) // 2]
        sh, un, b = _benchmark_energies(X, kernel, edges)
        assert sh == shape
        un_r, b_r = _loop_energies(X, edges, mode_3d)
        assert np.array_equal(un, un_r)
        assert np.allclose(b, b_r, rtol=0, atol=1e-15)


def test_run_benchmark_png_cpu_solver(tmp_path, capsys):
    from matplotlib.image import imsave
    from latticeflow_amd.pylatticeflow import _runBenchmark
    rng = np.random.default_rng(1)
    path = str(tmp_path / "img.png")
    imsave(path, _image(rng, 30, 40))
    t = _runBenchmark(path, "Full2d_8", "BKGCEnergyMinimizer", check_result=True)
    out = capsys.readouterr().out
    assert "Final result:" in out and "Solutions exactly match." in out
    assert t >= 0
