"""bench.py's contract on the CPU: the multi-GPU refusal, and that the default
2048x2048 segmentation line finds its calibrated HBM traffic record (the
counter runs in profiles/traffic.json were taken at the default tuning)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_default_line_has_measured_traffic():
    import bench
    r, it, ts, cap = bench.DEFAULT_TUNING["seg8"]
    tuning = {"rounds_per_relabel": r, "local_iters": it, "max_cycles": 0, "tile_shape": ts,
              "relabel_launches": cap}
    traffic = bench.measured_traffic("seg8 2048x2048 Full2d_8", tuning)
    assert traffic is not None and traffic > 0
    # another tuning is not matched to these counters
    other = dict(tuning, tile_shape=ts % 3 + 1)
    assert bench.measured_traffic("seg8 2048x2048 Full2d_8", other) is None
