"""Benchmark: exact lattice min-cut throughput on one MI355X.

A step is one exact solve (lf_run: working state from the resident network,
push/relabel rounds with exact global relabels, S_min flood, flow value) of
one synthetic lattice whose energies were ingested into HBM before timing.

Default workload: 2048x2048 Full2d_8, int32 capacities, contrast-sensitive
segmentation energy (64 random rectangles, seed 2).  Prints one JSON line with
the metric value (Medges/s), the roofline of the dominant kernel (the push /
relabel round, HBM-bound) measured with hipEvents on the solver stream, the
same for the exact global-relabel BFS and for the whole solve, and the
single-thread CPU baseline (the oracle's PRFlow restatement) timed on this host.
The line also carries the other two parts of the metric, each measured in the
same run: "parts" holds the 256^3 26-connected min-cut (Medges/s, parity on a
sample), the same volume unthresholded (a harder 3-D min-cut), the 1024^2 x
32-lambda TV path, plain and learned, and the
event-driven TV path (the whole path over lambda, read at the same 32
lambdas) on a 128^2 image of the same kind (--no-parts leaves them out).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Min-cut Medges/s (2048² 8-conn, 256³ 26-conn); TV-path time learned vs plain"
HBM_PEAK_GBS = 8000.0


# Tuning of each min-cut workload's bench line: the fastest schedule of a sweep
# on it, lf_autotune's candidates among them (rounds per relabel, local
# iterations, tile shape, relabel launches), fixed so that the line and its
# rocprofv3 profile run the same schedule; --autotune runs the search instead.
DEFAULT_TUNING = {"seg8": (4, 8, 1, 2), "blobs26": (1, 1, 1, 2), "grid4": (12, 4, 2, 2), "noise26": (0, 0, 1, 0)}


def make_workload(name, shape):
    from latticeflow_amd import workloads as W
    if name == "seg8":
        return W.seg8(seed=2, shape=shape or (2048, 2048)), 4
    if name == "grid4":
        return W.grid4_256(seed=1, shape=shape or (256, 256)), 4
    if name == "blobs26":
        return W.blobs26(seed=3, shape=shape or (256, 256, 256)), 8
    if name == "noise26":
        return W.noise26(seed=3, shape=shape or (256, 256, 256)), 8
    raise SystemExit("unknown workload " + name)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="seg8")
    ap.add_argument("--shape", default="")
    ap.add_argument("--rounds", type=int, default=0)
    ap.add_argument("--iters", type=int, default=0)
    ap.add_argument("--tile-shape", type=int, default=0)
    ap.add_argument("--relabel-launches", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true", help="batch workload: keep the default tuning")
    ap.add_argument("--autotune", action="store_true", help="single-lattice workloads: run lf_autotune first")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parts", action="store_true", help="default line: leave out the 3-D and TV-path parts")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device-to-device 5 times (calibrates the FETCH_SIZE / WRITE_SIZE "
                         "counters against a kernel of known traffic)")
    a = ap.parse_args()
    if a.copy_probe:
        return copy_probe()
    if a.workload == "tvpath":
        return bench_tvpath(a)
    if a.workload == "batch512":
        return bench_batch(a)
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (the solver has no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from latticeflow_amd import LatticeFlow, workloads as W

    shape = tuple(int(x) for x in a.shape.split("x")) if a.shape else None
    shape_given = shape is not None
    t0 = time.time()
    w, cap_bytes = make_workload(a.workload, shape)
    if a.verbose:
        print("generated %s in %.1fs" % (a.workload, time.time() - t0), file=sys.stderr, flush=True)
    shape = w["shape"]
    lf = LatticeFlow(w["kernel"], "LatticeEnergyMinimizer", shape, numerical_precision=1.0,
                     cap_bytes=cap_bytes)
    explicit = bool(a.rounds or a.iters or a.tile_shape or a.relabel_launches)
    if explicit:
        lf.setTuning(rounds_per_relabel=a.rounds, local_iters=a.iters, tile_shape=a.tile_shape,
                     relabel_launches=a.relabel_launches)
    elif not a.autotune and a.workload in DEFAULT_TUNING:
        r, it, ts, cap = DEFAULT_TUNING[a.workload]
        lf.setTuning(rounds_per_relabel=r, local_iters=it, tile_shape=ts, relabel_launches=cap)
    # inputs resident in HBM before the timed region
    du = torch.from_numpy(np.ascontiguousarray(w["unary"])).cuda()
    dp = torch.from_numpy(np.ascontiguousarray(w["pairwise"])).cuda()
    torch.cuda.synchronize()
    lf.addEnergyDevice(du.data_ptr(), dp.data_ptr())
    del du, dp
    offs = lf.directions()
    edges = W.n_edges(offs, shape)
    if a.autotune and not explicit:
        # the library's autotuner picks the tile shape and relabel schedule
        # (untimed; every candidate gives the same flow and cuts)
        lf.autotune(reps=4)
    tuning = lf.getTuning()

    for _ in range(a.warmup):
        lf.run()
    torch.cuda.synchronize()
    push_s = 0.0
    tiles = 0
    launches = 0
    bfs_s = 0.0
    bfs_visits = 0
    bfs_launches = 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        lf.run()
        st = lf.stats()
        push_s += st["push_seconds"]
        tiles += st["tiles_pushed"]
        launches += st["rounds"]
        bfs_s += st["bfs_device_seconds"]
        bfs_visits += st["bfs_visits"]
        bfs_launches += st["bfs_launches"]
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    flow = lf.flow()
    if a.verbose:
        print(json.dumps(st), file=sys.stderr, flush=True)

    bytes_alg = tiles * st["tile_nodes"] * st["bytes_per_node"]
    achieved = bytes_alg / push_s / 1e9 if push_s > 0 else 0.0
    bfs_bytes = bfs_visits * st["bfs_bytes_per_visit"]
    bfs_achieved = bfs_bytes / bfs_s / 1e9 if bfs_s > 0 else 0.0
    traffic, traffic_src = measured_traffic("%s %s %s" % (a.workload, "x".join(map(str, shape)), w["kernel"]),
                                            tuning, with_source=True)
    # the same kernel under the schedule with the least LDS work per staged
    # byte (4 local iterations per round): the fraction of the HBM roofline
    # a round reaches when its time goes mostly to staging the tile
    push4 = None
    if a.workload == "seg8" and not explicit and not a.autotune:
        lf.setTuning(rounds_per_relabel=4, local_iters=4, tile_shape=tuning["tile_shape"],
                     relabel_launches=tuning["relabel_launches"])
        lf.run()
        p_s, p_t, p_l = 0.0, 0, 0
        for _ in range(3):
            lf.run()
            s4 = lf.stats()
            p_s += s4["push_seconds"]
            p_t += s4["tiles_pushed"]
            p_l += s4["rounds"]
        lf.setTuning(**{k: tuning[k] for k in ("rounds_per_relabel", "local_iters", "tile_shape", "relabel_launches")})
        b4 = p_t * s4["tile_nodes"] * s4["bytes_per_node"]
        push4 = {"local_iters": 4, "rounds_per_relabel": 4, "avg_launch_us": round(p_s / max(p_l, 1) * 1e6, 2),
                 "nodes_per_launch": round(p_t * s4["tile_nodes"] / max(p_l, 1), 1),
                 "achieved": round(b4 / p_s / 1e9, 1) if p_s > 0 else 0.0,
                 "frac": round(b4 / p_s / 1e9 / HBM_PEAK_GBS, 4) if p_s > 0 else 0.0}
    roofline = {
        "bound": "hbm",
        "kernel": "k_push (push/relabel round)",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_source": traffic_src,
        "avg_launch_us": round(push_s / max(launches, 1) * 1e6, 2),
        "bytes_per_node": st["bytes_per_node"],
        "nodes_per_launch": round(tiles * st["tile_nodes"] / max(launches, 1), 1),
        "push_share_of_step": round(push_s / a.steps / dt, 3),
        # the exact global relabel: nodes its tile BFS visits x (k c + 4 k + 4)
        # bytes, over the device time of its launches (hipEvents per batch)
        "bfs": {"kernel": "k_bfs_tile (exact global-relabel BFS)", "achieved": round(bfs_achieved, 1),
                "frac": round(bfs_achieved / HBM_PEAK_GBS, 4), "bytes_per_visit": st["bfs_bytes_per_visit"],
                "bytes_model": "per node touched: k residuals (k c), k neighbour labels (4 k) and its own "
                               "label (4); the kernel reads 1-bit arc masks kept current by the push "
                               "write-back instead of the k residuals, so the model counts bytes it never "
                               "moves (26-direction int64 lattices exceed the HBM peak on it)",
                "visits_per_launch": round(bfs_visits / max(bfs_launches, 1), 1),
                "avg_launch_us": round(bfs_s / max(bfs_launches, 1) * 1e6, 2),
                "bfs_share_of_step": round(bfs_s / a.steps / dt, 3)},
        # push + BFS algorithmic bytes over the whole solve's wall time
        "step": {"achieved": round((bytes_alg + bfs_bytes) / a.steps / dt / 1e9, 1),
                 "frac": round((bytes_alg + bfs_bytes) / a.steps / dt / 1e9 / HBM_PEAK_GBS, 4)},
    }
    if push4:
        roofline["k_push_at_4_local_iters"] = push4

    cpu = None
    parity = None
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        O.build()
        # bounded sample: the whole lattice when it is small, else its first half
        # along the slowest axis (the oracle solve takes tens of seconds there)
        n = int(np.prod(shape))
        cshape = shape if n <= 8 * 1024 * 1024 else (shape[0] // 2,) + tuple(shape[1:])
        cn = int(np.prod(cshape))
        red, alpha = O.reparam(cshape, offs, w["unary"][:cn], w["pairwise"][:cn])
        tc = time.perf_counter()
        ref = O.prflow(cshape, offs, red, alpha)
        tcpu = time.perf_counter() - tc
        cedges = W.n_edges(offs, cshape)
        cpu = {"value": round(cedges / tcpu / 1e6, 3), "unit": "Medges/s", "cores": 1, "kind": "port",
               "sample": "%s %s lattice, one single-thread solve of the PRFlow restatement (oracle/lf_oracle.c),"
                         " %.2f s" % ("full" if cshape == shape else "first-half", "x".join(map(str, cshape)), tcpu),
               **host_cpu()}
        if cshape == shape:
            parity = bool(ref["flow"] == flow and np.array_equal(
                lf.getCut().reshape(-1), ref["smin"].astype(np.int32)))
        else:
            parity = sample_parity(w, cshape, cap_bytes, ref)

    out = {
        "metric": METRIC,
        "value": round(edges / dt / 1e6, 3),
        "unit": "Medges/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32" if cap_bytes == 4 else "int64",
        "data": "synthetic",
        "config": {"workload": "%s %s %s" % (a.workload, "x".join(map(str, shape)), w["kernel"]),
                   "edges": edges, "nodes": int(np.prod(shape)), "seed": w["seed"],
                   "flow": flow, "parity_vs_cpu": parity,
                   "cycles": st["cycles"], "rounds": st["rounds"], "bfs_launches": st["bfs_launches"],
                   "tuning": tuning,
                   "tuning_source": ("explicit" if explicit else "lf_autotune" if a.autotune else
                                     "fixed (the fastest of a schedule sweep on this workload, lf_autotune's candidates among them)")},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if a.workload == "seg8" and not shape_given and not a.no_parts:
        del lf
        out["parts"] = {"blobs26_256x256x256": part_blobs26(a), "noise26_256x256x256": part_noise26(a),
                        "tvpath_1024x1024_32": part_tvpath(a), "tvpath_events_128x128": part_tvpath_events(a)}
    print(json.dumps(out), flush=True)


def host_cpu():
    """Host CPU model and the cores this process may use."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"host_cpu": model, "nproc": avail}


def sample_parity(w, cshape, cap_bytes, ref):
    """Device solve of the same leading sample the CPU baseline solved, compared
    with the oracle's flow and S_min on it."""
    import numpy as np
    from latticeflow_amd import LatticeFlow
    cn = int(np.prod(cshape))
    lf = LatticeFlow(w["kernel"], "LatticeEnergyMinimizer", cshape, numerical_precision=1.0, cap_bytes=cap_bytes)
    lf.addEnergyInt(w["unary"][:cn], w["pairwise"][:cn])
    lf.run()
    return bool(lf.flow() == ref["flow"] and np.array_equal(lf.getCut().reshape(-1), ref["smin"].astype(np.int32)))


def part_blobs26(a):
    """The 256^3 26-connected min-cut of the metric (blobs volume, int64), timed
    like the default line, with its CPU baseline and sample parity."""
    return part_volume26(a, "blobs26", 2)


def part_noise26(a):
    """The same 256^3 volume without the threshold into blobs (soft boundaries,
    flow travels): a harder 26-connected min-cut beside the metric's; CPU
    baseline and parity on the leading 1/32 of the volume."""
    return part_volume26(a, "noise26", 32)


def part_volume26(a, name, sample_div):
    import numpy as np
    import torch
    from latticeflow_amd import LatticeFlow, workloads as W
    w, cap_bytes = make_workload(name, None)
    shape = w["shape"]
    lf = LatticeFlow(w["kernel"], "LatticeEnergyMinimizer", shape, numerical_precision=1.0, cap_bytes=cap_bytes)
    r, it, ts, cap = DEFAULT_TUNING[name]
    lf.setTuning(rounds_per_relabel=r, local_iters=it, tile_shape=ts, relabel_launches=cap)
    du = torch.from_numpy(np.ascontiguousarray(w["unary"])).cuda()
    dp = torch.from_numpy(np.ascontiguousarray(w["pairwise"])).cuda()
    torch.cuda.synchronize()
    lf.addEnergyDevice(du.data_ptr(), dp.data_ptr())
    del du, dp
    edges = W.n_edges(lf.directions(), shape)
    steps = max(3, min(a.steps, 10))
    lf.run()
    torch.cuda.synchronize()
    push_s, tiles, launches = 0.0, 0, 0
    t0 = time.perf_counter()
    for _ in range(steps):
        lf.run()
        st = lf.stats()
        push_s += st["push_seconds"]
        tiles += st["tiles_pushed"]
        launches += st["rounds"]
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    achieved = tiles * st["tile_nodes"] * st["bytes_per_node"] / push_s / 1e9 if push_s > 0 else 0.0
    res = {"workload": "%s %s %s int64" % (name, "x".join(map(str, shape)), w["kernel"]),
           "value": round(edges / dt / 1e6, 3), "unit": "Medges/s", "ms_per_step": round(dt * 1e3, 3),
           "steps": steps, "edges": edges, "flow": lf.flow(), "k_push_frac": round(achieved / HBM_PEAK_GBS, 4),
           "cycles": st["cycles"], "rounds": st["rounds"], "bfs_launches": st["bfs_launches"]}
    del lf
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        O.build()
        cshape = (shape[0] // sample_div,) + tuple(shape[1:])
        cn = int(np.prod(cshape))
        offs = O.edge_order(w["kernel"], cshape)[0]
        red, alpha = O.reparam(cshape, offs, w["unary"][:cn], w["pairwise"][:cn])
        tc = time.perf_counter()
        ref = O.prflow(cshape, offs, red, alpha)
        tcpu = time.perf_counter() - tc
        res["cpu_baseline"] = {"value": round(W.n_edges(offs, cshape) / tcpu / 1e6, 3), "unit": "Medges/s",
                               "cores": 1, "kind": "port",
                               "sample": "leading %s, one single-thread PRFlow-restatement solve, %.2f s"
                                         % ("x".join(map(str, cshape)), tcpu), **host_cpu()}
        res["parity_vs_cpu_sample"] = sample_parity(w, cshape, cap_bytes, ref)
    return res


def part_tvpath(a):
    """The TV-path part of the metric: 1024^2 Voronoi image, Full2d_8, 32
    lambdas log-spaced in [0.01, 2]; plain and learned wall time (one timed
    path each after a warm-up), their output difference, and the oracle's
    parity on two of the lambdas."""
    import argparse
    ns = argparse.Namespace(**vars(a))
    ns.steps, ns.warmup, ns.shape = 1, 1, ""
    return tvpath_record(ns)


def part_tvpath_events(a, n=128, cells=16):
    """The event-driven TV path (TVRegPath: the exact solution at lambda_max =
    2, then joins and device-checked splits down to lambda = 0) of an n x n
    Voronoi image (``cells`` cells + N(0, 0.1)), read at the 32 lambdas
    log-spaced in [0.01, 2]; timed against the per-lambda layered path on the
    same lambdas, with the oracle's parity on two of them."""
    import numpy as np
    import torch
    torch.zeros(1, device="cuda")
    from latticeflow_amd import TVRegPath, workloads as W
    from latticeflow_amd.pylatticeflow import _tv_path
    f = W.tv_voronoi(seed=5, shape=(n, n), cells=cells)
    lams = W.tv_lambdas()
    lmax = float(lams.max())
    TVRegPath(f[:32, :32], lmax, "Full2d_8")  # warm-up of the device code paths
    t0 = time.perf_counter()
    path = TVRegPath(f, lmax, "Full2d_8")
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    R = path(lams)
    t_eval = time.perf_counter() - t0
    st = path.stats()
    _tv_path(f, lams, "Full2d_8")
    t0 = time.perf_counter()
    Rl, _ = _tv_path(f, lams, "Full2d_8")
    t_layered = time.perf_counter() - t0
    S = 2 * float(np.abs(f - f.mean()).max())
    out = {"workload": "tvpath events %dx%d Full2d_8 Voronoi(%d cells)+N(0,0.1), lambda_max 2, read at 32 lambdas "
                       "log[0.01,2]" % (n, n, cells),
           "build_seconds": round(t_build, 3), "eval_seconds": round(t_eval, 4),
           "layered_seconds_same_lambdas": round(t_layered, 3),
           "max_diff_vs_layered_over_scale": float(np.abs(R - Rl).max() / S),
           "stats": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()}}
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        O.build()
        pick = [0, 31]
        tc = time.perf_counter()
        Rr = O.tv_path(f, lams[pick], "Full2d_8")
        tcpu = time.perf_counter() - tc
        err = float(np.abs(R[pick] - Rr).max() / S)
        out["parity_vs_cpu"] = {"max_err_over_scale": err, "within_1e-9": err <= 1e-9}
        out["cpu_baseline"] = {"value": round(tcpu / len(pick) * len(lams), 2),
                               "unit": "s (32 lambdas, extrapolated)", "cores": 1, "kind": "port",
                               "sample": "lambdas %s with the oracle's per-lambda divide and conquer, %.2f s for %d "
                                         "lambdas, scaled to 32" % ([round(float(lams[i]), 4) for i in pick], tcpu,
                                                                    len(pick)),
                               **host_cpu()}
    return out


def measured_traffic(workload, tuning, with_source=False):
    """HBM bytes per k_push launch from the calibrated FETCH_SIZE / WRITE_SIZE
    counter runs recorded in profiles/traffic.json, when they were taken on
    this workload at this tuning; else None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    none = (None, None) if with_source else None
    try:
        rec = json.load(open(path)).get(workload)
    except (OSError, ValueError):
        return none
    if not rec or any(tuning.get(k) != v for k, v in rec.get("tuning", {}).items()):
        return none
    # the counters were taken on this kernel's source only
    if rec.get("kernel_source_sha256") != push_kernel_digest():
        return none
    if with_source:
        return rec["bytes_per_launch"], ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this command on this "
                                         "kernel source, calibrated by a copy kernel (%s)" % rec.get("source", ""))
    return rec["bytes_per_launch"]


def push_kernel_digest():
    """SHA-256 of the source that k_push is compiled from: lf_mincut.hip up to
    the end of the kernel (its state, helpers and body; the host code after it
    does not change what a launch reads and writes)."""
    import hashlib
    src = os.path.join(ROOT, "latticeflow_amd", "csrc", "lf_mincut.hip")
    try:
        text = open(src, "rb").read()
    except OSError:
        return None
    end = text.find(b"// Work list of the tiles whose flag is set")
    return hashlib.sha256(text[:end] if end > 0 else text).hexdigest()


def copy_probe():
    import torch
    n = 1 << 28  # 2^28 floats = 1 GiB each way
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n, "launches": 5}), flush=True)


def bench_batch(a):
    """Batched min-cuts: 512 Full2d_8 segmentation lattices of uneven sizes
    (sides log-uniform in [64, 1024], aspect ratio in [0.5, 2], each lattice
    with its own seed), packed on one device canvas and solved by one min-cut.
    A step is one lf_run over the whole batch; value = sum of edges / step time."""
    import numpy as np
    import torch  # noqa: F401
    from concurrent.futures import ThreadPoolExecutor
    from latticeflow_amd import LatticeFlowBatch, workloads as W
    count = int(a.shape) if a.shape else 512
    inst = W.batch_instances(seed=4, count=count)
    shapes = [s for s, _ in inst]
    threads = min(16, os.cpu_count() or 1)
    t0 = time.time()
    bt = LatticeFlowBatch("Full2d_8", shapes, numerical_precision=1.0, cap_bytes=4)
    offs = None
    edges = 0
    keep = {}
    sample = list(range(0, count, max(1, count // 128)))[:128]
    with ThreadPoolExecutor(threads) as ex:
        for i, w in enumerate(ex.map(lambda t: W.seg8(seed=t[1], shape=t[0]), inst)):
            bt.addEnergyInt(i, w["unary"], w["pairwise"])
            if offs is None:
                offs = bt.edges(i)
            e = bt.edges(i)
            edges += W.n_edges(np.concatenate([e, -e]), shapes[i])
            if i in sample and not a.no_cpu_baseline:
                keep[i] = w
    if a.verbose:
        print("generated and ingested %d lattices in %.1fs, canvas %s" % (count, time.time() - t0, bt.canvas()),
              file=sys.stderr, flush=True)
    if a.rounds or a.iters or a.tile_shape or a.relabel_launches:
        bt.setTuning(rounds_per_relabel=a.rounds, local_iters=a.iters, tile_shape=a.tile_shape,
                     relabel_launches=a.relabel_launches)
    elif not a.no_autotune:
        bt.autotune(reps=2)
    tuning = bt.getTuning()
    for _ in range(a.warmup):
        bt.run()
    push_s = 0.0
    tiles = 0
    launches = 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        bt.run()
        flows = bt.flows()
        st = bt.stats()
        push_s += st["push_seconds"]
        tiles += st["tiles_pushed"]
        launches += st["rounds"]
    dt = (time.perf_counter() - t0) / a.steps
    bytes_alg = tiles * st["tile_nodes"] * st["bytes_per_node"]
    achieved = bytes_alg / push_s / 1e9 if push_s > 0 else 0.0
    roofline = {"bound": "hbm", "kernel": "k_push (push/relabel round)", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_us": round(push_s / max(launches, 1) * 1e6, 2), "bytes_per_node": st["bytes_per_node"],
                "nodes_per_launch": round(tiles * st["tile_nodes"] / max(launches, 1), 1),
                "push_share_of_step": round(push_s / a.steps / dt, 3)}
    cpu = None
    parity = None
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        O.build()
        prep = {}
        for i, w in keep.items():
            o = np.concatenate([bt.edges(i), -bt.edges(i)])
            prep[i] = (w["shape"], o) + O.reparam(w["shape"], o, w["unary"], w["pairwise"])
        tc = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            refs = dict(zip(prep, ex.map(lambda i: O.prflow(*prep[i]), prep)))
        tcpu = time.perf_counter() - tc
        sedges = sum(W.n_edges(prep[i][1], prep[i][0]) for i in prep)
        cuts = bt.getCuts()
        parity = all(refs[i]["flow"] == flows[i] and np.array_equal(cuts[i].reshape(-1), refs[i]["smin"].astype(np.int32))
                     for i in refs)
        cpu = {"value": round(sedges / tcpu / 1e6, 3), "unit": "Medges/s", "cores": threads, "kind": "port",
               "sample": "%d of the %d lattices (every %dth, %d edges), one PRFlow-restatement solve each "
                         "(oracle/lf_oracle.c) on a pool of %d threads, %.2f s"
                         % (len(prep), count, max(1, count // 128), sedges, threads, tcpu),
               **host_cpu()}
    hc, wc = bt.canvas()
    out = {
        "metric": "Batched min-cut Medges/s (%d Full2d_8 lattices, sides 64..1024)" % count,
        "value": round(edges / dt / 1e6, 3), "unit": "Medges/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "batch%d seg8 Full2d_8 (seeds 40000+i)" % count, "lattices": count,
                   "edges": edges, "nodes": bt.size, "canvas": [hc, wc],
                   "packing": round(bt.size / (hc * wc), 4), "flow_sum": int(flows.sum()),
                   "parity_vs_cpu_sample": parity, "cycles": st["cycles"], "rounds": st["rounds"],
                   "bfs_launches": st["bfs_launches"], "tuning": tuning},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_tvpath(a):
    """Exact TV regularization path: 1024x1024 Voronoi image, Full2d_8, 32 lambdas
    log-spaced in [0.01, 2], all lambdas solved together on the device.  Plain:
    divide and conquer with the mean threshold; learned: thresholds from the
    shipped net's predictions (same output).  value = plain path time."""
    print(json.dumps(tvpath_record(a, full_line=True)), flush=True)


def tvpath_record(a, full_line=False):
    import numpy as np
    import torch
    torch.zeros(1, device="cuda")
    from latticeflow_amd import workloads as W
    from latticeflow_amd.pylatticeflow import _tv_path
    from latticeflow_amd.tvnet import TVNet, tv_path_learned
    shape = tuple(int(x) for x in a.shape.split("x")) if a.shape else (1024, 1024)
    f = W.tv_voronoi(seed=5, shape=shape)
    lams = W.tv_lambdas()
    net = TVNet.load()
    for i in range(a.warmup):
        _tv_path(f, lams, "Full2d_8")
        if i == 0:
            # the net's first forward pass loads its kernels (about 0.18 s once)
            tv_path_learned(f, lams, "Full2d_8", net=net)
    times, ltimes = [], []
    R = Rl = st = gs = None
    for _ in range(max(a.steps, 1)):
        t0 = time.perf_counter()
        R, st = _tv_path(f, lams, "Full2d_8")
        times.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        Rl, stl, gs = tv_path_learned(f, lams, "Full2d_8", net=net)
        torch.cuda.synchronize()
        ltimes.append(time.perf_counter() - t0)
    dt = sum(times) / len(times)
    dl = sum(ltimes) / len(ltimes)
    same = float(np.abs(Rl - R).max() / np.abs(R).max())
    cpu = None
    parity = None
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        O.build()
        pick = [16, 31]
        tc = time.perf_counter()
        Rr = O.tv_path(f, lams[pick], "Full2d_8")
        tcpu = time.perf_counter() - tc
        cpu = {"value": round(tcpu / len(pick) * len(lams), 2), "unit": "s (32-lambda path, extrapolated)",
               "cores": 1, "kind": "port",
               "sample": "lambdas %s on the full %s image with the oracle's divide and conquer, %.2f s for %d "
                         "lambdas, scaled to 32" % ([round(float(lams[i]), 4) for i in pick],
                                                     "x".join(map(str, shape)), tcpu, len(pick)),
               **host_cpu()}
        rel = float(np.abs(R[pick] - Rr).max() / np.abs(Rr).max())
        parity = {"max_rel_err": rel, "within_1e-9": rel <= 1e-9}
    if not full_line:
        return {"workload": "tvpath %s Full2d_8 Voronoi(256 cells)+N(0,0.1), 32 lambdas log[0.01,2]"
                % "x".join(map(str, shape)), "plain_seconds": round(dt, 3), "learned_seconds": round(dl, 3),
                "speedup_learned_vs_plain": round(dt / dl, 3), "max_rel_diff_learned_vs_plain": same,
                "parity_vs_cpu": parity, "guided": gs, "cpu_baseline": cpu}
    out = {
        "metric": "TV-path wall time (exact, 32 lambdas), plain divide and conquer",
        "value": round(dt, 3), "unit": "s", "n_gpus": 1, "steps": len(times), "warmup": a.warmup,
        "ms_per_step": round(dt * 1e3, 1), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64", "data": "synthetic",
        "config": {"workload": "tvpath %s Full2d_8 Voronoi(256 cells)+N(0,0.1), 32 lambdas log[0.01,2]"
                   % "x".join(map(str, shape)), "levels_max": st["reserved"], "cycles": st["cycles"],
                   "rounds": st["rounds"], "bfs_launches": st["bfs_launches"], "parity_vs_cpu": parity,
                   "learned": {"seconds": round(dl, 3), "speedup_vs_plain": round(dt / dl, 3),
                               "max_rel_diff_vs_plain": same, "guided": gs,
                               "net": "3-layer 3x3 conv, 16 channels, bf16 MFMA (latticeflow_amd/data/tvnet_params.npy)"}},
        "cpu_baseline": cpu,
    }
    return out


if __name__ == "__main__":
    main()
