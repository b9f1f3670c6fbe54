"""Benchmark of the MI355X tiled push-relabel grid cut.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--tile 32]

One step = one plain cut of the 4096x4096 workload (three foreground
ellipses, noise sigma 30, seed 1): init from the row-major int32 capacities
already resident in HBM, push-relabel to termination, minimal source set,
labels written to HBM.  Prints one JSON line; `value` is Mpix/s of the plain
cut.  The CPU baseline times the oracle's restatement of the reference's
push-relabel on the 512x512 workload, on one host core.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "cl-util_amd"))

import numpy as np  # noqa: E402

METRIC = ("graph-cut Mpix/s + p50/p99 stroke-to-mask latency, "
          "served/batched vs plain")
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per pixel touched, per kernel class:
#   push_first 16 read per image pixel (source, sink, right, down) + 25 written
#              per padded pixel (e, h, 4 caps, residual mask)
#   push       49 per pixel per later residency: 24 read + 25 written
#   bfs_init    5: residual mask read, h written
#   bfs_round   9: mask + h read, h written, per tile pass
#   flood       2: mask read, reach written, per tile pass
#   labels      2: reach read, label written
ALG_BYTES_PER_PX = {"push": 49, "bfs_init": 5, "bfs_round": 9, "labels": 2}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--tile", type=int, default=32)
    ap.add_argument("--sweeps", type=int, default=0)
    ap.add_argument("--bfs", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-parity", action="store_true", help="skip the oracle check")
    ap.add_argument("--no-sweep", action="store_true", help="skip the tile sweep")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    return ap.parse_args()


def time_cut(gc, steps, warmup, sweeps, bfs, profile=False):
    for _ in range(warmup):
        gc.cut(sweeps=sweeps, bfs=bfs)
    gc.profile(profile)
    totals = {}
    stats = None
    t0 = time.perf_counter()
    for _ in range(steps):
        gc.cut(sweeps=sweeps, bfs=bfs)
        if profile:
            for k, (ms, n) in gc.kernel_times().items():
                a = totals.setdefault(k, [0.0, 0])
                a[0] += ms
                a[1] += n
    dt = time.perf_counter() - t0
    stats = gc.stats()
    gc.profile(False)
    return dt / steps, totals, stats


def roofline(totals, stats, tile, W, H, steps):
    """Dominant kernel class by total time; algorithmic bytes per launch over
    its mean launch duration."""
    T2 = tile * tile
    ntiles = stats["tiles"]
    # algorithmic bytes each class moves per run
    per_run = {
        "push_first": 16 * W * H + 25 * ntiles * T2,
        "push": ALG_BYTES_PER_PX["push"] * (stats["residencies"] - ntiles) * T2,
        "bfs_init": ALG_BYTES_PER_PX["bfs_init"] * stats["global_relabels"] * ntiles * T2,
        "bfs_round": ALG_BYTES_PER_PX["bfs_round"]
        * (stats["bfs_passes"] - stats["global_relabels"] * ntiles) * T2,
        "labels": ALG_BYTES_PER_PX["labels"] * W * H,
    }
    dom = max((k for k in totals if per_run.get(k) and totals[k][1]),
              key=lambda k: totals[k][0])
    ms, launches = totals[dom]
    per_launch_bytes = per_run[dom] * steps / launches
    avg_s = ms / launches / 1e3
    achieved = per_launch_bytes / avg_s / 1e9
    return {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "alg_bytes_per_launch": int(per_launch_bytes),
            "avg_launch_us": round(avg_s * 1e6, 2)}


def cpu_baseline(seconds):
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import gcmi_workloads as G
    import oracle as O
    w = G.centred(512, 512, seed=0)
    a = w.arrays()
    O.pr_cut(*a)  # warm (load, page in)
    n = 0
    t0 = time.perf_counter()
    while True:
        O.pr_cut(*a, bfs=5)
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": round(n * w.pixels / dt / 1e6, 3), "unit": "Mpix/s", "cores": 1,
            "kind": "port",
            "sample": "%d cuts of the 512x512 centred-ellipse workload (seed 0) by the "
                      "oracle's C restatement of the reference push-relabel, %.1f s" % (n, dt)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has a single MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2
    import gcmi_workloads as G
    from GraphCut import GraphCut
    from gcmi_lib import DeviceBuffer

    W = H = args.size
    print("# generating %dx%d workload" % (W, H), file=sys.stderr, flush=True)
    work = G.three_ellipses(W, H, seed=1)
    dev = [DeviceBuffer(a.nbytes).upload(a) for a in work.arrays()]
    gc = GraphCut(None, None, (W, H), tile=args.tile)
    gc.upload_device(*dev)
    sweeps = args.sweeps or None
    bfs = args.bfs or None
    print("# timing", file=sys.stderr, flush=True)
    t_step, totals, stats = time_cut(gc, args.steps, args.warmup, sweeps, bfs, profile=True)
    t_plain, _, _ = time_cut(gc, args.steps, 1, sweeps, bfs, profile=False)
    value = W * H / t_plain / 1e6
    flow = gc.flow
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_plain * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": "plain single cut, %dx%d, 3 ellipses, sigma 30, seed 1" % (W, H),
                   "tile": args.tile, "sweeps": sweeps or 4, "bfs": bfs or 1,
                   "pixels": W * H},
        "flow": flow,
        "run_stats": stats,
        "roofline": roofline(totals, stats, args.tile, W, H, args.steps),
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in totals.items()},
        "ms_per_step_profiled": round(t_step * 1e3, 4),
    }
    if not args.no_parity:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle as O
        fo, mo = O.bk_cut(*work.arrays())
        out["parity"] = {"flow_equal": fo == flow,
                         "mask_equal": bool((gc.labels() == mo).all()),
                         "oracle": "Boykov-Kolmogorov (oracle/gc_oracle.c)"}
    if not args.no_sweep:
        sweep = {}
        for t in (16, 32, 64):
            g2 = GraphCut(None, None, (W, H), tile=t)
            g2.upload_device(*dev)
            ts, _, _ = time_cut(g2, max(3, args.steps // 4), 1, sweeps, bfs)
            sweep[str(t)] = round(W * H / ts / 1e6, 1)
            g2.close()
        out["tile_sweep_mpix_s"] = sweep
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
