"""Benchmark of the MI355X tiled push-relabel grid cut.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--tile 32]
                    [--no-batch] [--no-warm] [--no-serve] [--no-cpu] ...

One step = one plain cut of the 4096x4096 workload (three foreground
ellipses, noise sigma 30, seed 1): init from the row-major int32 capacities
already resident in HBM, push-relabel to termination, minimal source set,
labels written to HBM.  `value` is Mpix/s of that plain cut.  Beside it, in
the same JSON line:

  batched   64 ragged images (W, H ~ U[256, 2048]) solved together, tile /
            sweeps / push rounds per relabel picked by the autotuner, against
            64 sequential plain cuts;
  warm      a 2048x2048 image and 100 seeded brush strokes: stroke-to-mask
            latency (device brush, reflow, packed mask in host memory) warm
            vs from scratch;
  served    256 resident sessions (512..4096 squared, weights 4:3:2:1),
            Poisson arrivals, 10% new images / 90% strokes: stroke latency
            p50/p99 and sustained stroke rate per serving mode;
  cpu_baseline  the oracle's restatement of the reference push-relabel on
            the 512x512 workload, on one host core.
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


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--tile", type=int, default=32)
    ap.add_argument("--sweeps", type=int, default=0)
    ap.add_argument("--bfs", type=int, default=0)
    ap.add_argument("--local", type=int, default=1,
                    help="plain cut: local relabels after the first (1) or full relabels (0)")
    ap.add_argument("--rounds", type=int, default=2,
                    help="plain cut: device-driven BFS/flood rounds per relabel (0: host-driven)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-parity", action="store_true", help="skip the oracle check")
    ap.add_argument("--no-sweep", action="store_true", help="skip the tile sweep")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched leg")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-start leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--serve-modes", default="plain,cold,batched",
                    help="serving modes to run (comma-separated)")
    ap.add_argument("--serve-rates", default="",
                    help="replace each mode's rate sweep (comma-separated requests/s)")
    ap.add_argument("--serve-image-frac", type=float, default=0.1,
                    help="share of new-image requests in the serving mix")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--serve-max-batch", type=int, default=64,
                    help="most requests per served solve")
    return ap.parse_args()


# ------------------------------------------------------------------ plain


def time_cut(gc, steps, warmup, sweeps, bfs, profile=False):
    for _ in range(warmup):
        gc.cut(sweeps=sweeps, bfs=bfs)
    gc.profile(profile)
    totals = {}
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
    its mean launch duration (HIP events on the cut's stream)."""
    T2 = tile * tile
    ntiles = stats["tiles"]
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
    traffic, source = pmc_traffic(dom, W, tile) if W == H else (None, None)
    return {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": source,
            "alg_bytes_per_launch": int(per_launch_bytes),
            "avg_launch_us": round(avg_s * 1e6, 2)}


def pmc_traffic(kernel, size, tile):
    """HBM bytes per launch of `kernel` on the size x size workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (each counter scaled
    by what it reads for a copy of known size), or None if not profiled."""
    path = os.path.join(HERE, "profiles", "plain_cut_%d_t%d_pmc_traffic.md" % (size, tile))
    key = "traffic_bytes_per_launch %s %d %d =" % (kernel, size, tile)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key):
                    return int(line.split("=")[1]), os.path.relpath(path, HERE)
    except OSError:
        pass
    return None, None


# ---------------------------------------------------------------- batched


def batched_leg(args):
    """64 ragged images: autotuned batched run vs 64 sequential plain cuts."""
    import gcmi_workloads as G
    from GraphCut import GraphCut
    from GraphCutBatch import GraphCutBatch
    from gcmi_lib import DeviceBuffer
    log("batched: generating 64 ragged images")
    works = G.ragged_batch(64, 256, 2048, seed=2)
    dims = [(w.shape[1], w.shape[0]) for w in works]
    pixels = sum(w.pixels for w in works)
    dev = [[DeviceBuffer(a.nbytes).upload(a) for a in w.arrays()] for w in works]

    def make_batch(tile):
        b = GraphCutBatch(dims, tile=tile)
        for i, d in enumerate(dev):
            b.upload_device(i, *d)
        return b

    def time_batch(b, sweeps, bfs, reps):
        # an image not edited since its last solve is cut from scratch again
        b.cut(sweeps=sweeps, bfs=bfs)
        t0 = time.perf_counter()
        for _ in range(reps):
            b.cut(sweeps=sweeps, bfs=bfs)
        return (time.perf_counter() - t0) / reps

    # autotuner: tile x sweeps x push rounds per relabel x recording (0:
    # host-driven; r > 0: one replayed hipGraph with r BFS rounds per
    # relabel) x relabels after the first (full or local), two timed runs each
    log("batched: autotuning")
    tune = []
    batches = {}
    for tile in (16, 32, 64):
        batches[tile] = make_batch(tile)
        for sweeps in (2, 4, 8):
            for bfs in (1, 2):
                for graph in (0, 2, 4, 8):
                    for local in (0, 1):
                        batches[tile].set_graph(graph)
                        batches[tile].set_local_relabel(bool(local))
                        dt = time_batch(batches[tile], sweeps, bfs, 2)
                        tune.append({"tile": tile, "sweeps": sweeps, "bfs": bfs, "graph": graph,
                                     "local": local, "ms": round(dt * 1e3, 3)})
    best = min(tune, key=lambda r: r["ms"])
    b = batches[best["tile"]]
    b.set_graph(best["graph"])
    b.set_local_relabel(bool(best["local"]))
    reps = max(3, args.steps // 2)
    t_batch = time_batch(b, best["sweeps"], best["bfs"], reps)
    host = min((r for r in tune if r["graph"] == 0), key=lambda r: r["ms"])
    flows = b.cut(sweeps=best["sweeps"], bfs=best["bfs"])
    stats = b.stats()
    # sequential plain cuts with the same tile, sweeps and push rounds, and
    # the plain cut's own run settings (local relabels, 2 device rounds)
    log("batched: 64 sequential plain cuts")
    cuts = []
    for i, d in enumerate(dev):
        g = GraphCut(None, None, dims[i], tile=best["tile"])
        g.set_local_relabel(bool(args.local))
        g.set_device_rounds(args.rounds)
        g.upload_device(*d)
        cuts.append(g)
    for g in cuts:
        g.cut(sweeps=best["sweeps"], bfs=best["bfs"])
    t0 = time.perf_counter()
    for _ in range(reps):
        for g in cuts:
            g.cut(sweeps=best["sweeps"], bfs=best["bfs"])
    t_seq = (time.perf_counter() - t0) / reps
    same = all(g.flow == f for g, f in zip(cuts, flows))
    for g in cuts:
        g.close()
    for x in batches.values():
        x.close()
    return {"workload": "64 images, W,H ~ U[256,2048] (seed 2), centred ellipse, sigma 20",
            "pixels": pixels, "autotune": best, "autotune_trials": len(tune),
            "batched_ms": round(t_batch * 1e3, 3),
            "batched_mpix_s": round(pixels / t_batch / 1e6, 1),
            "sequential_ms": round(t_seq * 1e3, 3),
            "sequential_mpix_s": round(pixels / t_seq / 1e6, 1),
            "speedup": round(t_seq / t_batch, 2), "flows_equal_sequential": bool(same),
            "best_host_driven": host, "host_syncs": stats["host_syncs"],
            "rounds": stats["rounds"], "global_relabels": stats["global_relabels"]}


# ------------------------------------------------------------------- warm


def warm_leg(args):
    """2048x2048, 100 strokes: warm reflow vs from scratch, per stroke
    paint + cut + packed mask readback."""
    import gcmi_workloads as G
    from GraphCut import GraphCut
    W = H = 2048
    log("warm: generating the 2048x2048 image and 100 strokes")
    work = G.centred(W, H, seed=3)
    strokes = G.strokes(W, H, count=100, seed=3, raster=False)
    res = {}
    masks = {}
    flows = {}
    # warm reflows host-driven with full relabels, with local relabels, and
    # device-driven with local relabels (2 or 1 rounds per relabel); scratch:
    # every stroke from scratch, with the plain cut's settings
    variants = {"warm_host": {}, "warm_local": dict(local=True),
                "warm_device": dict(local=True, rounds=2),
                "warm_device1": dict(local=True, rounds=1),
                "scratch": dict(local=True, rounds=2)}
    for mode, opt in variants.items():
        gc = GraphCut(None, None, (W, H), tile=args.tile)
        gc.set_local_relabel(opt.get("local", False))
        gc.set_device_rounds(opt.get("rounds", 0))
        gc.upload(*work.arrays())
        gc.cut()
        gc.labels_packed()
        lat = []
        fl = []
        for st in strokes:
            s, t = st.capacities()
            t0 = time.perf_counter()
            gc.paint(st.points, st.width, s, t)
            if mode == "scratch":
                gc.forget()
            fl.append(gc.cut())
            bits = gc.labels_packed()
            lat.append((time.perf_counter() - t0) * 1e3)
        masks[mode] = bits
        flows[mode] = fl
        res[mode] = {"p50_ms": round(float(np.percentile(lat, 50)), 3),
                     "p99_ms": round(float(np.percentile(lat, 99)), 3),
                     "mean_ms": round(float(np.mean(lat)), 3)}
        gc.close()
    best = min((m for m in variants if m != "scratch"), key=lambda m: res[m]["mean_ms"])
    out = {"workload": "2048x2048 centred ellipse (seed 3), 100 strokes (seed 3), "
                       "device brush, packed mask to host", "tile": args.tile,
           "warm": dict(res[best], variant=best), "scratch": res["scratch"],
           "variants": res,
           "speedup_mean": round(res["scratch"]["mean_ms"] / res[best]["mean_ms"], 2),
           "flows_equal": all(flows[m] == flows["scratch"] for m in variants),
           "final_mask_equal": all(bool((masks[m] == masks["scratch"]).all()) for m in variants)}
    if not args.no_parity:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle as O
        cur = work.copy()
        for st in G.strokes(W, H, count=100, seed=3):
            G.apply_stroke(cur, st)
        fo, mo = O.bk_cut(*cur.arrays())
        packed = np.packbits(mo.reshape(-1), bitorder="little")
        packed = np.pad(packed, (0, (-packed.size) % 4)).view(np.uint32)
        out["parity_last_stroke"] = {"flow_equal": fo == flows[best][-1],
                                     "mask_equal": bool((packed == masks[best]).all())}
    return out


# ---------------------------------------------------------------- serving


def serve_leg(args):
    """The 256-session mix at swept Poisson rates.  Modes: plain (one request
    at a time, each cut from scratch: the baseline), cold (continuous
    batching, every request cut from scratch) and batched (continuous
    batching, strokes reflowed warm: device-driven runs with local
    relabels).  New images arrive as 8-bit colour images in every mode."""
    import gcmi_serving as S
    from GraphCutServer import GraphCutServer
    log("served: generating the 256-session mix")
    mix = S.Mix(nsess=256, seed=4)
    # every mode runs the plain cut's fastest settings (local relabels after
    # the first, 2 device rounds per relabel)
    fast = dict(local_relabel=True, device_rounds=2)
    modes = {"plain": (dict(mode="plain", **fast), (2000, 2500, 3000, 3500, 4000)),
             "cold": (dict(mode="cold", **fast), (2000, 3000, 4000, 6000)),
             "batched": (dict(mode="batched", **fast),
                         (16000, 20000, 24000, 26000, 28000, 30000, 32000, 34000))}
    stroke_only = {"plain": (1500, 2000, 2500, 3000, 3500),
                   "batched": (30000, 35000, 40000, 45000, 50000)}
    out = {"workload": "256 sessions, sizes {512,1024,2048,4096}^2 weights 4:3:2:1 (seed 4), "
                       "Poisson arrivals, %g%% new images (8-bit colour, capacities built on "
                       "the device) / the rest polyline strokes, %.1f s per rate"
                       % (100 * args.serve_image_frac, args.serve_seconds),
           "resident_pixels": mix.pixels, "hbm_budget_bytes": 200 * 10**9, "modes": {},
           "stroke_only": {}}
    only = [m for m in args.serve_modes.split(",") if m]
    modes = {m: v for m, v in modes.items() if m in only}
    if args.serve_rates:
        rates = tuple(float(x) for x in args.serve_rates.split(","))
        modes = {m: (kw, rates) for m, (kw, _) in modes.items()}
        stroke_only = {}
    for mode, (kw, rs) in modes.items():
        log("served: mode %s" % mode)
        with GraphCutServer(tile=32, hbm_budget=200 * 10**9, max_sessions=256,
                            max_batch=args.serve_max_batch, **kw) as srv:
            sids = S.open_sessions(srv, mix, rgb=True)
            # an unmeasured warm-up trace (first-use allocations, page-in)
            S.run_trace(srv, sids, mix, rs[0], duration=0.5, seed=7,
                        image_frac=args.serve_image_frac, rgb=True)
            pts = [S.run_trace(srv, sids, mix, r, duration=args.serve_seconds, seed=11 + k,
                               image_frac=args.serve_image_frac, rgb=True)
                   for k, r in enumerate(rs)]
            out["modes"][mode] = {"points": pts, "sustained_2048_p99_20ms": S.sustained(pts),
                                  "device_bytes": srv.stats()["device_bytes"]}
            if mode in stroke_only:
                # the same sessions, strokes only
                log("served: mode %s, strokes only" % mode)
                pts = [S.run_trace(srv, sids, mix, r, duration=args.serve_seconds,
                                   seed=51 + k, image_frac=0.0)
                       for k, r in enumerate(stroke_only[mode])]
                out["stroke_only"][mode] = {"points": pts,
                                            "sustained_2048_p99_20ms": S.sustained(pts)}
        # the closed server's device memory is released (and cleared by the
        # driver) in the background: let that finish before the next mode
        time.sleep(3.0)

    def rate(entry):
        return (entry["sustained_2048_p99_20ms"] or {}).get("stroke_rate")

    sus = {m: rate(v) for m, v in out["modes"].items()}
    out["sustained_stroke_rate"] = sus
    if sus.get("plain"):
        best = max((v for m, v in sus.items() if m != "plain" and v), default=None)
        out["served_vs_plain"] = round(best / sus["plain"], 2) if best else None
    so = {m: rate(v) for m, v in out["stroke_only"].items()}
    out["stroke_only_sustained_stroke_rate"] = so
    if so.get("plain") and so.get("batched"):
        out["stroke_only_served_vs_plain"] = round(so["batched"] / so["plain"], 2)
    return out


# ------------------------------------------------------------------- CPU


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
    log("generating %dx%d workload" % (W, H))
    work = G.three_ellipses(W, H, seed=1)
    dev = [DeviceBuffer(a.nbytes).upload(a) for a in work.arrays()]
    gc = GraphCut(None, None, (W, H), tile=args.tile)
    gc.set_local_relabel(bool(args.local))
    gc.set_device_rounds(args.rounds)
    gc.upload_device(*dev)
    sweeps = args.sweeps or None
    bfs = args.bfs or None
    log("timing the plain cut")
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
                   "local_relabel": bool(args.local), "device_rounds": args.rounds,
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
            g2.set_local_relabel(bool(args.local))
            g2.set_device_rounds(args.rounds)
            g2.upload_device(*dev)
            ts, _, _ = time_cut(g2, max(3, args.steps // 4), 1, sweeps, bfs)
            sweep[str(t)] = round(W * H / ts / 1e6, 1)
            g2.close()
        out["tile_sweep_mpix_s"] = sweep
    gc.close()
    del dev
    # (serving runs before the batched leg: after the autotuner's many short
    # runs, the served throughput came out lower in the same process)
    if not args.no_warm:
        out["warm"] = warm_leg(args)
    if not args.no_serve:
        out["served"] = serve_leg(args)
    if not args.no_batch:
        out["batched"] = batched_leg(args)
    if not args.no_cpu:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
