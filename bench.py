"""Benchmark of the MI355X tiled push-relabel grid cut.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--tile 32]
                    [--no-batch] [--no-warm] [--no-serve] [--no-cpu] ...

One step = one plain cut of the 4096x4096 workload (three foreground
ellipses, noise sigma 30, seed 1): init from the row-major int32 capacities
already resident in HBM, push-relabel to termination, minimal source set,
labels written to HBM.  `value` is Mpix/s of that plain cut.  Beside it, in
the same JSON line:

  roofline  the dominant kernel's algorithmic bytes per launch over its mean
            launch time (HIP events on the cut's stream), and the whole
            cut's algorithmic bytes over its step time;
  tiles     the same cut at tiles 16, 32 and 64;
  warm      a 2048x2048 image and 100 seeded brush strokes: stroke-to-mask
            latency (device brush, reflow, packed mask in host memory) warm
            vs from scratch;
  served    256 resident sessions (512..4096 squared, weights 4:3:2:1),
            Poisson arrivals, 10% new images / 90% strokes drawn afresh:
            stroke latency p50/p99 and the sustained stroke rate per serving
            mode, found by bisection;
  batched   64 ragged images (W, H ~ U[256, 2048]) solved together, tile /
            sweeps / push rounds per relabel picked by the autotuner, against
            64 sequential plain cuts;
  cpu_baseline  the oracle's restatement of the reference push-relabel on
            the 512x512 workload, on the host's cores;
  parity    (last) every checked result against the oracle's answers for
            these workloads (tests/golden/bench_expected.json, written by
            tests/golden/make_bench_golden.py): flow values and sha256 of the
            packed minimal source sets.
"""
import argparse
import concurrent.futures as cf
import glob
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
EXPECTED = os.path.join(HERE, "tests", "golden", "bench_expected.json")

# Algorithmic HBM bytes per pixel touched, per kernel class:
#   push_first 16 read per image pixel (source, sink, right, down) + 21 written
#              per padded pixel (e, 4 caps, residual mask; with one push round
#              per relabel its heights are not stored: the full relabel that
#              follows rewrites them)
#   push       49 per pixel per later residency: 24 read + 25 written
#   bfs_init    5 per pixel of a relabel's first pass: residual mask read, h written
#   bfs_round   9 per pixel of a later BFS pass: mask + h read, h written
#   flood       2 per pixel of a flood pass: mask read, reach written
#   labels      2 per image pixel: reach read, label written (+ 1 bit packed)
ALG_BYTES_PER_PX = {"push": 49, "bfs_init": 5, "bfs_round": 9, "flood": 2, "labels": 2}


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
    ap.add_argument("--rounds", type=int, default=1,
                    help="plain cut: device-driven BFS/flood rounds per relabel (0: host-driven)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-sweep", action="store_true", help="skip the tile sweep")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched leg")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-start leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--serve-modes", default="plain,cold,batched",
                    help="serving modes to run (comma-separated)")
    ap.add_argument("--serve-rates", default="",
                    help="probe these rates (comma-separated requests/s) instead of bisecting")
    ap.add_argument("--serve-image-frac", type=float, default=0.1,
                    help="share of new-image requests in the serving mix")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0: the usable cores, at most 16)")
    ap.add_argument("--serve-max-batch", type=int, default=64,
                    help="most requests per served solve")
    ap.add_argument("--serve-inline", action="store_true",
                    help="run the serving leg in this process (default: a child process)")
    ap.add_argument("--serve-child", action="store_true", help=argparse.SUPPRESS)
    return ap.parse_args()


def load_expected():
    with open(EXPECTED) as f:
        return json.load(f)


def answer(flow, bits, exp):
    """Parity record of one result against the oracle's answer."""
    import gcmi_serving as S
    sha = S.mask_sha(bits)
    return {"flow_equal": int(flow) == exp["flow"], "mask_equal": sha == exp["mask_sha"],
            "mask_sha": sha}


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


def class_bytes(stats, tile, W, H, bfs=None):
    """Algorithmic bytes per run of each kernel class (from the run's
    counters; round 1 stores its heights with more than one push round per
    relabel)."""
    T2 = tile * tile
    ntiles = stats["tiles"]
    first_written = 25 if (bfs or 1) > 1 else 21
    return {
        "push_first": 16 * W * H + first_written * ntiles * T2,
        "push": ALG_BYTES_PER_PX["push"] * (stats["residencies"] - ntiles) * T2,
        "bfs_init": ALG_BYTES_PER_PX["bfs_init"] * stats["relabel_tiles"] * T2,
        "bfs_round": ALG_BYTES_PER_PX["bfs_round"] * stats["bfs_round_passes"] * T2,
        "flood": ALG_BYTES_PER_PX["flood"] * stats["flood_passes"] * T2,
        "labels": ALG_BYTES_PER_PX["labels"] * W * H,
    }


def kernel_table(totals, stats, tile, W, H, steps, bfs=None):
    """Per kernel class of a profiled (host-driven) run: time per step, mean
    launch time, algorithmic bytes per launch, GB/s and fraction of peak."""
    per_run = class_bytes(stats, tile, W, H, bfs)
    out = {}
    for k, (ms, launches) in totals.items():
        if not launches:
            continue
        row = {"ms_per_step": round(ms / steps, 4), "launches_per_step": launches / steps,
               "avg_launch_us": round(ms / launches * 1e3, 2)}
        if per_run.get(k):
            b = per_run[k] * steps / launches
            gbs = b / (ms / launches / 1e3) / 1e9
            row.update(alg_bytes_per_launch=int(b), achieved_gbs=round(gbs, 1),
                       frac=round(gbs / HBM_PEAK_GBS, 4))
        out[k] = row
    return out


def roofline(table, W, tile):
    """The dominant kernel class (by time) that has an algorithmic byte
    count."""
    dom = max((k for k in table if "achieved_gbs" in table[k]),
              key=lambda k: table[k]["ms_per_step"])
    r = table[dom]
    traffic, source = pmc_traffic(dom, W, tile)
    return {"kernel": dom, "bound": "hbm", "achieved": r["achieved_gbs"],
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": r["frac"], "traffic": traffic,
            "traffic_source": source, "alg_bytes_per_launch": r["alg_bytes_per_launch"],
            "avg_launch_us": r["avg_launch_us"]}


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


def tile_leg(args, dev, W, H, exp, sweeps, bfs):
    """The plain cut at tiles 16, 32 and 64: Mpix/s, whole-cut algorithmic
    GB/s, the first push round's GB/s, flow and mask against the oracle."""
    from GraphCut import GraphCut
    out = {}
    for t in (16, 32, 64):
        g = GraphCut(None, None, (W, H), tile=t)
        g.set_local_relabel(bool(args.local))
        g.set_device_rounds(args.rounds)
        g.upload_device(*dev)
        n = max(3, args.steps // 4)
        ts, _, st = time_cut(g, n, 1, sweeps, bfs)
        flow = g.flow
        bits = g.labels_packed()
        _, tot, pst = time_cut(g, 3, 0, sweeps, bfs, profile=True)
        tab = kernel_table(tot, pst, t, W, H, 3, bfs)
        out[str(t)] = {"ms": round(ts * 1e3, 4), "mpix_s": round(W * H / ts / 1e6, 1),
                       "whole_cut_gbs": round(st["alg_bytes"] / ts / 1e9, 1),
                       "push_first_gbs": tab.get("push_first", {}).get("achieved_gbs"),
                       "parity": answer(flow, bits, exp)}
        g.close()
    return out


# ------------------------------------------------------------------- warm


def warm_leg(args, exp):
    """2048x2048, 100 strokes: warm reflow vs from scratch, per stroke
    paint + cut + packed mask readback; every stroke's flow and mask
    checked (outside the timed region)."""
    import gcmi_lib as L
    import gcmi_serving as S
    import gcmi_workloads as G
    from GraphCut import GraphCut
    W = H = 2048
    log("warm: generating the 2048x2048 image and 100 strokes")
    work = G.centred(W, H, seed=3)
    strokes = G.strokes(W, H, count=100, seed=3, raster=False)
    res = {}
    bad = {}
    # warm reflows host-driven with full relabels, with local relabels, and
    # device-driven with local relabels (2 or 1 rounds per relabel); scratch:
    # every stroke from scratch, with the plain cut's settings
    variants = {"warm_host": {}, "warm_local": dict(local=True),
                "warm_device": dict(local=True, rounds=2),
                "warm_device1": dict(local=True, rounds=1),
                "scratch": dict(local=True, rounds=1)}
    for mode, opt in variants.items():
        gc = GraphCut(None, None, (W, H), tile=args.tile)
        gc.set_local_relabel(opt.get("local", False))
        gc.set_device_rounds(opt.get("rounds", 0))
        gc.upload(*work.arrays())
        gc.cut()
        # the mask lands in a page-locked buffer, as a latency-bound client's would
        mbuf = L.pinned_empty(((W * H + 31) // 32,), np.uint32)
        gc.labels_packed(out=mbuf)
        lat = []
        nbad = 0
        for k, st in enumerate(strokes):
            s, t = st.capacities()
            t0 = time.perf_counter()
            gc.paint(st.points, st.width, s, t)
            if mode == "scratch":
                gc.forget()
            flow = gc.cut()
            bits = gc.labels_packed(out=mbuf)
            lat.append((time.perf_counter() - t0) * 1e3)
            a = answer(flow, bits, exp[k])
            nbad += not (a["flow_equal"] and a["mask_equal"])
        bad[mode] = nbad
        res[mode] = {"p50_ms": round(float(np.percentile(lat, 50)), 3),
                     "p99_ms": round(float(np.percentile(lat, 99)), 3),
                     "mean_ms": round(float(np.mean(lat)), 3)}
        gc.close()
    best = min((m for m in variants if m != "scratch"), key=lambda m: res[m]["mean_ms"])
    out = {"workload": "2048x2048 centred ellipse (seed 3), 100 strokes (seed 3), "
                       "device brush, packed mask to host", "tile": args.tile,
           "warm": dict(res[best], variant=best), "scratch": res["scratch"],
           "variants": res,
           "speedup_mean": round(res["scratch"]["mean_ms"] / res[best]["mean_ms"], 2)}
    parity = {"strokes_checked": len(strokes) * len(variants),
              "mismatches": sum(bad.values()), "per_variant_mismatches": bad}
    return out, parity


# ---------------------------------------------------------------- serving


def verify_served(srv, sids, mix, exp):
    """The verification trace (gcmi_serving.verification_trace): every
    session gets its base image again, then the trace is submitted from one
    thread; every request's flow and every session's final mask against the
    oracle's answers."""
    import gcmi_serving as S
    S.reset_sessions(srv, sids, mix, rgb=True)
    items = S.verification_trace(mix)
    t0 = time.monotonic_ns() + int(0.1e9)
    ids = srv.replay([(t0 + x[0], sids[x[1]], x[2], x[3]) for x in items])
    flows_bad = 0
    for j, rid in enumerate(ids):
        info = srv.wait(int(rid))
        srv.forget(int(rid))
        flows_bad += info.flow != exp["request_flows"][j]
    masks_bad = 0
    for i in range(mix.nsess):
        bits, _ = srv.mask_bits(sids[i])
        e = exp["sessions"][i]
        masks_bad += S.mask_sha(bits) != e["mask_sha"]
    return {"requests_checked": len(ids), "request_flow_mismatches": int(flows_bad),
            "sessions_checked": mix.nsess, "session_mask_mismatches": int(masks_bad)}


def serve_leg(args, exp):
    """The 256-session mix.  Modes: plain (one request at a time, each cut
    from scratch: the baseline), cold (continuous batching, every request
    cut from scratch) and batched (continuous batching, strokes reflowed
    warm: device-driven runs with local relabels).  New images arrive as
    8-bit colour images in every mode.  Each mode's sustained rate (p99 <=
    20 ms on 2048^2 sessions, stroke and new-image throughputs within 10% of
    the offer) is found by bisection to 2%; then the verification trace is
    checked."""
    import gcmi_serving as S
    from GraphCutServer import GraphCutServer
    log("served: generating the 256-session mix")
    mix = S.Mix(nsess=256, seed=4)
    # every mode runs with local relabels after the first and 1 device BFS
    # round per relabel (measured best for the plain server's per-request
    # cuts and the batched server's solves alike)
    fast = dict(local_relabel=True, device_rounds=1)
    modes = {"plain": (dict(mode="plain", **fast), (2000, 4000)),
             "cold": (dict(mode="cold", **fast), (2000, 6000)),
             "batched": (dict(mode="batched", **fast), (16000, 36000))}
    stroke_only = {"plain": (2000, 4000), "batched": (24000, 48000)}
    out = {"workload": "256 sessions, sizes {512,1024,2048,4096}^2 weights 4:3:2:1 (seed 4), "
                       "Poisson arrivals, %g%% new images (8-bit colour, capacities built on "
                       "the device) / the rest polyline strokes drawn afresh per request, "
                       "%.1f s per probed rate" % (100 * args.serve_image_frac,
                                                   args.serve_seconds),
           "resident_pixels": mix.pixels, "hbm_budget_bytes": 200 * 10**9, "modes": {},
           "stroke_only": {}}
    only = [m for m in args.serve_modes.split(",") if m]
    modes = {m: v for m, v in modes.items() if m in only}
    fixed = tuple(float(x) for x in args.serve_rates.split(",")) if args.serve_rates else None
    parity = {}
    seed = [11]

    def search(srv, sids, bracket, image_frac, rgb):
        def probe(rate):
            # every probe starts from the sessions' base images, so that the
            # strokes earlier probes painted (many, at overloaded rates) do
            # not carry over, whatever order the bisection probes in
            seed[0] += 1
            S.reset_sessions(srv, sids, mix, rgb=True)
            S.run_trace(srv, sids, mix, rate, duration=0.3, seed=1000 + seed[0],
                        image_frac=image_frac, rgb=rgb)
            pt = S.run_trace(srv, sids, mix, rate, duration=args.serve_seconds, seed=seed[0],
                             image_frac=image_frac, rgb=rgb)
            log("served: %s rate %g: p99(2048) %s ms, throughput %s/s" % (
                srv.mode, rate, pt.get("stroke_p99_ms_2048"), pt["stroke_throughput"]))
            return pt
        if fixed:
            pts = [probe(r) for r in fixed]
            return pts, S.sustained(pts)
        return S.find_sustained(probe, *bracket)

    for mode, (kw, bracket) in modes.items():
        log("served: mode %s" % mode)
        with GraphCutServer(tile=32, hbm_budget=200 * 10**9, max_sessions=256,
                            max_batch=args.serve_max_batch, **kw) as srv:
            sids = S.open_sessions(srv, mix, rgb=True)
            # an unmeasured warm-up trace (first-use allocations, page-in)
            S.run_trace(srv, sids, mix, bracket[0], duration=0.5, seed=7,
                        image_frac=args.serve_image_frac, rgb=True)
            pts, best = search(srv, sids, bracket, args.serve_image_frac, True)
            out["modes"][mode] = {"points": pts, "sustained_2048_p99_20ms": best,
                                  "device_bytes": srv.stats()["device_bytes"]}
            if mode in stroke_only and not fixed:
                log("served: mode %s, strokes only" % mode)
                pts, best = search(srv, sids, stroke_only[mode], 0.0, False)
                out["stroke_only"][mode] = {"points": pts, "sustained_2048_p99_20ms": best}
            log("served: mode %s, verification trace" % mode)
            parity[mode] = verify_served(srv, sids, mix, exp)
        # the closed server's device memory is released (and cleared by the
        # driver) in the background: let that finish before the next mode
        if mode != list(modes)[-1]:
            time.sleep(3.0)

    summarise(out)
    return out, parity


def summarise(out):
    """The sustained stroke rates of the served modes and their ratios to
    the plain mode's."""
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


# ---------------------------------------------------------------- batched


def batched_leg(args, exp):
    """64 ragged images: autotuned batched run vs 64 sequential plain cuts."""
    import gcmi_serving as S
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

    # autotuner: tile x sweeps x push rounds per relabel x run mode (host-
    # driven; device-driven with r BFS / flood rounds per relabel and one
    # synchronisation; the same recorded once as a hipGraph and replayed) x
    # relabels after the first (full or local), two timed runs each
    log("batched: autotuning")
    modes = [(0, 0), (1, 0), (2, 0), (0, 1), (0, 2), (0, 4)]  # (device rounds, graph rounds)

    def set_mode(b, device, graph):
        b.set_graph(graph)  # (0: host-driven, no recording)
        if device:
            b.set_device_rounds(device)

    tune = []
    batches = {}
    for tile in (16, 32, 64):
        batches[tile] = make_batch(tile)
        for sweeps in (2, 4, 8):
            for bfs in (1, 2):
                for device, graph in modes:
                    for local in (0, 1):
                        set_mode(batches[tile], device, graph)
                        batches[tile].set_local_relabel(bool(local))
                        dt = time_batch(batches[tile], sweeps, bfs, 2)
                        tune.append({"tile": tile, "sweeps": sweeps, "bfs": bfs, "device": device,
                                     "graph": graph, "local": local, "ms": round(dt * 1e3, 3)})
    best = min(tune, key=lambda r: r["ms"])
    b = batches[best["tile"]]
    set_mode(b, best["device"], best["graph"])
    b.set_local_relabel(bool(best["local"]))
    reps = max(3, args.steps // 2)
    t_batch = time_batch(b, best["sweeps"], best["bfs"], reps)
    host = min((r for r in tune if r["graph"] == 0 and r["device"] == 0), key=lambda r: r["ms"])
    flows = b.cut(sweeps=best["sweeps"], bfs=best["bfs"])
    stats = b.stats()
    nbad = 0
    for i in range(len(works)):
        a = answer(flows[i], S.pack_mask(b.labels(i)), exp[i])
        nbad += not (a["flow_equal"] and a["mask_equal"])
    # sequential plain cuts with the same tile, sweeps and push rounds, and
    # the plain cut's own run settings (local relabels, 1 device round)
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
    seq_bad = sum(g.flow != e["flow"] for g, e in zip(cuts, exp))
    for g in cuts:
        g.close()
    for x in batches.values():
        x.close()
    out = {"workload": "64 images, W,H ~ U[256,2048] (seed 2), centred ellipse, sigma 20",
           "pixels": pixels, "autotune": best, "autotune_trials": len(tune),
           "batched_ms": round(t_batch * 1e3, 3),
           "batched_mpix_s": round(pixels / t_batch / 1e6, 1),
           "sequential_ms": round(t_seq * 1e3, 3),
           "sequential_mpix_s": round(pixels / t_seq / 1e6, 1),
           "speedup": round(t_seq / t_batch, 2),
           "best_host_driven": host, "host_syncs": stats["host_syncs"],
           "rounds": stats["rounds"], "global_relabels": stats["global_relabels"]}
    parity = {"images_checked": len(works), "mismatches": int(nbad),
              "sequential_flow_mismatches": int(seq_bad)}
    return out, parity


# ------------------------------------------------------------------- CPU


def host_cpu():
    """nproc (the cores this process may use), the CPU model, the OpenCL
    ICDs installed and any pocl library (the OpenCL CPU runtime that running
    the reference's own kernels on the host would need)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    icds = sorted(glob.glob("/etc/OpenCL/vendors/*.icd"))
    pocl = sorted(set(sum((glob.glob(os.path.join(d, "libpocl*"))
                           for d in ("/usr/lib", "/usr/lib/x86_64-linux-gnu", "/usr/local/lib",
                                     "/opt/rocm/lib", "/usr/lib64")), [])))
    return {"nproc": len(os.sched_getaffinity(0)), "cpu_model": model,
            "opencl_icds": icds, "pocl_libraries": pocl}


def cpu_baseline(seconds, threads):
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import gcmi_workloads as G
    import oracle as O
    info = host_cpu()
    if threads <= 0:
        threads = max(1, min(16, info["nproc"]))
    w = G.centred(512, 512, seed=0)
    a = w.arrays()
    O.pr_cut(*a)  # warm (load, page in)
    stop = time.perf_counter() + seconds

    def worker(_):
        n = 0
        while time.perf_counter() < stop:
            O.pr_cut(*a, bfs=5)  # ctypes releases the GIL during the cut
            n += 1
        return n

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        n = sum(ex.map(worker, range(threads)))
    dt = time.perf_counter() - t0
    # The reference's kernels would need an OpenCL CPU device (pocl); an
    # installed ICD of AMD's OpenCL (amdocl64) exposes GPUs only.
    if info["pocl_libraries"]:
        runtime = "pocl is installed on this host but not used"
    else:
        runtime = "no OpenCL CPU runtime (pocl) on this host"
        if info["opencl_icds"]:
            runtime += "; the OpenCL ICDs present are GPU runtimes"
    return {"value": round(n * w.pixels / dt / 1e6, 3), "unit": "Mpix/s", "cores": threads,
            "kind": "port",
            "sample": "%d cuts of the 512x512 centred-ellipse workload (seed 0) by the "
                      "oracle's C restatement of the reference push-relabel, %d threads of "
                      "independent cuts, %.1f s (oracle CPU, not the reference: %s)"
                      % (n, threads, dt, runtime),
            "host": info}


def serve_in_child(args):
    """Each serving mode in a child process of its own, as a server runs:
    its HIP context holds only that server's streams.  (A process's streams
    share the device's four hardware queues.  With another stream in the
    process -- the null stream of earlier copies, or a server created and
    closed before -- the served lanes were measured to lose about a third
    of their throughput.)"""
    import subprocess
    out, parity = None, {}
    for mode in [m for m in args.serve_modes.split(",") if m]:
        cmd = [sys.executable, os.path.abspath(__file__), "--serve-child",
               "--serve-seconds", str(args.serve_seconds), "--serve-modes", mode,
               "--serve-image-frac", str(args.serve_image_frac),
               "--serve-max-batch", str(args.serve_max_batch)]
        if args.serve_rates:
            cmd += ["--serve-rates", args.serve_rates]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, check=True)
        d = json.loads(r.stdout.strip().splitlines()[-1])
        if out is None:
            out = d["served"]
        else:
            out["modes"].update(d["served"]["modes"])
            out["stroke_only"].update(d["served"]["stroke_only"])
        parity.update(d["parity"])
    summarise(out)
    return out, parity


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has a single MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2
    if args.serve_child:
        served, parity = serve_leg(args, load_expected()["served"])
        print(json.dumps({"served": served, "parity": parity}), flush=True)
        return 0
    import gcmi_serving as S
    import gcmi_workloads as G
    from GraphCut import GraphCut
    from gcmi_lib import DeviceBuffer

    exp = load_expected()
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
    t_prof, totals, pstats = time_cut(gc, args.steps, args.warmup, sweeps, bfs, profile=True)
    t_plain, _, stats = time_cut(gc, args.steps, 1, sweeps, bfs, profile=False)
    value = W * H / t_plain / 1e6
    flow = gc.flow
    parity = {}
    if W == 4096 and H == 4096:
        parity["plain_4096"] = answer(flow, gc.labels_packed(), exp["plain_4096"])
    table = kernel_table(totals, pstats, args.tile, W, H, args.steps, bfs)
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
        "roofline": roofline(table, W, args.tile),
        "whole_cut": {"alg_bytes_per_cut": stats["alg_bytes"],
                      "achieved_gbs": round(stats["alg_bytes"] / t_plain / 1e9, 1),
                      "frac": round(stats["alg_bytes"] / t_plain / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels_host_driven": table,
        "ms_per_step_profiled_host_driven": round(t_prof * 1e3, 4),
    }
    if not args.no_sweep and W == 4096 and H == 4096:
        out["tiles"] = tile_leg(args, dev, W, H, exp["plain_4096"], sweeps, bfs)
        parity["tiles"] = {t: v["parity"] for t, v in out["tiles"].items()}
    gc.close()
    del dev
    # (serving runs before the batched leg: after the autotuner's many short
    # runs, the served throughput came out lower in the same process)
    if not args.no_warm:
        out["warm"], parity["warm_2048"] = warm_leg(args, exp["warm_2048"])
    if not args.no_serve:
        if args.serve_inline:
            out["served"], parity["served"] = serve_leg(args, exp["served"])
        else:
            out["served"], parity["served"] = serve_in_child(args)
    if not args.no_batch:
        out["batched"], parity["batched_64"] = batched_leg(args, exp["batched_64"])
    if not args.no_cpu:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds, args.cpu_threads)
    parity["all_equal"] = all_equal(parity)
    parity["expected"] = os.path.relpath(EXPECTED, HERE)
    out["parity"] = parity
    print(json.dumps(out), flush=True)
    return 0


def all_equal(parity):
    """True if every checked flow and mask equals the oracle's answer."""
    def ok(x):
        if isinstance(x, dict):
            for k, v in x.items():
                if k.endswith("_equal") and v is not True:
                    return False
                if "mismatches" in k and k != "per_variant_mismatches" and v:
                    return False
                if isinstance(v, dict) and not ok(v):
                    return False
        return True
    return ok(parity)


if __name__ == "__main__":
    sys.exit(main())
