"""Serving workload: resident sessions, Poisson request traces, latency
statistics (host side).

The mix: `nsess` sessions with square sizes drawn from SIZES with weights
4:3:2:1 (seeded).  Session i starts from base image i mod K of its size
(three_ellipses with a per-size seed; K images per size), so that the
4096x4096 images are not generated 26 times.  Requests arrive as a Poisson
process; each picks a session uniformly; 10% are new images (the session's
next base image, cut from scratch; sent as its 8-bit colour image, whose
capacities the server builds on the device, or as the four int32 capacity
arrays) and 90% are brush strokes: polylines drawn afresh for every request
from the trace's seeded generator (3-8 segments with vertices uniform in the
image, width 8-32, foreground and background hard constraints alternating
per session), rasterised on the device.

Latency of a request = completion (mask in host memory) - scheduled
arrival, so any lag of the load generator counts against the server.
"""
import threading
import time

import numpy as np

import gcmi_workloads as G

SIZES = (512, 1024, 2048, 4096)
WEIGHTS = (4, 3, 2, 1)
BASES = {512: 4, 1024: 3, 2048: 2, 4096: 2}


class Mix:
    def __init__(self, nsess=256, seed=4, sizes=SIZES, weights=WEIGHTS, bases=BASES):
        rng = np.random.default_rng(seed)
        w = np.asarray(weights, float)
        self.sizes = [int(x) for x in rng.choice(sizes, size=nsess, p=w / w.sum())]
        self.bases = {n: [G.three_ellipses(n, n, seed=4000 + 10 * n + k) for k in range(bases[n])]
                      for n in sizes if n in self.sizes}
        self.nsess = nsess
        # strokes painted per session so far: the next one is foreground if even
        self.painted = np.zeros(nsess, np.int64)

    @property
    def pixels(self):
        return sum(n * n for n in self.sizes)

    def base(self, i, k=0):
        n = self.sizes[i]
        b = self.bases[n]
        return b[(i + k) % len(b)]


# Page-locked copies of the colour images the load generator replays (a few
# base images, each sent many times), so that their uploads run at the
# link's rate; made on first use, on the GPU side of a run.
_PINNED = {}


def pinned(a):
    hit = _PINNED.get(id(a))
    if hit is None or hit[0] is not a:
        import gcmi_lib as L
        hit = (a, L.pinned_copy(a))
        _PINNED[id(a)] = hit
    return hit[1]


def _pin_payload(kind, payload):
    return pinned(payload) if kind == "rgb" else payload


def image_request(work, rgb):
    """(kind, payload) of a new-image request for a workload."""
    return ("rgb", work.image) if rgb else ("image", work.arrays())


def open_sessions(srv, mix, rgb=False):
    """Opens every session and cuts its first image; returns the ids."""
    sids = [srv.open(n, n) for n in mix.sizes]
    reset_sessions(srv, sids, mix, rgb)
    return sids


def reset_sessions(srv, sids, mix, rgb=False):
    """Gives every session its first base image again (a known state: the
    strokes of earlier traces are gone) and waits for the cuts."""
    trace = []
    for i in range(mix.nsess):
        kind, payload = image_request(mix.base(i), rgb)
        trace.append((0, sids[i], kind, _pin_payload(kind, payload)))
    for r in srv.replay(trace):
        srv.wait(int(r))
        srv.forget(int(r))
    mix.painted[:] = 0


def make_trace(mix, rate, duration, seed, t0_ns, image_frac=0.1, rgb=False):
    """(strokes, images): lists of (t_ns, session index, kind, payload, size,
    source) where source is the Workload of a new image or the Stroke.
    Every stroke is drawn afresh (gcmi_workloads.random_stroke) and
    alternates foreground / background per session (mix.painted)."""
    rng = np.random.default_rng(seed)
    n = max(1, int(rate * duration))
    ts = t0_ns + np.cumsum(rng.exponential(1e9 / rate, size=n)).astype(np.int64)
    strokes, images = [], []
    nimg = np.zeros(mix.nsess, int)
    for k in range(n):
        i = int(rng.integers(mix.nsess))
        size = mix.sizes[i]
        if rng.random() < image_frac:
            nimg[i] += 1
            work = mix.base(i, int(nimg[i]))
            kind, payload = image_request(work, rgb)
            images.append((int(ts[k]), i, kind, payload, size, work))
        else:
            st = G.random_stroke(rng, size, size, mix.painted[i] % 2 == 0)
            mix.painted[i] += 1
            s, t = st.capacities()
            strokes.append((int(ts[k]), i, "polyline", (st.points, st.width, s, t), size, st))
    return strokes, images


IMAGE_THREADS = 2


def replay_trace(srv, sids, strokes, images):
    """Submits a trace's strokes from one generator thread and its new
    images from IMAGE_THREADS more (sessions split between them, so each
    session's images stay in order), so that an image's upload does not
    delay the strokes behind it; colour images go out from page-locked
    copies.  Returns {"s": stroke request ids, "i": image request ids}."""
    ids = {}
    errors = []
    parts = [[k for k, x in enumerate(images) if x[1] % IMAGE_THREADS == j]
             for j in range(IMAGE_THREADS)]

    def go(name, tr):
        try:
            ids[name] = srv.replay([(x[0], sids[x[1]], x[2], _pin_payload(x[2], x[3]))
                                    for x in tr]) if tr else np.zeros(0, np.int64)
        except Exception as e:  # reported by the caller
            errors.append(e)

    th = [threading.Thread(target=go, args=("s", strokes))]
    th += [threading.Thread(target=go, args=(j, [images[k] for k in part]))
           for j, part in enumerate(parts)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    if errors:
        raise errors[0]
    img = np.zeros(len(images), np.int64)
    for j, part in enumerate(parts):
        img[part] = ids.pop(j)
    ids["i"] = img
    return ids


def run_trace(srv, sids, mix, rate, duration=2.0, seed=11, image_frac=0.1, lead=0.2, rgb=False):
    """Replays a Poisson trace (replay_trace) and returns latency statistics
    in ms.  The trace is generated first and starts `lead` seconds after
    that, so the generator's own time is not counted as latency."""
    strokes, images = make_trace(mix, rate, duration, seed, 0, image_frac, rgb)
    t0 = time.monotonic_ns() + int(lead * 1e9)
    strokes = [(x[0] + t0,) + tuple(x[1:]) for x in strokes]
    images = [(x[0] + t0,) + tuple(x[1:]) for x in images]
    ids = replay_trace(srv, sids, strokes, images)
    lat = {"stroke": [], "image": []}
    by_size = {}
    batch = []
    t_end = {"s": t0, "i": t0}
    for name, tr in (("s", strokes), ("i", images)):
        for rid, item in zip(ids[name], tr):
            info = srv.wait(int(rid))
            srv.forget(int(rid))
            ms = (info.t_done - item[0]) / 1e6
            lat["stroke" if name == "s" else "image"].append(ms)
            if name == "s":
                by_size.setdefault(item[4], []).append(ms)
            batch.append(info.batch)
            t_end[name] = max(t_end[name], info.t_done)

    def pct(v, q):
        return round(float(np.percentile(v, q)), 3) if v else None

    out = {"rate": rate, "requests": len(strokes) + len(images),
           "strokes": len(strokes), "images": len(images),
           "stroke_rate": round(rate * len(strokes) / max(1, len(strokes) + len(images)), 1),
           "stroke_throughput": round(len(strokes) / max((t_end["s"] - t0) / 1e9, 1e-9), 1),
           "image_rate": round(rate * len(images) / max(1, len(strokes) + len(images)), 1),
           "image_throughput": round(len(images) / max((t_end["i"] - t0) / 1e9, 1e-9), 1),
           "stroke_p50_ms": pct(lat["stroke"], 50), "stroke_p99_ms": pct(lat["stroke"], 99),
           "image_p50_ms": pct(lat["image"], 50), "image_p99_ms": pct(lat["image"], 99),
           "mean_batch": round(float(np.mean(batch)), 2)}
    for size in sorted(by_size):
        out["stroke_p50_ms_%d" % size] = pct(by_size[size], 50)
        out["stroke_p99_ms_%d" % size] = pct(by_size[size], 99)
    return out


def sustained(points, p99_ms=20.0, size=2048):
    """Highest offered rate whose strokes on `size` sessions kept p99 within
    p99_ms and whose stroke and new-image throughputs kept up with the offer
    (within 10%)."""
    best = None
    for p in points:
        p99 = p.get("stroke_p99_ms_%d" % size)
        if p99 is not None and p99 <= p99_ms and \
                p["stroke_throughput"] >= 0.9 * p["stroke_rate"] and \
                p.get("image_throughput", 0) >= 0.9 * p.get("image_rate", 0):
            if best is None or p["rate"] > best["rate"]:
                best = p
    return best


def find_sustained(probe, lo, hi, tol=0.02, p99_ms=20.0, size=2048, expand=3):
    """Bisects the offered rate for the highest one that passes `sustained`'s
    rule: probe(rate) -> run_trace statistics.  Starts from [lo, hi]: lo is
    halved while it fails and hi doubled while it passes (at most `expand`
    times each), then the bracket is halved until (hi - lo) <= tol * lo.
    Returns (points probed, the best passing point or None)."""
    points = []

    def ok(rate):
        pt = probe(rate)
        points.append(pt)
        return sustained([pt], p99_ms, size) is not None

    lo_ok = ok(lo)
    tries = 0
    while not lo_ok and tries < expand:
        hi, lo = lo, lo / 2.0
        lo_ok = ok(lo)
        tries += 1
    if not lo_ok:
        return points, None
    hi_fails = tries > 0  # hi was probed and failed
    tries = 0
    while not hi_fails and tries < expand:
        if ok(hi):
            lo, hi = hi, hi * 2.0
            tries += 1
        else:
            hi_fails = True
    if not hi_fails:
        return points, sustained(points, p99_ms, size)
    while hi - lo > tol * lo:
        mid = round((lo + hi) / 2.0)
        if ok(mid):
            lo = mid
        else:
            hi = mid
    return points, sustained(points, p99_ms, size)


# The fixed trace whose every answer the benchmark checks: each session
# first gets its base image again (a known state), then this Poisson trace is
# submitted from one thread in time order, so each session's requests are
# served in trace order.
VERIFY = dict(rate=20000, duration=0.075, seed=99, image_frac=0.1)


def verification_trace(mix, rgb=True):
    """The verification trace's requests in time order (t_ns from 0), with
    the per-session stroke alternation restarted."""
    mix.painted[:] = 0
    strokes, images = make_trace(mix, VERIFY["rate"], VERIFY["duration"], VERIFY["seed"], 0,
                                 VERIFY["image_frac"], rgb)
    return sorted(strokes + images, key=lambda x: x[0])


def mask_sha(words):
    """First 16 hex digits of the sha256 of a packed mask (uint32 words,
    little-endian)."""
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(words, "<u4").tobytes()).hexdigest()[:16]


def pack_mask(mask):
    """uint8 [H][W] labels -> packed uint32 words (bit p % 32 of word p // 32)."""
    b = np.packbits(np.asarray(mask, np.uint8).reshape(-1), bitorder="little")
    return np.pad(b, (0, (-b.size) % 4)).view(np.uint32)
