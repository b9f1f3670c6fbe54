"""Seeded synthetic workloads for the grid cut (host side, numpy).

Every workload is a set of int32 arrays, all shaped [H][W] and C-contiguous:

    source[y][x]  capacity s -> p  (cost of labelling p background)
    sink[y][x]    capacity p -> t  (cost of labelling p foreground)
    right[y][x]   w(p, p+1)  (last column is 0)
    down[y][x]    w(p, p+W)  (last row is 0)

Image model: background mean (60,80,100), foreground ellipses mean
(180,150,90), per-channel Gaussian noise, clipped and rounded to uint8.
Unaries: round(10*min(||I-mu_k||^2/100, 100)), toward the opposite class mean
(source capacity uses the background mean, sink capacity the foreground mean).
Pairwise: round(50*exp(-||Ip-Iq||^2/(2*beta))), beta = mean squared
neighbour difference over all horizontal and vertical pairs.

Pixel coordinates are the integer column/row indices; an ellipse with centre
(cx, cy) and semi-axes (ax, ay) covers ((x-cx)/ax)^2 + ((y-cy)/ay)^2 <= 1.
Rounding is numpy rint (half to even).

The reference computes its pairwise weights on the device from an RGBA image
(graphcut.cl:1018-1131) and takes the unaries as one float excess image
(test_GraphCut.py:44-47); here both are int32 and uploaded, so the device path
is pure integer.
"""
import numpy as np

MU_BG = np.array([60.0, 80.0, 100.0])
MU_FG = np.array([180.0, 150.0, 90.0])
HARD = 1 << 30


class Workload:
    """One cut's inputs plus the RGB image they came from."""

    def __init__(self, image, source, sink, right, down):
        self.image = image
        self.source = source
        self.sink = sink
        self.right = right
        self.down = down

    @property
    def shape(self):
        return self.source.shape

    @property
    def pixels(self):
        return int(self.source.size)

    def arrays(self):
        return self.source, self.sink, self.right, self.down

    def copy(self):
        return Workload(self.image, self.source.copy(), self.sink.copy(),
                        self.right, self.down)


def make_image(rng, W, H, ellipses, sigma):
    """uint8 RGB [H][W][3] with the given (cx, cy, ax, ay) ellipses."""
    ys = np.arange(H, dtype=np.float64)[:, None]
    xs = np.arange(W, dtype=np.float64)[None, :]
    fg = np.zeros((H, W), dtype=bool)
    for cx, cy, ax, ay in ellipses:
        fg |= ((xs - cx) / ax) ** 2 + ((ys - cy) / ay) ** 2 <= 1.0
    mean = np.where(fg[..., None], MU_FG, MU_BG)
    noise = rng.normal(0.0, sigma, size=(H, W, 3))
    return np.clip(np.rint(mean + noise), 0, 255).astype(np.uint8)


def unaries(image):
    img = image.astype(np.float64)
    dbg = ((img - MU_BG) ** 2).sum(-1)
    dfg = ((img - MU_FG) ** 2).sum(-1)
    source = np.rint(10.0 * np.minimum(dbg / 100.0, 100.0)).astype(np.int32)
    sink = np.rint(10.0 * np.minimum(dfg / 100.0, 100.0)).astype(np.int32)
    return source, sink


def pairwise(image):
    img = image.astype(np.float64)
    H, W, _ = img.shape
    dh = ((img[:, 1:] - img[:, :-1]) ** 2).sum(-1)
    dv = ((img[1:, :] - img[:-1, :]) ** 2).sum(-1)
    count = dh.size + dv.size
    beta = (dh.sum() + dv.sum()) / count if count else 1.0
    if beta <= 0:
        beta = 1.0
    right = np.zeros((H, W), dtype=np.int32)
    down = np.zeros((H, W), dtype=np.int32)
    right[:, :-1] = np.rint(50.0 * np.exp(-dh / (2.0 * beta))).astype(np.int32)
    down[:-1, :] = np.rint(50.0 * np.exp(-dv / (2.0 * beta))).astype(np.int32)
    return right, down


def caps_from_rgb(image, mu_bg=MU_BG, mu_fg=MU_FG, unary_scale=10.0, unary_div=100.0,
                  unary_cap=100.0, pair="exp", pair_scale=50.0, inv_beta=0.1, inv_eps=0.05):
    """Capacities of a uint8 [H][W][3 or 4] colour image under a colour
    model (the host statement of gcmi_rgb_model; float64, each operation
    rounded once): unaries rint(unary_scale * min(|I - mu|^2 / unary_div,
    unary_cap)); pairwise "exp": rint(pair_scale * exp(-d2 / (2 beta))),
    beta = mean d2 (1 if 0); "inv": rint(pair_scale / (inv_beta * sqrt(d2) +
    inv_eps)), the reference's weight (graphcut.cl:1018-1031)."""
    img = image[..., :3].astype(np.float64)
    H, W, _ = img.shape

    def unary(mu):
        e = img - np.asarray(mu, dtype=np.float64)
        d = (e[..., 0] * e[..., 0] + e[..., 1] * e[..., 1]) + e[..., 2] * e[..., 2]
        return np.rint(unary_scale * np.minimum(d / unary_div, unary_cap)).astype(np.int32)

    dh = ((img[:, 1:] - img[:, :-1]) ** 2).sum(-1)
    dv = ((img[1:, :] - img[:-1, :]) ** 2).sum(-1)
    if pair == "exp":
        count = dh.size + dv.size
        beta = (dh.sum() + dv.sum()) / count if count else 1.0
        if beta <= 0:
            beta = 1.0
        weight = lambda d: np.rint(pair_scale * np.exp(-d / (2.0 * beta)))
    elif pair == "inv":
        weight = lambda d: np.rint(pair_scale / (inv_beta * np.sqrt(d) + inv_eps))
    else:
        raise ValueError(pair)
    right = np.zeros((H, W), dtype=np.int32)
    down = np.zeros((H, W), dtype=np.int32)
    right[:, :-1] = weight(dh).astype(np.int32)
    down[:-1, :] = weight(dv).astype(np.int32)
    return Workload(image, unary(mu_bg), unary(mu_fg), right, down)


def from_image(image):
    source, sink = unaries(image)
    right, down = pairwise(image)
    return Workload(image, source, sink, right, down)


def centred(W, H, seed=0, sigma=20.0):
    """One centred ellipse, semi-axes (0.3W, 0.2H).  The 512x512 single cut
    (seed 0) and the 2048x2048 warm-start image (seed 3) use this."""
    rng = np.random.default_rng(seed)
    ell = [((W - 1) / 2.0, (H - 1) / 2.0, 0.3 * W, 0.2 * H)]
    return from_image(make_image(rng, W, H, ell, sigma))


def three_ellipses(W, H, seed=1, sigma=30.0):
    """Three ellipses, centres U[0.1,0.9]*(W,H), semi-axes U[0.05,0.25]*(W,H)
    (the 4096x4096 single cut, seed 1)."""
    rng = np.random.default_rng(seed)
    ell = []
    for _ in range(3):
        cx, cy = rng.uniform(0.1, 0.9) * W, rng.uniform(0.1, 0.9) * H
        ax, ay = rng.uniform(0.05, 0.25) * W, rng.uniform(0.05, 0.25) * H
        ell.append((cx, cy, ax, ay))
    return from_image(make_image(rng, W, H, ell, sigma))


def ragged_batch(count=64, lo=256, hi=2048, seed=2, sigma=20.0):
    """`count` images with W, H each drawn from U[lo, hi]; every image is a
    centred ellipse with its own seed (seed*1000 + i)."""
    rng = np.random.default_rng(seed)
    dims = rng.integers(lo, hi + 1, size=(count, 2))
    return [centred(int(w), int(h), seed=seed * 1000 + i, sigma=sigma)
            for i, (w, h) in enumerate(dims)]


def tiny(W, H, seed, cap_hi=20, w_hi=10):
    """Small random integer instance (no image) for exhaustive checks."""
    rng = np.random.default_rng(seed)
    source = rng.integers(0, cap_hi + 1, size=(H, W)).astype(np.int32)
    sink = rng.integers(0, cap_hi + 1, size=(H, W)).astype(np.int32)
    right = rng.integers(0, w_hi + 1, size=(H, W)).astype(np.int32)
    down = rng.integers(0, w_hi + 1, size=(H, W)).astype(np.int32)
    right[:, -1] = 0
    down[-1, :] = 0
    return Workload(None, source, sink, right, down)


# ---------------------------------------------------------------- strokes


class Stroke:
    """A brush stroke: flat pixel indices it covers and its side, plus the
    polyline (points, width) it was rasterised from, if any."""

    def __init__(self, index, foreground, points=None, width=None):
        self.index = index            # int32 flat indices y*W + x, ascending
        self.foreground = foreground  # True: hard source; False: hard sink
        self.points = points          # float64 [npts][2] (x, y)
        self.width = width

    def capacities(self):
        """(source, sink) capacity painted by the stroke."""
        return (HARD, 0) if self.foreground else (0, HARD)

    def values(self):
        n = self.index.size
        if self.foreground:
            return (np.full(n, HARD, np.int32), np.zeros(n, np.int32))
        return (np.zeros(n, np.int32), np.full(n, HARD, np.int32))


def _capsule(xs, ys, x0, y0, dx, dy, L2, r):
    """The capsule test of one segment at pixel centres (xs, ys), float64,
    each operation rounded once in this order."""
    if L2 > 0:
        t = np.clip(((xs - x0) * dx + (ys - y0) * dy) / L2, 0.0, 1.0)
    else:
        t = np.zeros(np.broadcast(xs, ys).shape)
    px, py = x0 + t * dx, y0 + t * dy
    return (xs - px) ** 2 + (ys - py) ** 2 <= r * r


def _segment_box(W, H, x0, y0, x1, y1, r):
    xa = max(int(np.floor(min(x0, x1) - r)), 0)
    xb = min(int(np.ceil(max(x0, x1) + r)), W - 1)
    ya = max(int(np.floor(min(y0, y1) - r)), 0)
    yb = min(int(np.ceil(max(y0, y1) + r)), H - 1)
    return xa, xb, ya, yb


def rasterise_polyline_box(W, H, pts, width):
    """Flat indices of pixels whose centre lies within width/2 of the
    polyline (a union of capsules): every pixel of each segment's bounding
    box tested.  The plain statement that rasterise_polyline must equal."""
    r = width / 2.0
    hit = np.zeros((H, W), dtype=bool)
    for (x0, y0), (x1, y1) in zip(pts[:-1], pts[1:]):
        xa, xb, ya, yb = _segment_box(W, H, x0, y0, x1, y1, r)
        if xa > xb or ya > yb:
            continue
        ys = np.arange(ya, yb + 1, dtype=np.float64)[:, None]
        xs = np.arange(xa, xb + 1, dtype=np.float64)[None, :]
        dx, dy = x1 - x0, y1 - y0
        hit[ya:yb + 1, xa:xb + 1] |= _capsule(xs, ys, x0, y0, dx, dy, dx * dx + dy * dy, r)
    return np.flatnonzero(hit).astype(np.int32)


def rasterise_polyline(W, H, pts, width):
    """The pixel set of rasterise_polyline_box, testing only a band around
    each segment: in row y the capsule lies within r of the segment points
    whose y is within r of y, so its columns lie within r of those points'
    x range (computed with one pixel of margin).  The same test on a
    superset of the hit pixels gives the same set."""
    r = width / 2.0
    hit = np.zeros((H, W), dtype=bool)
    for (x0, y0), (x1, y1) in zip(pts[:-1], pts[1:]):
        xa, xb, ya, yb = _segment_box(W, H, x0, y0, x1, y1, r)
        if xa > xb or ya > yb:
            continue
        dx, dy = x1 - x0, y1 - y0
        rows = np.arange(ya, yb + 1, dtype=np.float64)
        m = r + 1.0
        if abs(dy) > 1e-9:
            ta = (rows - m - y0) / dy
            tb = (rows + m - y0) / dy
            tlo = np.clip(np.minimum(ta, tb), 0.0, 1.0)
            thi = np.clip(np.maximum(ta, tb), 0.0, 1.0)
            xl = x0 + np.minimum(tlo * dx, thi * dx)
            xh = x0 + np.maximum(tlo * dx, thi * dx)
            lo = np.maximum(np.floor(xl - m), xa).astype(np.int64)
            hi = np.minimum(np.ceil(xh + m), xb).astype(np.int64)
        else:
            lo = np.full(rows.size, xa, np.int64)
            hi = np.full(rows.size, xb, np.int64)
        keep = hi >= lo
        if not keep.any():
            continue
        rows, lo, hi = rows[keep], lo[keep], hi[keep]
        span = int((hi - lo).max()) + 1
        cols = lo[:, None] + np.arange(span)[None, :]
        inside = cols <= hi[:, None]
        xs = cols.astype(np.float64)
        ys = rows[:, None]
        h = _capsule(xs, ys, x0, y0, dx, dy, dx * dx + dy * dy, r) & inside
        yy, kk = np.nonzero(h)
        hit[rows[yy].astype(np.int64), cols[yy, kk]] = True
    return np.flatnonzero(hit).astype(np.int32)


def random_stroke(rng, W, H, foreground, raster=False):
    """One stroke drawn from `rng`: a polyline of 3-8 segments with vertices
    uniform in the image and width 8-32 px.  With raster=False the pixel set
    is not computed (index is None)."""
    nseg = int(rng.integers(3, 9))
    pts = list(zip(rng.uniform(0, W - 1, nseg + 1),
                   rng.uniform(0, H - 1, nseg + 1)))
    width = int(rng.integers(8, 33))
    idx = rasterise_polyline(W, H, pts, width) if raster else None
    return Stroke(idx, foreground, np.array(pts, dtype=np.float64), width)


def strokes(W, H, count=100, seed=3, first_foreground=True, raster=True):
    """`count` seeded strokes (random_stroke), alternating
    foreground/background."""
    rng = np.random.default_rng(seed + 7919)
    return [random_stroke(rng, W, H, (i % 2 == 0) == first_foreground, raster)
            for i in range(count)]


def rasterise(stroke, W, H):
    """The stroke with its pixel set computed."""
    if stroke.index is None:
        stroke.index = rasterise_polyline(W, H, stroke.points, stroke.width)
    return stroke


def apply_stroke(work, stroke):
    """Apply a stroke's hard constraints to a workload in place."""
    s, t = stroke.values()
    work.source.reshape(-1)[stroke.index] = s
    work.sink.reshape(-1)[stroke.index] = t
