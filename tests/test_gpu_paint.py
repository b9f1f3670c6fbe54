"""Brush strokes rasterised on the device (paint): the painted pixel set is
the host rasteriser's (a union of capsules tested in float64), checked
through the cut -- a stroke paints hard constraints, so one pixel more or
less changes the flow or the minimal source set -- against a from-scratch
oracle cut of the capacities with the host-rasterised stroke applied."""
import numpy as np
import pytest

import gcmi_workloads as G
import oracle as O
from GraphCut import GraphCut
from GraphCutBatch import GraphCutBatch
from GraphCutServer import GraphCutServer

pytestmark = pytest.mark.gpu


def apply_polyline(work, pts, width, fg):
    idx = G.rasterise_polyline(work.shape[1], work.shape[0], pts, width)
    G.apply_stroke(work, G.Stroke(idx, fg))
    return idx


def edge_strokes(W, H):
    """Polylines that probe the rasteriser's corners: ties at exactly r,
    zero width, repeated points, points outside the image, one point."""
    return [
        ([(2.0, 3.5), (W - 3.0, 3.5)], 1, True),            # d = 0.5 = r on rows 3 and 4
        ([(0.0, 0.0), (W - 1.0, H - 1.0)], 0, False),         # width 0: centres on the line
        ([(5.0, 5.0), (5.0, 5.0), (9.25, 7.75)], 3, True),   # zero-length segment
        ([(-20.0, H / 2.0), (W + 20.0, H / 3.0)], 6, False),  # leaves the image
        ([(W / 2.0, H / 2.0)], 9, True),                      # one point: paints nothing
        ([(1.5, 1.5), (1.5, H - 1.5), (W - 1.5, H - 1.5)], 2.0, True),
        ([(-50.0, -50.0), (-10.0, -40.0)], 8, True),          # wholly outside
        ([(0.1 * W, 0.3 * H), (0.7 * W, 0.9 * H), (0.2 * W, 0.6 * H), (0.9 * W, 0.1 * H)],
         17, False),
    ]


@pytest.mark.parametrize("tile", [16, 32, 64])
def test_paint_edge_strokes_warm(tile):
    W, H = 173, 141
    work = G.centred(W, H, seed=41)
    gc = GraphCut(None, None, (W, H), tile=(tile, tile))
    gc.upload(*work.arrays())
    gc.cut()
    for k, (pts, width, fg) in enumerate(edge_strokes(W, H)):
        s, t = (G.HARD, 0) if fg else (0, G.HARD)
        gc.paint(pts, width, s, t)
        apply_polyline(work, pts, width, fg)
        f = gc.cut()
        fo, mo = O.bk_cut(*work.arrays())
        assert f == fo, (tile, k)
        assert (gc.labels() == mo).all(), (tile, k)


def test_paint_before_first_cut_and_soft_values():
    W, H = 300, 200
    work = G.three_ellipses(W, H, seed=43)
    gc = GraphCut(None, None, (W, H))
    gc.upload(*work.arrays())
    pts = [(10.0, 20.0), (250.5, 150.25), (40.0, 180.0)]
    gc.paint(pts, 11, 123, 7)  # fresh: only rewrites the unaries
    idx = G.rasterise_polyline(W, H, pts, 11)
    work.source.reshape(-1)[idx] = 123
    work.sink.reshape(-1)[idx] = 7
    assert gc.cut() == O.bk_cut(*work.arrays())[0]
    assert (gc.labels() == O.bk_cut(*work.arrays())[1]).all()


def wide_strokes(W, H):
    """Polylines whose rows span several 64-column passes and several
    256-column pieces, and whose segments overlap the earlier ones."""
    return [
        ([(3.0, 10.0), (W - 4.0, 12.5)], 9, True),                        # near-horizontal, full width
        ([(W - 2.0, 40.0), (5.5, 48.25), (W - 30.0, 55.0)], 21, False),   # doubles back over itself
        ([(10.0, 5.0), (W / 2.0, H - 5.0), (W - 10.0, 5.0), (12.0, 6.0)], 31, True),
        ([(0.0, H / 2.0), (W - 1.0, H / 2.0 + 0.5)], 64, False),          # 64 wide, rows of W
        ([(W / 3.0, 0.0), (W / 3.0 + 0.25, H - 1.0)], 3, True),            # steep: one pass per row
        ([(-30.0, -8.0), (W + 30.0, H + 8.0)], 12, False),                # clipped at both ends
    ]


@pytest.mark.parametrize("tile", [16, 32, 64])
def test_paint_pixel_set_direct(tile):
    """The painted pixels read back from the device's unaries equal the host
    rasteriser's, pixel for pixel (an image not cut yet: painting only
    rewrites its unaries), for strokes wider than a pass and a piece."""
    W, H = 1100, 90
    work = G.centred(W, H, seed=47)
    gc = GraphCut(None, None, (W, H), tile=(tile, tile))
    for k, (pts, width, fg) in enumerate(wide_strokes(W, H)):
        gc.upload(*work.arrays())
        s, t = (1000 + k, 3 + k) if fg else (5 + k, 2000 + k)
        gc.paint(pts, width, s, t)
        src, snk, right, down = gc.capacities()
        es, et = work.source.copy(), work.sink.copy()
        idx = G.rasterise_polyline(W, H, pts, width)
        es.reshape(-1)[idx] = s
        et.reshape(-1)[idx] = t
        assert idx.size > 0, k
        assert (src == es).all() and (snk == et).all(), (tile, k, int((src != es).sum()))
        assert (right == work.right).all() and (down == work.down).all(), (tile, k)
    gc.close()


@pytest.mark.parametrize("tile", [32, 64])
def test_paint_strokes_2048(tile):
    """The 2048x2048 warm-start workload painted on the device: after
    every stroke the cut equals the oracle's."""
    W = H = 2048
    work = G.centred(W, H, seed=3)
    strokes = G.strokes(W, H, count=6, seed=3)
    gc = GraphCut(None, None, (W, H), tile=(tile, tile))
    gc.upload(*work.arrays())
    gc.cut()
    for k, st in enumerate(strokes):
        s, t = st.capacities()
        gc.paint(st.points, st.width, s, t)
        G.apply_stroke(work, st)
        f = gc.cut()
        fo, mo = O.bk_cut(*work.arrays())
        assert f == fo, k
        assert (gc.labels() == mo).all(), k


def test_batch_paint():
    dims = [(200, 120), (64, 300), (333, 77)]
    works = [G.centred(w, h, seed=60 + i) for i, (w, h) in enumerate(dims)]
    b = GraphCutBatch(dims, tile=32)
    for i, w in enumerate(works):
        b.upload(i, *w.arrays())
    b.cut()
    for i, (w, h) in enumerate(dims):
        st = G.strokes(w, h, count=1, seed=70 + i)[0]
        s, t = st.capacities()
        b.paint(i, st.points, st.width, s, t)
        G.apply_stroke(works[i], st)
    flows = b.cut()
    for i, w in enumerate(works):
        fo, mo = O.bk_cut(*w.arrays())
        assert flows[i] == fo and (b.labels(i) == mo).all(), i


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_server_replay_polylines(mode):
    """A replayed trace of images and polyline strokes over several
    sessions: each session's final mask and every request's flow equal the
    oracle's."""
    dims = [(256, 256), (300, 200), (128, 128), (400, 100)]
    works = [G.centred(w, h, seed=80 + i) for i, (w, h) in enumerate(dims)]
    strokes = [G.strokes(w, h, count=4, seed=90 + i) for i, (w, h) in enumerate(dims)]
    with GraphCutServer(tile=32, hbm_budget=1 << 30, max_sessions=8, mode=mode) as srv:
        sids = [srv.open(w, h) for w, h in dims]
        trace = [(0, sids[i], "image", works[i].arrays()) for i in range(len(dims))]
        for r in range(4):
            for i in range(len(dims)):
                st = strokes[i][r]
                s, t = st.capacities()
                trace.append((0, sids[i], "polyline", (st.points, st.width, s, t)))
        ids = srv.replay(trace)
        infos = [srv.wait(int(r), timeout=60) for r in ids]
        cur = [w.copy() for w in works]
        for j, (_, sid, kind, _) in enumerate(trace):
            i = sids.index(sid)
            if kind == "polyline":
                G.apply_stroke(cur[i], strokes[i][(j - len(dims)) // len(dims)])
            assert infos[j].flow == O.bk_cut(*cur[i].arrays())[0], (mode, j)
        for i, sid in enumerate(sids):
            mask, rid = srv.mask(sid)
            assert rid == ids[-len(dims) + i]
            assert (mask == O.bk_cut(*cur[i].arrays())[1]).all(), (mode, i)


def test_packed_labels_and_forget():
    """labels_packed() is labels() packed little-endian into uint32 words;
    forget() makes the next cut start from the capacities (same answer)."""
    W, H = 333, 97  # W*H not a multiple of 32
    work = G.centred(W, H, seed=44)
    gc = GraphCut(None, None, (W, H))
    gc.upload(*work.arrays())
    f0 = gc.cut()
    bits = gc.labels_packed()
    lab = gc.labels()
    ref = np.packbits(lab.reshape(-1), bitorder="little")
    ref = np.pad(ref, (0, (-ref.size) % 4)).view(np.uint32)
    assert bits.shape == ((W * H + 31) // 32,)
    assert (bits == ref).all()
    st = G.strokes(W, H, count=1, seed=45)[0]
    s, t = st.capacities()
    gc.paint(st.points, st.width, s, t)
    gc.forget()
    f1 = gc.cut()
    G.apply_stroke(work, st)
    fo, mo = O.bk_cut(*work.arrays())
    assert f1 == fo and f0 != f1
    assert (gc.labels() == mo).all()
