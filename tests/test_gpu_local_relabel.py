"""Warm reflow with the local relabel option (only the tiles that changed
are relabelled for the first relabels of a run):
after every edit the cut equals a from-scratch oracle cut, exactly as with
full relabels."""
import numpy as np
import pytest

import gcmi_workloads as G
import oracle as O
from GraphCut import GraphCut
from GraphCutBatch import GraphCutBatch
from GraphCutServer import GraphCutServer

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tile", [16, 32, 64])
def test_local_relabel_pixel_strokes(tile):
    W, H = 384, 256
    w = G.centred(W, H, seed=13)
    gc = GraphCut(None, None, (W, H), tile=tile)
    gc.set_local_relabel(True)
    gc.upload(*w.arrays())
    gc.cut()
    cur = w.copy()
    for k, st in enumerate(G.strokes(W, H, count=10, seed=15)):
        G.apply_stroke(cur, st)
        s, t = st.values()
        gc.edit(st.index, s, t)
        f = gc.cut()
        fo, mo = O.bk_cut(*cur.arrays())
        assert f == fo, (tile, k)
        assert (gc.labels() == mo).all(), (tile, k)


def test_local_relabel_soft_edits_and_random_grids():
    """Soft unary changes (not only hard constraints) on random grids."""
    rng = np.random.default_rng(5)
    for trial in range(4):
        W, H = [(97, 61), (200, 33), (64, 64), (130, 170)][trial]
        w = G.tiny(W, H, 40 + trial, cap_hi=300, w_hi=60)
        gc = GraphCut(None, None, (W, H), tile=16)
        gc.set_local_relabel(True)
        gc.upload(*w.arrays())
        gc.cut()
        cur = w.copy()
        for k in range(6):
            n = int(rng.integers(1, 200))
            idx = rng.choice(W * H, size=n, replace=False).astype(np.int32)
            s = rng.integers(0, 400, size=n).astype(np.int32)
            t = rng.integers(0, 400, size=n).astype(np.int32)
            cur.source.reshape(-1)[idx] = s
            cur.sink.reshape(-1)[idx] = t
            gc.edit(idx, s, t)
            f = gc.cut()
            fo, mo = O.bk_cut(*cur.arrays())
            assert f == fo, (trial, k)
            assert (gc.labels() == mo).all(), (trial, k)


def test_local_relabel_2048_brush():
    W = H = 2048
    work = G.centred(W, H, seed=3)
    strokes = G.strokes(W, H, count=6, seed=21)
    gc = GraphCut(None, None, (W, H), tile=32)
    gc.set_local_relabel(True)
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


def test_local_relabel_batch_and_server():
    dims = [(256, 256), (300, 200), (128, 128), (400, 100)]
    works = [G.three_ellipses(w, h, seed=30 + i) for i, (w, h) in enumerate(dims)]
    strokes = [G.strokes(w, h, count=3, seed=40 + i) for i, (w, h) in enumerate(dims)]
    b = GraphCutBatch(dims, tile=32)
    b.set_local_relabel(True)
    for i, w in enumerate(works):
        b.upload(i, *w.arrays())
    b.cut()
    cur = [w.copy() for w in works]
    for r in range(3):
        for i in range(len(dims)):
            st = strokes[i][r]
            s, t = st.capacities()
            b.paint(i, st.points, st.width, s, t)
            G.apply_stroke(cur[i], st)
        flows = b.cut()
        for i in range(len(dims)):
            fo, mo = O.bk_cut(*cur[i].arrays())
            assert flows[i] == fo and (b.labels(i) == mo).all(), (r, i)
    with GraphCutServer(tile=32, hbm_budget=1 << 30, max_sessions=8,
                        local_relabel=True) as srv:
        sids = [srv.open(w, h) for w, h in dims]
        trace = [(0, sids[i], "image", works[i].arrays()) for i in range(len(dims))]
        for r in range(3):
            for i in range(len(dims)):
                st = strokes[i][r]
                s, t = st.capacities()
                trace.append((0, sids[i], "polyline", (st.points, st.width, s, t)))
        ids = srv.replay(trace)
        for r in ids:
            srv.wait(int(r), timeout=60)
        for i, sid in enumerate(sids):
            mask, _ = srv.mask(sid)
            assert (mask == O.bk_cut(*cur[i].arrays())[1]).all(), i


@pytest.mark.parametrize("device_rounds", [0, 1, 2, 4])
def test_local_relabel_mixed_fresh_and_warm(device_rounds):
    """A run that cuts some images from new capacities and reflows strokes on
    the others, with local relabels (host- or device-driven).  With few
    device rounds the device-driven part of a solve often does not finish,
    and the host goes on from its state: the flood must then run once, over
    the finished solve."""
    dims = [(300, 220), (128, 300), (257, 129), (200, 200), (64, 500)]
    works = [G.three_ellipses(w, h, seed=60 + i) for i, (w, h) in enumerate(dims)]
    strokes = [G.strokes(w, h, count=4, seed=70 + i) for i, (w, h) in enumerate(dims)]
    b = GraphCutBatch(dims, tile=32)
    b.set_local_relabel(True)
    if device_rounds:
        b.set_device_rounds(device_rounds)
    for i, w in enumerate(works):
        b.upload(i, *w.arrays())
    b.cut()
    cur = [w.copy() for w in works]
    rng = np.random.default_rng(device_rounds)
    for r in range(4):
        for i in range(len(dims)):
            if (i + r) % 3 == 0:
                cur[i] = G.tiny(dims[i][0], dims[i][1], 500 + 10 * r + i, cap_hi=400, w_hi=50)
                b.upload(i, *cur[i].arrays())
            else:
                st = strokes[i][r]
                s, t = st.capacities()
                b.paint(i, st.points, st.width, s, t)
                G.apply_stroke(cur[i], st)
        which = sorted(rng.choice(len(dims), size=4, replace=False).tolist()) if r == 2 else None
        flows = b.cut(which=which)
        for k, i in enumerate(which if which is not None else range(len(dims))):
            fo, mo = O.bk_cut(*cur[i].arrays())
            assert flows[k] == fo and (b.labels(i) == mo).all(), (r, i)


@pytest.mark.parametrize("tile,device_rounds", [(16, 2), (16, 6), (32, 3)])
def test_far_excess_behind_zero_excess_corridor(tile, device_rounds):
    """A sink appears at one end of a long strip whose only excess sits at
    the other end, past pixels with no excess (nothing there pushes, so no
    tile between them turns dirty).  The relabel's distances must reach the
    far excess however few BFS rounds a device-driven relabel runs: a
    relabel whose queue did not drain hands its queued tiles to the next
    one."""
    W, H = 700, 40
    z = np.zeros((H, W), np.int32)
    src, snk = z.copy(), z.copy()
    src[:, W - 20:] = 9  # excess only at the far right
    right = np.full((H, W), 6, np.int32)
    down = np.full((H, W), 6, np.int32)
    right[:, -1] = 0
    down[-1, :] = 0
    w = G.Workload(None, src, snk, right, down)
    gc = GraphCut(None, None, (W, H), tile=tile)
    gc.set_local_relabel(True)
    gc.set_device_rounds(device_rounds)
    gc.upload(*w.arrays())
    fo, mo = O.bk_cut(*w.arrays())
    assert fo == 0 and gc.cut() == 0 and (gc.labels() == mo).all()
    cur = w.copy()
    # sinks at the far left, then also in the middle, then the left one goes
    for k, (x, t) in enumerate([(3, 1 << 20), (350, 50), (3, 0), (200, 7)]):
        idx = np.arange(H, dtype=np.int32) * W + x
        cur.sink.reshape(-1)[idx] = t
        gc.edit(idx, np.zeros(H, np.int32), np.full(H, t, np.int32))
        fo, mo = O.bk_cut(*cur.arrays())
        assert fo > 0
        assert gc.cut() == fo, (tile, device_rounds, k)
        assert (gc.labels() == mo).all(), (tile, device_rounds, k)


@pytest.mark.parametrize("seed", [21, 28, 31, 33, 35, 37])
def test_unfinished_device_run_floods_once(seed):
    """Batches mixing images cut from new capacities with reflowed strokes,
    with 1-3 device BFS rounds per relabel, so that the device-driven part of
    a solve often ends unfinished and the host goes on from its state.  The
    minimal source set must come from one flood over the finished solve: a
    flood over the unfinished state would leave reach the later incremental
    flood trusts.  (These seeds gave wrong masks, with the right flows, when
    the device-driven flood ran regardless of the done flag.)"""
    rng0 = np.random.default_rng(seed)
    n = 5
    dims = [tuple(int(v) for v in rng0.integers(64, 320, 2)) for _ in range(n)]
    dr = int(rng0.integers(1, 4))
    works = [G.three_ellipses(w, h, seed=1000 * seed + i) for i, (w, h) in enumerate(dims)]
    strokes = [G.strokes(w, h, count=3, seed=2000 * seed + i) for i, (w, h) in enumerate(dims)]
    b = GraphCutBatch(dims, tile=32)
    b.set_local_relabel(True)
    b.set_device_rounds(dr)
    for i, w in enumerate(works):
        b.upload(i, *w.arrays())
    b.cut()
    cur = [w.copy() for w in works]
    for r in range(3):
        for i in range(n):
            if (i + r + seed) % 3 == 0:
                cur[i] = G.tiny(dims[i][0], dims[i][1], 500 + 10 * r + i + 100 * seed, cap_hi=400,
                                w_hi=50)
                b.upload(i, *cur[i].arrays())
            else:
                st = strokes[i][r]
                s, t = st.capacities()
                b.paint(i, st.points, st.width, s, t)
                G.apply_stroke(cur[i], st)
        flows = b.cut()
        for i in range(n):
            fo, mo = O.bk_cut(*cur[i].arrays())
            assert flows[i] == fo and (b.labels(i) == mo).all(), (r, i)
