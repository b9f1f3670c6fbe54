"""Batched cuts and warm-start reflow against the oracle: every image of a
batch, and the cut after every brush stroke, must equal a from-scratch
oracle cut of the same capacities (flow value and minimal source set)."""
import numpy as np
import pytest

import gcmi_lib as L
import gcmi_workloads as G
import oracle as O
from GraphCut import GraphCut
from GraphCutBatch import GraphCutBatch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tile", [16, 32, 64])
def test_ragged_batch(tile):
    rng = np.random.default_rng(tile)
    dims = [(int(w), int(h)) for w, h in rng.integers(20, 400, size=(9, 2))] + [(1, 1), (3, 500)]
    works = [G.tiny(w, h, 100 + i, cap_hi=300, w_hi=60) if i % 2 else
             G.centred(w, h, seed=200 + i) for i, (w, h) in enumerate(dims)]
    b = GraphCutBatch(dims, tile=tile)
    for i, w in enumerate(works):
        b.upload(i, *w.arrays())
    flows = b.cut()
    for i, w in enumerate(works):
        fo, mo = O.bk_cut(*w.arrays())
        assert flows[i] == fo, i
        assert (b.labels(i) == mo).all(), i
    # a subset re-solves warm with the same answers; a repeat is identical
    sub = b.cut(which=[3, 0])
    assert sub == [flows[3], flows[0]]
    assert b.cut() == flows


def test_batch_of_tiny_vs_brute_force():
    dims = [(3, 3), (4, 4), (1, 7), (5, 3), (2, 2), (4, 3)] * 3
    works = [G.tiny(w, h, 7 * i, cap_hi=[5, 20, 50][i % 3], w_hi=[3, 10, 30][i % 3])
             for i, (w, h) in enumerate(dims)]
    b = GraphCutBatch(dims, tile=16)
    for i, w in enumerate(works):
        b.upload(i, *w.arrays())
    flows = b.cut()
    for i, w in enumerate(works):
        fb, mb = O.brute_force(*w.arrays())
        assert flows[i] == fb and (b.labels(i) == mb).all(), i


def test_batch_matches_plain():
    w = G.three_ellipses(700, 500, seed=9)
    gc = GraphCut(None, None, (700, 500))
    gc.upload(*w.arrays())
    f = gc.cut()
    b = GraphCutBatch([(700, 500)])
    b.upload(0, *w.arrays())
    assert b.cut() == [f]
    assert (b.labels(0) == gc.labels()).all()


@pytest.mark.parametrize("tile", [16, 32, 64])
def test_warm_start_strokes(tile):
    W, H = 384, 256
    w = G.centred(W, H, seed=3)
    gc = GraphCut(None, None, (W, H), tile=tile)
    gc.upload(*w.arrays())
    gc.cut()
    cur = w.copy()
    for k, s in enumerate(G.strokes(W, H, count=12, seed=5)):
        G.apply_stroke(cur, s)
        src, snk = s.values()
        gc.edit(s.index, src, snk)
        flow = gc.cut()
        fo, mo = O.bk_cut(*cur.arrays())
        assert flow == fo, k
        assert (gc.labels() == mo).all(), k


def test_warm_start_soft_edits():
    # edits that lower and raise terminal capacities (flow must be returned)
    W, H = 200, 150
    w = G.three_ellipses(W, H, seed=11)
    gc = GraphCut(None, None, (W, H))
    gc.upload(*w.arrays())
    gc.cut()
    cur = w.copy()
    rng = np.random.default_rng(1)
    for k in range(10):
        idx = rng.choice(W * H, size=500, replace=False).astype(np.int32)
        s = rng.integers(0, 1200, idx.size).astype(np.int32)
        t = rng.integers(0, 1200, idx.size).astype(np.int32)
        cur.source.reshape(-1)[idx] = s
        cur.sink.reshape(-1)[idx] = t
        gc.edit(idx, s, t)
        flow = gc.cut()
        fo, mo = O.bk_cut(*cur.arrays())
        assert flow == fo and (gc.labels() == mo).all(), k


def test_warm_start_2048_strokes():
    W = H = 2048
    w = G.centred(W, H, seed=3)
    gc = GraphCut(None, None, (W, H))
    gc.upload(*w.arrays())
    gc.cut()
    cur = w.copy()
    for k, s in enumerate(G.strokes(W, H, count=6, seed=3)):
        G.apply_stroke(cur, s)
        gc.edit(s.index, *s.values())
        flow = gc.cut()
        fo, mo = O.bk_cut(*cur.arrays())
        assert flow == fo and (gc.labels() == mo).all(), k


def test_batch_warm_mix():
    # some images edited (warm), one re-uploaded (fresh), solved together
    dims = [(300, 200), (128, 512), (257, 129)]
    works = [G.centred(w, h, seed=40 + i) for i, (w, h) in enumerate(dims)]
    b = GraphCutBatch(dims)
    for i, w in enumerate(works):
        b.upload(i, *w.arrays())
    b.cut()
    s0 = G.strokes(300, 200, count=1, seed=1)[0]
    G.apply_stroke(works[0], s0)
    b.edit(0, s0.index, *s0.values())
    works[2] = G.three_ellipses(257, 129, seed=99)
    b.upload(2, *works[2].arrays())
    flows = b.cut()
    for i, w in enumerate(works):
        fo, mo = O.bk_cut(*w.arrays())
        assert flows[i] == fo and (b.labels(i) == mo).all(), i


def test_edit_validation():
    gc = GraphCut(None, None, (16, 16))
    z = np.zeros((16, 16), np.int32)
    with pytest.raises(L.GcmiError):
        gc.edit([0], [1], [1])  # before upload
    gc.upload(z, z, z, z)
    with pytest.raises(L.GcmiError):
        gc.edit([3, 3], [1, 2], [0, 0])  # duplicate index
    with pytest.raises(L.GcmiError):
        gc.edit([256], [1], [0])  # out of range
    with pytest.raises(L.GcmiError):
        gc.edit([0], [-1], [0])
    gc.edit([5], [7], [0])  # before the first cut: rewrites capacities only
    assert gc.cut() == 0
    assert gc.labels().reshape(-1)[5] == 1
