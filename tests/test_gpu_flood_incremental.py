"""The minimal source set after warm reflows, where the flood recomputes
only the tiles that changed and the tiles whose reach came through them:
every answer equals a from-scratch oracle cut.  The cases make reach travel
far from its excess -- through pixels with no excess of their own, across
many tiles -- and then cut it off, move it, or give it back."""
import numpy as np
import pytest

import gcmi_workloads as G
import oracle as O
from GraphCut import GraphCut
from GraphCutBatch import GraphCutBatch

pytestmark = pytest.mark.gpu


def corridor(W, H, seed_xy, w=5):
    """No unaries but one seed of excess; strong pairwise arcs everywhere,
    so the seed's leftover excess reaches the whole image."""
    z = np.zeros((H, W), np.int32)
    src, snk = z.copy(), z.copy()
    x, y = seed_xy
    src[y, x] = 1000
    snk[H - 1, W - 1] = 1
    right = np.full((H, W), w, np.int32)
    down = np.full((H, W), w, np.int32)
    right[:, -1] = 0
    down[-1, :] = 0
    return G.Workload(None, src, snk, right, down)


@pytest.mark.parametrize("tile,device_rounds", [(16, 0), (32, 0), (64, 0), (16, 6)])
def test_reach_through_many_tiles_toggled(tile, device_rounds):
    W, H = 300, 170
    w = corridor(W, H, (3, 4))
    gc = GraphCut(None, None, (W, H), tile=tile)
    if device_rounds:
        gc.set_device_rounds(device_rounds)
    gc.upload(*w.arrays())
    fo, mo = O.bk_cut(*w.arrays())
    assert gc.cut() == fo and (gc.labels() == mo).all()
    assert mo.sum() > W * H // 2  # the reach spans the image
    cur = w.copy()
    moves = [((3, 4), 0, 0), ((150, 80), 1000, 0), ((150, 80), 0, 0), ((299, 0), 7, 0),
             ((10, 160), 500, 0), ((299, 0), 0, 3), ((10, 160), 0, 0), ((3, 4), 1000, 0)]
    for k, ((x, y), s, t) in enumerate(moves):
        i = y * W + x
        cur.source.reshape(-1)[i] = s
        cur.sink.reshape(-1)[i] = t
        gc.edit(np.array([i], np.int32), np.array([s], np.int32), np.array([t], np.int32))
        fo, mo = O.bk_cut(*cur.arrays())
        assert gc.cut() == fo, (tile, k)
        assert (gc.labels() == mo).all(), (tile, k)


def test_wall_cuts_reach_off_and_back():
    """A hard sink wall across the image cuts the far side's reach off; a
    gap in it gives the reach back."""
    W, H = 260, 200
    w = corridor(W, H, (5, 5), w=9)
    gc = GraphCut(None, None, (W, H), tile=32)
    gc.set_local_relabel(True)
    gc.upload(*w.arrays())
    gc.cut()
    cur = w.copy()
    wall = np.array([y * W + 130 for y in range(H)], np.int32)
    steps = [(wall, 0, 1 << 30), (wall[90:110], 0, 0), (wall[90:110], 0, 1 << 30),
             (wall, 0, 0)]
    for k, (idx, s, t) in enumerate(steps):
        cur.source.reshape(-1)[idx] = s
        cur.sink.reshape(-1)[idx] = t
        gc.edit(idx, np.full(idx.size, s, np.int32), np.full(idx.size, t, np.int32))
        fo, mo = O.bk_cut(*cur.arrays())
        assert gc.cut() == fo, k
        assert (gc.labels() == mo).all(), k


@pytest.mark.parametrize("local", [False, True])
def test_random_soft_edits_batch(local):
    """Many small soft edits on random grids in a batch, some images left
    out of some runs (their edits pile up), reach from sparse excess."""
    rng = np.random.default_rng(11 + local)
    dims = [(97, 61), (130, 170), (64, 64), (200, 33)]
    works = []
    for i, (w, h) in enumerate(dims):
        x = G.tiny(w, h, 70 + i, cap_hi=40, w_hi=60)
        x.source[rng.random((h, w)) < 0.9] = 0  # sparse excess
        works.append(x)
    b = GraphCutBatch(dims, tile=16)
    b.set_local_relabel(local)
    for i, x in enumerate(works):
        b.upload(i, *x.arrays())
    b.cut()
    cur = [x.copy() for x in works]
    for r in range(8):
        which = sorted(rng.choice(len(dims), size=int(rng.integers(1, 5)), replace=False).tolist())
        for i in which:
            w, h = dims[i]
            n = int(rng.integers(1, 60))
            idx = rng.choice(w * h, size=n, replace=False).astype(np.int32)
            s = (rng.integers(0, 80, size=n) * (rng.random(n) < 0.3)).astype(np.int32)
            t = rng.integers(0, 80, size=n).astype(np.int32)
            cur[i].source.reshape(-1)[idx] = s
            cur[i].sink.reshape(-1)[idx] = t
            b.edit(i, idx, s, t)
        flows = b.cut(which=which)
        for k, i in enumerate(which):
            fo, mo = O.bk_cut(*cur[i].arrays())
            assert flows[k] == fo and (b.labels(i) == mo).all(), (r, i)
