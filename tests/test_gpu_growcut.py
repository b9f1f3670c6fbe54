"""GrowCut on the device against the synchronous float32 oracle: labels and
strengths equal after each of several step counts and at the fixed point,
for ragged image sizes, several weights and 3 or 4 channels; and the
visualisation filters (IncrementalTileList.execute, GraphCut.execute)."""
import numpy as np
import pytest

import gcmi_workloads as G
import growcut_oracle as GC
from GraphCut import GraphCut
from GrowCut import GrowCut
from IncrementalTileList import IncrementalTileList, Operator

pytestmark = pytest.mark.gpu


def seeds(W, H, seed):
    rng = np.random.default_rng(seed)
    st = G.strokes(W, H, count=4, seed=seed)
    return [(s.index, 1 + (k % 2)) for k, s in enumerate(st)] + \
        [(rng.integers(0, W * H, 20).astype(np.int32), 3)]


@pytest.mark.parametrize("W,H,kind,ch", [(150, 110, 0, 3), (64, 64, 1, 4), (301, 77, 4, 3),
                                         (33, 200, 2, 4), (97, 51, 3, 3)])
def test_growcut_equals_synchronous_oracle(W, H, kind, ch):
    img = G.three_ellipses(W, H, seed=W + H).image
    if ch == 4:
        img = np.concatenate([img, np.full((H, W, 1), 200, np.uint8)], axis=2)
    sd = seeds(W, H, W * 3 + H)
    gc = GrowCut(None, None, img, weight=kind)
    for idx, l in sd:
        gc.label(idx, label=l)
    done = 0
    for k in (1, 2, 5, 12):
        gc.evolve(k - done)
        done = k
        lab, st, _ = GC.run(img, sd, steps=k, kind=kind)
        assert (gc.labels() == lab).all(), k
        assert (gc.strength() == st).all(), k
    gc.evolve()
    assert gc.isComplete
    lab, st, changed = GC.run(img, sd, kind=kind)
    assert (gc.labels() == lab).all() and (gc.strength() == st).all()
    assert gc.iterations >= changed + 1
    # new seeds after completion evolve again from the current state
    extra = [(np.arange(0, W * H, 7, dtype=np.int32)[:50], 9)]
    gc.label(extra[0][0], label=9)
    gc.evolve()
    c = GC.colours(img)
    lab2, st2 = lab.copy(), st.copy()
    lab2.reshape(-1)[extra[0][0]] = 9
    st2.reshape(-1)[extra[0][0]] = 1.0
    while True:
        nl, ns = GC.step(c, lab2, st2, kind)
        if (nl == lab2).all() and (ns == st2).all():
            break
        lab2, st2 = nl, ns
    assert (gc.labels() == lab2).all() and (gc.strength() == st2).all()


def test_tile_list_filter_and_cut_fields():
    tl = IncrementalTileList(None, None, (200, 120), (16, 16))
    stamps = np.random.default_rng(5).integers(0, 4, (tl.dim[1], tl.dim[0])).astype(np.int32)
    tl.set_tiles(stamps)
    tl.iteration = 2
    img = tl.execute().get()
    # 200 x 120 in 16 x 16 tiles: 12 x 7 whole tiles, the rest has no tile
    ys, xs = np.mgrid[0:120, 0:200]
    inside = (xs < 192) & (ys < 112)
    hit = np.zeros((120, 200), bool)
    hit[inside] = stamps[ys[inside] // 16, xs[inside] // 16] == 2
    want = np.where(hit, 0xFF0000FF, 0xFF000000).astype(np.uint32)
    assert img.shape == (120, 200) and (img == want).all()
    w = G.centred(100, 70, seed=4)
    gc = GraphCut(None, None, (100, 70))
    gc.upload(*w.arrays())
    gc.cut()
    assert (gc.execute(None, "height") == gc.dHeight).all()
    e = gc.execute(None, "excess")
    assert e.shape == gc.dHeight.shape
    # the excess of pixels outside the minimal source set is <= 0
    lab = gc.labels()
    assert (e[:70, :100][lab == 0] <= 0).all()
