"""The reference's GraphCut step hooks (GraphCut.pyx:187-346) and its
constructor with lamda (GraphCut.pyx:80), against the oracle: a solve driven
by hand with push / relabel / startBfs / isCompleted, in the reference's own
loop order (GraphCut.pyx:348-418), gives the same flow and minimal source set
as cut() and Boykov-Kolmogorov."""
import numpy as np
import pytest

import gcmi_workloads as G
import oracle as O
from Buffer2D import Buffer2D
from GraphCut import GraphCut

pytestmark = pytest.mark.gpu


def reference_loop(gc, bfs=5, dExcess=None, cap=20000):
    """The driver of GraphCut.pyx:378-418: startBfs, then push; every bfs-th
    round startBfs + isCompleted, else relabel."""
    gc.startBfs(dExcess)
    it = 1
    while it < cap:
        gc.push(dExcess)
        if it % bfs == 0:
            gc.startBfs(dExcess)
            if gc.isCompleted(dExcess):
                return it
        else:
            gc.relabel(dExcess)
        it += 1
    raise AssertionError("did not complete")


CASES = [("tiny", 7, 5), ("tiny", 1, 300), ("centred", 300, 200), ("ellipses", 257, 255),
         ("tiny", 64, 64)]


@pytest.mark.parametrize("tile", [16, 32, 64])
@pytest.mark.parametrize("case", CASES)
def test_hooks_in_reference_order(case, tile):
    kind, W, H = case
    w = {"tiny": lambda: G.tiny(W, H, W * 7 + H, cap_hi=300, w_hi=60),
         "centred": lambda: G.centred(W, H, seed=5),
         "ellipses": lambda: G.three_ellipses(W, H, seed=6)}[kind]()
    fo, mo = O.bk_cut(*w.arrays())
    gc = GraphCut(None, None, (W, H), tile=tile)
    gc.upload(*w.arrays())
    assert not gc.isCompleted()
    reference_loop(gc)
    assert gc.flow == fo
    assert (gc.labels() == mo).all()
    assert gc.tilelistLoad.length == 0  # nothing left active
    # cut() on the same capacities agrees, and so does a second stepped solve
    assert gc.cut() == fo and (gc.labels() == mo).all()
    gc.upload(*w.arrays())
    reference_loop(gc, bfs=2)
    assert gc.flow == fo and (gc.labels() == mo).all()


def test_hooks_warm_after_strokes_and_hook_variants():
    """Stepped warm reflows after device brush strokes; only startBfs (no
    local relabels); isCompleted is False until a relabel leaves no active
    tile, and tile lists show the state."""
    W, H = 400, 300
    w = G.centred(W, H, seed=8)
    gc = GraphCut(None, None, (W, H))
    gc.upload(*w.arrays())
    gc.cut()
    cur = w.copy()
    for st in G.strokes(W, H, count=4, seed=9):
        G.apply_stroke(cur, st)
        s, t = st.capacities()
        gc.paint(st.points, st.width, s, t)
        gc.push()
        assert not gc.isCompleted()  # no relabel since the push
        assert gc.tilelistBfs.length > 0  # the stroke changed tiles
        while True:
            active = gc.startBfs()
            if active == 0:
                break
            assert gc.tilelistLoad.length == active
            gc.push()
        assert gc.isCompleted()
        fo, mo = O.bk_cut(*cur.arrays())
        assert gc.flow == fo and (gc.labels() == mo).all()


def test_lamda_constructor_from_image():
    """GraphCut(ctx, devs, dImg, lamda): the reference's weights built on the
    device from the image (also as packed RGBA in a Buffer2D), unaries from
    cut(dExcess); the same as the oracle on the host model's capacities."""
    W, H = 160, 120
    img = G.three_ellipses(W, H, seed=12).image
    rng = np.random.default_rng(3)
    excess = rng.integers(-400, 400, size=(H, W)).astype(np.int32)
    host = G.caps_from_rgb(img, pair="inv", pair_scale=45.0)
    src = np.maximum(excess, 0).astype(np.int32)
    snk = np.maximum(-excess, 0).astype(np.int32)
    fo, mo = O.bk_cut(src, snk, host.right, host.down)
    rgba = np.concatenate([img, np.full((H, W, 1), 255, np.uint8)], axis=2)
    packed = np.ascontiguousarray(rgba).view(np.uint32).reshape(H, W)
    for dimg in (img, Buffer2D(hostbuf=packed)):
        gc = GraphCut(None, None, dimg, 45)
        assert gc.dim == (W, H) and gc.lamda == 45.0
        assert gc.cut(excess) == fo
        assert (gc.labels() == mo).all()
        # and stepped through the hooks with the same dExcess
        gc2 = GraphCut(None, None, dimg, 45)
        reference_loop(gc2, dExcess=excess)
        assert gc2.flow == fo and (gc2.labels() == mo).all()
    # the reference's default lamda is 60
    assert GraphCut(None, None, img).lamda == 60.0
