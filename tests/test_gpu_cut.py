"""The HIP grid cut against the oracle: flow value and minimal source set,
bit for bit, on every tile size; fixtures from the reference's own data;
the full 4096x4096 workload; edge shapes; error paths."""
import numpy as np
import pytest

import gcmi_lib as L
import gcmi_workloads as G
import oracle as O
from conftest import crop_case
from GraphCut import GraphCut

pytestmark = pytest.mark.gpu

TILES = [16, 32, 64]


def run(w, tile, **kw):
    H, W = w.shape if hasattr(w, "shape") else np.shape(w[0])
    arrs = w.arrays() if hasattr(w, "arrays") else w
    gc = GraphCut(None, None, (W, H), tile=tile)
    gc.upload(*arrs)
    flow = gc.cut(**kw)
    return flow, gc.labels(), gc


@pytest.mark.parametrize("seed", range(24))
@pytest.mark.parametrize("tile", TILES)
def test_tiny_vs_brute_force(seed, tile):
    W, H = [(3, 3), (4, 4), (1, 7), (5, 3), (1, 1), (2, 2)][seed % 6]
    w = G.tiny(W, H, seed, cap_hi=[5, 20, 50][seed % 3], w_hi=[3, 10, 30][(seed // 3) % 3])
    fb, mb = O.brute_force(*w.arrays())
    flow, mask, _ = run(w, tile)
    assert flow == fb
    assert (mask == mb).all()


@pytest.mark.parametrize("shape", [(17, 9), (40, 33), (64, 64), (5, 70), (100, 37), (1, 300),
                                   (300, 1), (129, 65), (257, 255)])
@pytest.mark.parametrize("tile", TILES)
def test_random_grids(shape, tile):
    W, H = shape
    for seed in range(3):
        w = G.tiny(W, H, seed * 31 + W, cap_hi=[20, 200, 1000][seed], w_hi=[10, 60, 300][seed])
        fo, mo = O.bk_cut(*w.arrays())
        flow, mask, _ = run(w, tile, sweeps=[1, 4, 16][seed], bfs=[1, 2, 3][seed])
        assert flow == fo, (seed, flow, fo)
        assert (mask == mo).all(), (seed, int((mask != mo).sum()))


@pytest.mark.parametrize("tile", TILES)
def test_centred_512(tile):
    w = G.centred(512, 512, seed=0)
    fp, mp, _ = O.pr_cut(*w.arrays())
    flow, mask, gc = run(w, tile)
    assert flow == fp
    assert (mask == mp).all()
    assert O.cut_value(*w.arrays(), mask) == flow
    # run again on the same handle: identical bits
    assert gc.cut() == flow and (gc.labels() == mask).all()


@pytest.mark.parametrize("name", ["potts8", "ramp"])
@pytest.mark.parametrize("tile", TILES)
def test_reference_fixture(score_crop, name, tile):
    a, flow_ref, mask_ref = crop_case(score_crop, name)
    flow, mask, _ = run(a, tile)
    assert flow == flow_ref
    assert (mask == mask_ref).all()


def test_three_ellipses_4096():
    w = G.three_ellipses(4096, 4096, seed=1)
    fo, mo = O.bk_cut(*w.arrays())
    flow, mask, _ = run(w, 32)
    assert flow == fo
    assert (mask == mo).all()


@pytest.mark.parametrize("tile", TILES)
def test_ragged_1000x700(tile):
    w = G.three_ellipses(1000, 700, seed=5)
    fo, mo = O.bk_cut(*w.arrays())
    flow, mask, _ = run(w, tile)
    assert flow == fo and (mask == mo).all()


def test_hard_constraints():
    w = G.centred(640, 480, seed=3)
    for s in G.strokes(640, 480, count=6, seed=3):
        G.apply_stroke(w, s)
    fo, mo = O.bk_cut(*w.arrays())
    flow, mask, _ = run(w, 32)
    assert flow == fo and (mask == mo).all()


def test_long_path_and_round_cap():
    # flow has to cross a 2000-pixel corridor: needs many rounds
    W, H = 2000, 3
    src = np.zeros((H, W), np.int32)
    snk = np.zeros((H, W), np.int32)
    src[1, 0] = 500
    snk[1, -1] = 500
    right = np.zeros((H, W), np.int32)
    right[1, :-1] = 7
    down = np.zeros((H, W), np.int32)
    flow, mask, gc = run((src, snk, right, down), 32)
    assert flow == 7
    fo, mo = O.bk_cut(src, snk, right, down)
    assert (mask == mo).all()
    with pytest.raises(L.NotConverged):
        gc.cut(max_rounds=1)


def test_excess_interface():
    # cut(dExcess) with excess = source - sink, as test_GraphCut.py:44-47
    w = G.centred(256, 256, seed=4)
    e = w.source.astype(np.int64) - w.sink
    gc = GraphCut(None, None, (256, 256))
    gc.upload_pairwise(w.right, w.down)
    flow = gc.cut(e.astype(np.int32))
    cs, ct = np.maximum(e, 0).astype(np.int32), np.maximum(-e, 0).astype(np.int32)
    fo, mo = O.bk_cut(cs, ct, w.right, w.down)
    assert flow == fo and (gc.labels() == mo).all()


def test_heights_are_bfs_distances():
    w = G.tiny(70, 50, 3, cap_hi=100, w_hi=40)
    flow, mask, gc = run(w, 32)
    h = gc.dHeight[:50, :70]
    assert h.min() >= 0 and h.max() <= gc.maxHeight
    assert (h[mask == 1] == gc.maxHeight).all()  # source side cannot reach a sink


def test_input_validation():
    gc = GraphCut(None, None, (8, 8))
    z = np.zeros((8, 8), np.int32)
    neg = z.copy()
    neg[0, 0] = -1
    with pytest.raises(L.GcmiError):
        gc.upload(neg, z, z, z)
    big = z.copy()
    big[0, 0] = 2 ** 30
    wide = z.copy()
    wide[0, 0] = 2 ** 28
    with pytest.raises(L.GcmiError):
        gc.upload(big, z, wide, z)
    with pytest.raises(ValueError):
        gc.upload(z[:4], z, z, z)
    with pytest.raises(L.GcmiError):
        gc.labels()  # no cut yet
