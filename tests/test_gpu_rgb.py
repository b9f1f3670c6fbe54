"""Capacities built on the device from 8-bit colour images: bit-exact with
the host statement of the colour model (gcmi_workloads.caps_from_rgb) for
both pairwise models -- the workloads' exp(-d2 / 2 beta) and the
reference's lamda / (0.1 sqrt(d2) + 0.05) (graphcut.cl:1018-1031) -- and the
cut of those capacities equals the oracle's, in the plain, batched and
served forms."""
import numpy as np
import pytest

import gcmi_workloads as G
import oracle as O
from GraphCut import GraphCut
from GraphCutBatch import GraphCutBatch
from GraphCutServer import GraphCutServer
from gcmi_lib import PAIR_EXP, PAIR_INV, GcmiError, rgb_model

pytestmark = pytest.mark.gpu

MODELS = {
    "exp": ({}, {}),
    "inv60": (dict(pair_kind=PAIR_INV, pair_scale=60.0), dict(pair="inv", pair_scale=60.0)),
    "custom": (dict(mu_bg=(61.5, 79.25, 100.0), mu_fg=(200.0, 10.0, 33.3), unary_scale=7.0,
                    unary_div=37.0, unary_cap=250.0, pair_scale=123.0),
               dict(mu_bg=(61.5, 79.25, 100.0), mu_fg=(200.0, 10.0, 33.3), unary_scale=7.0,
                    unary_div=37.0, unary_cap=250.0, pair_scale=123.0)),
    "inv_custom": (dict(pair_kind=PAIR_INV, pair_scale=9.5, inv_beta=0.37, inv_eps=0.011),
                   dict(pair="inv", pair_scale=9.5, inv_beta=0.37, inv_eps=0.011)),
    # integer means, scale 1 and divisor 2: every odd squared distance puts
    # the unary exactly on a half-integer, where the device must fall back
    # from its reciprocal product to the division and round ties to even
    "halves": (dict(mu_bg=(60.0, 80.0, 100.0), mu_fg=(200.0, 10.0, 33.0), unary_scale=1.0,
                    unary_div=2.0, unary_cap=1e6),
               dict(mu_bg=(60.0, 80.0, 100.0), mu_fg=(200.0, 10.0, 33.0), unary_scale=1.0,
                    unary_div=2.0, unary_cap=1e6)),
}


def images():
    rng = np.random.default_rng(17)
    out = [G.centred(512, 384, seed=4).image, G.three_ellipses(300, 700, seed=8).image]
    for w, h in [(1, 1), (1, 37), (53, 1), (97, 61), (256, 256)]:
        out.append(rng.integers(0, 256, size=(h, w, 3), dtype=np.uint8))
    flat = np.full((40, 30, 3), 77, np.uint8)  # every d2 = 0: beta falls back to 1
    out.append(flat)
    # every sum of two squares up to 2 * 255^2 as a horizontal pair (black |
    # (a, b, 0)), and the vertical pairs between rows: the float32 first try
    # of the exponential weight and its float64 fallback over the whole range
    a, b = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    pairs = np.zeros((65536, 2, 3), np.uint8)
    pairs[:, 1, 0] = a.reshape(-1)
    pairs[:, 1, 1] = b.reshape(-1)
    out.append(pairs)
    rgba = np.concatenate([out[1], rng.integers(0, 256, size=out[1].shape[:2] + (1,),
                                                dtype=np.uint8)], axis=2)
    out.append(rgba)
    # odd widths with many whole groups of four pixels: the pair sums' word
    # reads of the row below start at every byte offset of a word (3 * W is
    # odd), and an RGBA image whose rows end inside a group
    out.append(rng.integers(0, 256, size=(257, 1001, 3), dtype=np.uint8))
    out.append(rng.integers(0, 256, size=(111, 333, 4), dtype=np.uint8))
    return out


@pytest.mark.parametrize("name", sorted(MODELS))
def test_rgb_capacities_bit_exact(name):
    dev, host = MODELS[name]
    m = rgb_model(**dev)
    for k, img in enumerate(images()):
        H, W = img.shape[:2]
        gc = GraphCut(None, None, (W, H), tile=32)
        gc.upload_rgb(img, m)
        ref = G.caps_from_rgb(img, **host)
        got = gc.capacities()
        for a, b, what in zip(got, ref.arrays(), ("source", "sink", "right", "down")):
            assert (a == b).all(), (name, k, what, int((a != b).sum()))
        gc.close()


@pytest.mark.parametrize("name", ["exp", "inv60"])
def test_rgb_cut_matches_oracle(name):
    dev, host = MODELS[name]
    img = G.three_ellipses(640, 480, seed=12).image
    gc = GraphCut(None, None, (640, 480), tile=32)
    gc.upload_rgb(img, rgb_model(**dev))
    ref = G.caps_from_rgb(img, **host)
    fo, mo = O.bk_cut(*ref.arrays())
    assert gc.cut() == fo
    assert (gc.labels() == mo).all()
    # the reference's calling convention: new excess, pairwise kept
    e = ref.source.astype(np.int64) - ref.sink
    e[100:140, 200:260] = 1 << 30
    e[300:320, 10:90] = -(1 << 30)
    e = e.astype(np.int32)
    fo2, mo2 = O.bk_cut(np.maximum(e, 0), np.maximum(-e.astype(np.int64), 0).astype(np.int32),
                        ref.right, ref.down)
    assert gc.cut(e) == fo2
    assert (gc.labels() == mo2).all()


def test_rgb_batch_and_server():
    rng = np.random.default_rng(3)
    dims = [(int(w), int(h)) for w, h in rng.integers(40, 300, size=(6, 2))]
    imgs = [G.centred(w, h, seed=90 + i).image for i, (w, h) in enumerate(dims)]
    refs = [G.caps_from_rgb(im) for im in imgs]
    b = GraphCutBatch(dims, tile=16)
    for i, im in enumerate(imgs):
        b.upload_rgb(i, im)
    flows = b.cut()
    for i, r in enumerate(refs):
        fo, mo = O.bk_cut(*r.arrays())
        assert flows[i] == fo and (b.labels(i) == mo).all(), i
    strokes = [G.strokes(w, h, count=2, seed=40 + i) for i, (w, h) in enumerate(dims)]
    with GraphCutServer(tile=32, hbm_budget=1 << 30, max_sessions=8, local_relabel=True,
                        device_rounds=6) as srv:
        sids = [srv.open(w, h) for w, h in dims]
        trace = [(0, sids[i], "rgb", imgs[i]) for i in range(len(dims))]
        cur = [r.copy() for r in refs]
        for rnd in range(2):
            for i in range(len(dims)):
                st = strokes[i][rnd]
                s, t = st.capacities()
                trace.append((0, sids[i], "polyline", (st.points, st.width, s, t)))
                G.apply_stroke(cur[i], st)
        for rid in srv.replay(trace):
            assert srv.wait(int(rid), timeout=60).status == 1
        for i, sid in enumerate(sids):
            mask, _ = srv.mask(sid)
            assert (mask == O.bk_cut(*cur[i].arrays())[1]).all(), i
        # submit_rgb with the reference's weights
        srv.set_rgb_model(rgb_model(pair_kind=PAIR_INV, pair_scale=60.0))
        rid = srv.submit_rgb(sids[0], imgs[0])
        info = srv.wait(rid, timeout=60)
        fo, mo = O.bk_cut(*G.caps_from_rgb(imgs[0], pair="inv", pair_scale=60.0).arrays())
        assert info.flow == fo and (srv.mask(sids[0])[0] == mo).all()


def test_rgb_arguments():
    gc = GraphCut(None, None, (20, 10), tile=16)
    img = np.zeros((10, 20, 3), np.uint8)
    with pytest.raises(ValueError):
        gc.upload_rgb(np.zeros((10, 20, 2), np.uint8))
    with pytest.raises(ValueError):
        gc.upload_rgb(np.zeros((11, 20, 3), np.uint8))
    for bad in (dict(unary_div=0.0), dict(unary_scale=1e6, unary_cap=1e6),
                dict(pair_kind=7), dict(pair_kind=PAIR_INV, inv_eps=0.0),
                dict(pair_scale=float("nan")), dict(pair_scale=1e9)):
        with pytest.raises(GcmiError):
            gc.upload_rgb(img, rgb_model(**bad))
    # a black image: source 10 * min(20000 / 100, 100) = 2000 and sink
    # 10 * min(63000 / 100, 100) = 1000 per pixel, so every pixel sends 1000
    gc.upload_rgb(img, rgb_model(pair_kind=PAIR_EXP))
    assert gc.cut() == 1000 * 20 * 10 == O.bk_cut(*G.caps_from_rgb(img).arrays())[0]
