"""The oracle pinned against brute force, scipy's Dinic, the committed
fixtures derived from the reference's own data, and closed forms."""
import hashlib
import json
import os

import numpy as np
import pytest

import gcmi_workloads as G
import oracle as O
from conftest import GOLDEN, crop_case

TINY = [(3, 3), (4, 4), (2, 5), (4, 3), (1, 7), (5, 3), (1, 1), (2, 2)]


def _tiny(seed):
    W, H = TINY[seed % len(TINY)]
    w = G.tiny(W, H, seed, cap_hi=[5, 20, 50][seed % 3], w_hi=[3, 10, 30][(seed // 3) % 3])
    if seed % 4 == 0:  # near-ties between the terminals
        d = np.random.default_rng(seed).integers(-1, 2, size=w.source.shape)
        w.sink = np.maximum(w.source + d, 0).astype(np.int32)
    return w


@pytest.mark.parametrize("seed", range(120))
def test_tiny_brute_force(seed):
    w = _tiny(seed)
    a = w.arrays()
    fb, mb = O.brute_force(*a)
    fp, mp, _ = O.pr_cut(*a, bfs=[1, 2, 5][seed % 3])
    fk, mk = O.bk_cut(*a)
    assert fp == fb and fk == fb
    assert (mp == mb).all() and (mk == mb).all()


@pytest.mark.parametrize("seed", range(12))
def test_medium_vs_scipy(seed):
    w = G.tiny(37 + seed, 23 + 2 * seed, seed, cap_hi=200, w_hi=80)
    a = w.arrays()
    fs, ms = O.scipy_cut(*a, with_mask=True)
    fp, mp, _ = O.pr_cut(*a)
    fk, mk = O.bk_cut(*a)
    assert fp == fs and fk == fs
    assert (mp == ms).all() and (mk == ms).all()
    assert O.cut_value(*a, mp) == fs  # duality certificate


def test_brute_force_catches_planted_mistake():
    # a dropped pairwise term must change the answer on this instance
    w = G.tiny(4, 4, 7, cap_hi=20, w_hi=30)
    a = list(w.arrays())
    f, m = O.brute_force(*a)
    a2 = [a[0], a[1], np.zeros_like(a[2]), a[3]]
    f2, _ = O.brute_force(*a2)
    assert f2 != f


@pytest.mark.parametrize("name", ["potts8", "ramp"])
def test_reference_fixture(score_crop, name):
    a, flow, mask = crop_case(score_crop, name)
    fk, mk = O.bk_cut(*a)
    fp, mp, _ = O.pr_cut(*a)
    assert fk == flow and fp == flow
    assert (mk == mask).all() and (mp == mask).all()
    assert O.cut_value(*a, mask) == flow


def test_generator_pinned():
    info = json.load(open(os.path.join(GOLDEN, "workloads.json")))["centred_512_seed0"]
    w = G.centred(512, 512, seed=0)
    for k, a in zip(("source", "sink", "right", "down"), w.arrays()):
        assert hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest() == info["sha256"][k]
    fk, mk = O.bk_cut(*w.arrays())
    assert fk == info["flow"]
    assert hashlib.sha256(mk.tobytes()).hexdigest() == info["mask_sha256"]


def test_generator_ranges():
    w = G.three_ellipses(300, 200, seed=1)
    assert w.source.dtype == np.int32 and w.source.shape == (200, 300)
    for a in (w.source, w.sink):
        assert a.min() >= 0 and a.max() <= 1000
    for a in (w.right, w.down):
        assert a.min() >= 0 and a.max() <= 50
    assert (w.right[:, -1] == 0).all() and (w.down[-1, :] == 0).all()


def test_strokes():
    ss = G.strokes(256, 128, count=6, seed=3)
    assert len(ss) == 6
    assert [s.foreground for s in ss] == [True, False] * 3
    for s in ss:
        assert s.index.size > 0 and (np.diff(s.index) > 0).all()
        assert s.index.max() < 256 * 128
    w = G.centred(256, 128, seed=3)
    G.apply_stroke(w, ss[0])
    assert (w.source.reshape(-1)[ss[0].index] == G.HARD).all()


# --------------------------------------------------------- closed forms


def test_scan_closed_form():
    rng = np.random.default_rng(0)
    for n in (0, 1, 7, 1000):
        a = rng.integers(0, 20, n)
        got, total = O.exclusive_scan(a)
        run, ref = 0, []
        for v in a:
            ref.append(run)
            run += int(v)
        assert list(got) == ref and total == run


def test_compact_closed_form():
    f = np.array([0, 1, 1, 0, 5, 0, 0, -1], np.int32)
    assert list(O.compact(f)) == [1, 2, 4, 7]


def test_tile_ops_closed_form():
    t = np.array([5, 10, 15, 10, -1], np.int32)
    assert list(O.tile_flag(t, O.EQUAL, 10)) == [0, 1, 0, 1, 0]
    assert list(O.tile_flag(t, O.GT, 10)) == [0, 0, 1, 0, 0]
    assert list(O.tile_flag(t, O.LT, 10)) == [1, 0, 0, 0, 1]
    assert list(O.tile_flag(t, O.GTE, 10)) == [0, 1, 1, 1, 0]
    assert list(O.tile_flag(t, O.LTE, 10)) == [1, 1, 0, 1, 1]
    assert list(O.tile_flag(t, 9, 10)) == [0, 0, 0, 0, 0]
    u = np.array([1, 2, 3, 4, 5], np.int32)
    assert list(O.tile_flag_logical(t, u, O.GTE, 10, O.LT, 4, O.AND)) == [0, 1, 1, 0, 0]
    assert list(O.tile_flag_logical(t, u, O.GTE, 10, O.LT, 4, O.OR)) == [1, 1, 1, 1, 0]
    assert list(O.tile_increment_logical(t, u, O.GTE, 10, O.LT, 4, O.AND, 99)) == [5, 99, 99, 10, -1]
