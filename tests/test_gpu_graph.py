"""Recorded (hipGraph) runs against the oracle: a run recorded once and
replayed -- for the same images with new capacities, after warm edits, for
changing subsets, and with too few rounds recorded, so that the host loop
finishes the run -- gives the same exact cut as a from-scratch oracle cut."""
import numpy as np
import pytest

import gcmi_workloads as G
import oracle as O
from GraphCut import GraphCut
from GraphCutBatch import GraphCutBatch

pytestmark = pytest.mark.gpu


def ragged(seed, count=7):
    rng = np.random.default_rng(seed)
    dims = [(int(w), int(h)) for w, h in rng.integers(24, 360, size=(count, 2))]
    works = [G.tiny(w, h, seed + i, cap_hi=300, w_hi=60) if i % 2 else
             G.centred(w, h, seed=seed + 50 + i) for i, (w, h) in enumerate(dims)]
    return dims, works


def check_all(b, works, flows):
    for i, w in enumerate(works):
        fo, mo = O.bk_cut(*w.arrays())
        assert flows[i] == fo, i
        assert (b.labels(i) == mo).all(), i


@pytest.mark.parametrize("tile,rounds,record", [(16, 8, True), (32, 8, True), (32, 1, True),
                                               (64, 4, True), (16, 3, False), (32, 8, False)])
def test_recorded_batch_replays(tile, rounds, record):
    dims, works = ragged(10 + tile + rounds)
    b = GraphCutBatch(dims, tile=tile)
    if record:
        b.set_graph(rounds)
    else:
        b.set_device_rounds(rounds)
    for i, w in enumerate(works):
        b.upload(i, *w.arrays())
    check_all(b, works, b.cut())
    # a replay of the same recording over new capacities of the same images
    works2 = [G.tiny(w, h, 900 + i, cap_hi=500, w_hi=40) for i, (w, h) in enumerate(dims)]
    for i, w in enumerate(works2):
        b.upload(i, *w.arrays())
    check_all(b, works2, b.cut())
    for i, w in enumerate(works):
        b.upload(i, *w.arrays())
    check_all(b, works, b.cut())


@pytest.mark.parametrize("record", [True, False])
def test_recorded_warm_strokes_and_subsets(record):
    dims, works = ragged(77, count=5)
    strokes = [G.strokes(w, h, count=4, seed=300 + i) for i, (w, h) in enumerate(dims)]
    b = GraphCutBatch(dims, tile=32)
    if record:
        b.set_graph(6)
    else:
        b.set_device_rounds(6)
        b.set_local_relabel(True)
    for i, w in enumerate(works):
        b.upload(i, *w.arrays())
    b.cut()
    cur = [w.copy() for w in works]
    for r in range(4):
        for i in range(len(dims)):
            st = strokes[i][r]
            s, t = st.capacities()
            b.paint(i, st.points, st.width, s, t)
            G.apply_stroke(cur[i], st)
        check_all(b, cur, b.cut())
    # alternating subsets re-record; each answer is still the oracle's
    for which in ([3, 1], [0, 2, 4], [3, 1]):
        for i in which:
            st = strokes[i][(i + len(which)) % 4]
            s, t = st.capacities()
            b.paint(i, st.points, st.width, s, t)
            G.apply_stroke(cur[i], st)
        flows = b.cut(which=which)
        for k, i in enumerate(which):
            fo, mo = O.bk_cut(*cur[i].arrays())
            assert flows[k] == fo and (b.labels(i) == mo).all(), (which, i)


def test_recorded_plain_cut_and_switch_off():
    w = G.three_ellipses(700, 500, seed=31)
    fo, mo = O.bk_cut(*w.arrays())
    gc = GraphCut(None, None, (700, 500), tile=32)
    gc.set_graph(8)
    gc.upload(*w.arrays())
    for _ in range(3):
        assert gc.cut() == fo
        assert (gc.labels() == mo).all()
    gc.set_graph(0)
    gc.upload(*w.arrays())
    assert gc.cut() == fo and (gc.labels() == mo).all()


def test_recorded_rounds_are_bounded():
    gc = GraphCut(None, None, (64, 64), tile=16)
    with pytest.raises(Exception):
        gc.set_graph(65)
    with pytest.raises(Exception):
        gc.set_graph(-1)
    with pytest.raises(Exception):
        gc.set_device_rounds(100)
