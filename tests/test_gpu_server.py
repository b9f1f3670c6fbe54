"""The serving loop: every request's answer -- flow value and packed minimal
source set -- equals a from-scratch oracle cut of the session's capacities
at that point, whether the session is served alone, in a mix of concurrent
sessions (continuous batching), or by the plain per-request server."""
import threading

import numpy as np
import pytest

import gcmi_lib as L
import gcmi_workloads as G
import oracle as O
from GraphCutServer import DONE, GraphCutServer

pytestmark = pytest.mark.gpu

BUDGET = 1 << 30


def session_script(k, W, H, nstrokes):
    """A session's requests: an image, strokes, a second image, strokes."""
    w1 = G.centred(W, H, seed=500 + k) if k % 2 else G.tiny(W, H, 600 + k, cap_hi=200, w_hi=40)
    w2 = G.three_ellipses(W, H, seed=700 + k)
    st = G.strokes(W, H, count=2 * nstrokes, seed=800 + k)
    return [("image", w1)] + [("stroke", s) for s in st[:nstrokes]] + \
        [("image", w2)] + [("stroke", s) for s in st[nstrokes:]]


def expected(script):
    """Oracle (flow, mask) after each request of a script."""
    out = []
    cur = None
    for kind, x in script:
        if kind == "image":
            cur = x.copy()
        else:
            G.apply_stroke(cur, x)
        f, m = O.bk_cut(*cur.arrays())
        out.append((f, m))
    return out


def run_script(srv, sid, script, sink):
    """Submits each request, waits for it, reads the mask back."""
    for kind, x in script:
        if kind == "image":
            rid = srv.submit_image(sid, *x.arrays())
        else:
            s, t = x.values()
            rid = srv.submit_stroke(sid, x.index, s, t)
        info = srv.wait(rid, timeout=60)
        mask, got = srv.mask(sid)
        assert got == rid
        sink.append((info, mask))


DIMS = [(37, 61), (256, 256), (300, 170), (128, 400), (513, 129), (64, 64)]


@pytest.fixture(scope="module")
def scripts():
    return [session_script(k, W, H, 4) for k, (W, H) in enumerate(DIMS)]


@pytest.fixture(scope="module")
def oracle_results(scripts):
    return [expected(s) for s in scripts]


def check(results, exp, tag):
    assert len(results) == len(exp)
    for j, ((info, mask), (f, m)) in enumerate(zip(results, exp)):
        assert info.status == DONE, (tag, j)
        assert info.flow == f, (tag, j, info.flow, f)
        assert (mask == m).all(), (tag, j)


@pytest.mark.parametrize("tile", [16, 32, 64])
def test_session_alone(tile, scripts, oracle_results):
    for k, (W, H) in enumerate(DIMS[:3]):
        with GraphCutServer(tile=tile, hbm_budget=BUDGET, max_sessions=2) as srv:
            sid = srv.open(W, H)
            res = []
            run_script(srv, sid, scripts[k], res)
            check(res, oracle_results[k], ("alone", tile, k))


@pytest.mark.parametrize("mode,lanes", [("batched", 0), ("plain", 0), ("batched", 1),
                                        ("batched", 2), ("cold", 0)])
def test_mix_equals_alone(mode, lanes, scripts, oracle_results):
    """All sessions submit concurrently from their own threads; each result
    is the oracle's (and so the same as the session served alone), with one
    lane or several solving concurrently."""
    with GraphCutServer(tile=32, hbm_budget=BUDGET, max_sessions=len(DIMS), mode=mode,
                        lanes=lanes) as srv:
        assert srv.stats()["lanes"] == (lanes or (1 if mode == "plain" else 3))
        sids = [srv.open(W, H) for W, H in DIMS]
        results = [[] for _ in DIMS]
        errors = []

        def client(k):
            try:
                run_script(srv, sids[k], scripts[k], results[k])
            except Exception as e:  # reported below
                errors.append((k, e))

        threads = [threading.Thread(target=client, args=(k,)) for k in range(len(DIMS))]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=300)
        assert not errors, errors
        for k in range(len(DIMS)):
            check(results[k], oracle_results[k], (mode, k))
        batches = [info.batch for r in results for info, _ in r]
        if mode == "plain":
            assert max(batches) == 1
        st = srv.stats()
        assert st["served"] == sum(len(s) for s in scripts)
        assert st["sessions"] == len(DIMS)


def test_continuous_batching_queued_requests():
    """Requests submitted without waiting are solved in shared runs; the
    last result of each session equals the oracle's."""
    dims = [(200 + 16 * k, 180 + 8 * k) for k in range(12)]
    works = [G.centred(W, H, seed=900 + k) for k, (W, H) in enumerate(dims)]
    strokes = [G.strokes(W, H, count=3, seed=950 + k) for k, (W, H) in enumerate(dims)]
    with GraphCutServer(tile=32, hbm_budget=BUDGET, max_sessions=16, max_batch=8) as srv:
        sids = [srv.open(W, H) for W, H in dims]
        reqs = []
        for k, sid in enumerate(sids):
            reqs.append(srv.submit_image(sid, *works[k].arrays()))
        for r in range(3):
            for k, sid in enumerate(sids):
                s, t = strokes[k][r].values()
                reqs.append(srv.submit_stroke(sid, strokes[k][r].index, s, t))
        infos = [srv.wait(r, timeout=60) for r in reqs]
        assert all(i.status == DONE for i in infos)
        assert max(i.batch for i in infos) > 1
        assert max(i.batch for i in infos) <= 8
        st = srv.stats()
        assert st["solves"] < len(reqs)
        for k, sid in enumerate(sids):
            cur = works[k].copy()
            for s in strokes[k]:
                G.apply_stroke(cur, s)
            f, m = O.bk_cut(*cur.arrays())
            mask, rid = srv.mask(sid)
            assert rid == reqs[len(sids) * 3 + k]
            assert srv.wait(rid).flow == f
            assert (mask == m).all(), k


def test_admission_by_budget():
    """Sessions open until the pool is full; closing one frees its room."""
    with GraphCutServer(tile=32, hbm_budget=64 << 20, max_sessions=64) as srv:
        cap = srv.stats()["tiles_capacity"]
        assert 0 < cap < 2048
        sids = []
        with pytest.raises(L.GcmiError) as e:
            for _ in range(64):
                sids.append(srv.open(256, 256))  # 64 tiles each
        assert e.value.code == L.EBUDGET
        # each lane holds whole sessions of its share of the tiles
        lanes = srv.stats()["lanes"]
        assert lanes == 3
        assert cap // 64 - lanes < len(sids) <= cap // 64
        w = G.centred(256, 256, seed=1)
        rid = srv.submit_image(sids[0], *w.arrays())
        f, m = O.bk_cut(*w.arrays())
        assert srv.wait(rid, timeout=60).flow == f
        srv.close(sids[-1])
        sid = srv.open(256, 256)
        rid = srv.submit_image(sid, *w.arrays())
        assert srv.wait(rid, timeout=60).flow == f
        assert (srv.mask(sid)[0] == m).all()
        # a session larger than the whole pool is refused
        with pytest.raises(L.GcmiError):
            srv.open(4096, 4096)


def test_request_errors():
    with GraphCutServer(tile=32, hbm_budget=BUDGET, max_sessions=2) as srv:
        sid = srv.open(50, 40)
        with pytest.raises(L.GcmiError):
            srv.submit_stroke(sid, [1, 2], 0, 5)  # before any image
        w = G.tiny(50, 40, 3)
        srv.wait(srv.submit_image(sid, *w.arrays()), timeout=60)
        with pytest.raises(L.GcmiError):
            srv.submit_stroke(sid, [1, 1], 0, 5)  # duplicate pixel
        with pytest.raises(L.GcmiError):
            srv.submit_stroke(sid, [2000], 0, 5)  # out of range
        with pytest.raises(L.GcmiError):
            srv.submit_stroke(sid, [3], -1, 5)  # negative capacity
        with pytest.raises(ValueError):
            srv.submit_image(sid, *[np.zeros((41, 50), np.int32)] * 4)  # wrong shape
        with pytest.raises(L.GcmiError):
            srv.wait(12345)
        with pytest.raises(L.GcmiError):
            srv.submit_stroke(7, [1], 0, 5)  # no such session
        # an empty stroke is a valid request (re-solve)
        info = srv.wait(srv.submit_stroke(sid, [], 0, 0), timeout=60)
        assert info.flow == O.bk_cut(*w.arrays())[0]


def test_submit_and_close_interleaved():
    """Client threads keep submitting strokes to a session while another
    thread closes it and reopens the slot at another size.  A submission
    either fails cleanly (EINVAL: no open session) or is queued before the
    close and served; close waits for every submission that passed its
    session check.  The reopened session then answers as the oracle."""
    W, H = 96, 80
    with GraphCutServer(tile=32, hbm_budget=BUDGET, max_sessions=2, max_batch=8) as srv:
        w = G.centred(W, H, seed=31)
        sid = srv.open(W, H)
        srv.wait(srv.submit_image(sid, *w.arrays()), timeout=60)
        st = G.strokes(W, H, count=8, seed=32)
        stop = threading.Event()
        ok, refused, other = [], [], []

        def client(k):
            j = 0
            while not stop.is_set():
                s = st[(k + j) % len(st)]
                a, b = s.capacities()
                try:
                    ok.append(srv.submit_polyline(sid, s.points, s.width, a, b))
                except L.GcmiError as e:
                    (refused if e.code == L.EINVAL else other).append(e)
                    return
                j += 1

        threads = [threading.Thread(target=client, args=(k,)) for k in range(3)]
        for t in threads:
            t.start()
        import time
        time.sleep(0.05)
        srv.close(sid)
        stop.set()
        for t in threads:
            t.join(timeout=60)
        assert not other, other
        assert ok  # some were queued before the close
        for rid in ok:
            assert srv.wait(rid, timeout=60).status == DONE
        # the slot reopens at another size and answers as the oracle
        w2 = G.three_ellipses(130, 70, seed=33)
        sid2 = srv.open(130, 70)
        rid = srv.submit_image(sid2, *w2.arrays())
        f, m = O.bk_cut(*w2.arrays())
        assert srv.wait(rid, timeout=60).flow == f
        assert (srv.mask(sid2)[0] == m).all()
