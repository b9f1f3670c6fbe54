"""The benchmark workloads at their full sizes, each result against the
oracle (Boykov-Kolmogorov, oracle/gc_oracle.c): flow value and minimal
source set, bit for bit.

  config 1  the 4096x4096 three-ellipse image (seed 1) at tiles 16 and 64,
            host-driven and with the benchmark's run settings
  config 2  64 ragged images, W, H ~ U[256, 2048] (seed 2), at the settings
            the benchmark's autotuner picks (tile 32, 2 sweeps, 1 push round
            per relabel, a recorded hipGraph with 2 rounds, local relabels),
            replayed, and host-driven
  config 3  the 2048x2048 image (seed 3) and all 100 strokes (seed 3), each
            painted by the device brush and reflowed warm with the
            benchmark's fastest variant (local relabels, 1 device round per
            relabel) and host-driven; the oracle after every stroke
  config 4  the 256-session serving mix (seed 4, sizes {512..4096}^2, 200 GB
            budget), new images as colour images, a Poisson trace of fresh
            strokes and new images; every request's flow and every session's
            final mask against the oracle of its capacities at that point
"""
import concurrent.futures as cf
import time

import numpy as np
import pytest

import gcmi_serving as S
import gcmi_workloads as G
import oracle as O
from GraphCut import GraphCut
from GraphCutBatch import GraphCutBatch
from GraphCutServer import DONE, GraphCutServer

pytestmark = pytest.mark.gpu


def packed(mask):
    """uint32 words, bit p % 32 of word p // 32 = mask pixel p."""
    b = np.packbits(np.asarray(mask, np.uint8).reshape(-1), bitorder="little")
    return np.pad(b, (0, (-b.size) % 4)).view(np.uint32)


def bk_many(arrays_list, threads=8):
    with cf.ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda a: O.bk_cut(*a), arrays_list))


# ---------------------------------------------------------------- config 1


@pytest.fixture(scope="module")
def ellipses_4096():
    w = G.three_ellipses(4096, 4096, seed=1)
    f, m = O.bk_cut(*w.arrays())
    return w, f, m


@pytest.mark.timeout(300)
@pytest.mark.parametrize("tile", [16, 64])
@pytest.mark.parametrize("bench_settings", [False, True])
def test_config1_4096_tiles(ellipses_4096, tile, bench_settings):
    w, f, m = ellipses_4096
    gc = GraphCut(None, None, (4096, 4096), tile=tile)
    if bench_settings:
        gc.set_local_relabel(True)
        gc.set_device_rounds(1)
    gc.upload(*w.arrays())
    for rep in range(2):  # a repeat cuts from scratch again
        assert gc.cut() == f, (tile, rep)
        assert (gc.labels() == m).all(), (tile, rep)
    gc.close()


# ---------------------------------------------------------------- config 2


@pytest.fixture(scope="module")
def ragged_64():
    works = G.ragged_batch(64, 256, 2048, seed=2)
    return works, bk_many([w.arrays() for w in works])


@pytest.mark.timeout(300)
@pytest.mark.parametrize("setting", ["autotuned", "host", "device1_local", "tile16_graph4",
                                     "tile64_device8"])
def test_config2_batch_64(ragged_64, setting):
    works, expect = ragged_64
    dims = [(w.shape[1], w.shape[0]) for w in works]
    tile = {"tile16_graph4": 16, "tile64_device8": 64}.get(setting, 32)
    b = GraphCutBatch(dims, tile=tile)
    for i, w in enumerate(works):
        b.upload(i, *w.arrays())
    sweeps, bfs, reps = 4, 1, 1
    if setting == "autotuned":
        b.set_graph(2)
        b.set_local_relabel(True)
        sweeps, reps = 2, 3
    elif setting == "device1_local":
        # (the bench's autotuner picks this or the host-driven run)
        b.set_device_rounds(1)
        b.set_local_relabel(True)
        bfs, reps = 2, 2
    elif setting == "tile16_graph4":
        b.set_graph(4)
        sweeps, bfs, reps = 8, 2, 2
    elif setting == "tile64_device8":
        b.set_device_rounds(8)
        b.set_local_relabel(True)
    for rep in range(reps):  # later runs replay the recording
        flows = b.cut(sweeps=sweeps, bfs=bfs)
        for i, (fo, mo) in enumerate(expect):
            assert flows[i] == fo, (setting, rep, i)
            assert (b.labels(i) == mo).all(), (setting, rep, i)
    b.close()


# ---------------------------------------------------------------- config 3


@pytest.fixture(scope="module")
def strokes_2048():
    W = H = 2048
    work = G.centred(W, H, seed=3)
    strokes = G.strokes(W, H, count=100, seed=3)
    cur = work.copy()
    expect = []
    for st in strokes:
        G.apply_stroke(cur, st)
        f, m = O.bk_cut(*cur.arrays())
        expect.append((f, packed(m)))
    return work, strokes, expect


@pytest.mark.timeout(300)
@pytest.mark.parametrize("variant", ["device1", "host"])
def test_config3_100_strokes(strokes_2048, variant):
    work, strokes, expect = strokes_2048
    gc = GraphCut(None, None, (2048, 2048), tile=32)
    if variant == "device1":
        gc.set_local_relabel(True)
        gc.set_device_rounds(1)
    gc.upload(*work.arrays())
    gc.cut()
    for k, st in enumerate(strokes):
        s, t = st.capacities()
        gc.paint(st.points, st.width, s, t)
        flow = gc.cut()
        bits = gc.labels_packed()
        assert flow == expect[k][0], (variant, k)
        assert (bits == expect[k][1]).all(), (variant, k)
    gc.close()


# ---------------------------------------------------------------- config 4


@pytest.mark.timeout(600)
def test_config4_served_mix():
    mix = S.Mix(nsess=256, seed=4)
    with GraphCutServer(tile=32, hbm_budget=200 * 10**9, max_sessions=256, max_batch=64,
                        mode="batched", local_relabel=True, device_rounds=1) as srv:
        sids = S.open_sessions(srv, mix, rgb=True)
        t0 = time.monotonic_ns() + int(0.2e9)
        strokes, images = S.make_trace(mix, 12000, 0.1, seed=11, t0_ns=t0, image_frac=0.1,
                                       rgb=True)
        ids = S.replay_trace(srv, sids, strokes, images)
        per = [[] for _ in range(mix.nsess)]  # (request id, source) per session
        infos = {}
        for name, tr in (("s", strokes), ("i", images)):
            for rid, item in zip(ids[name], tr):
                rid = int(rid)
                infos[rid] = srv.wait(rid, timeout=120)
                assert infos[rid].status == DONE, rid
                per[item[1]].append((rid, item[5]))
        assert max(i.batch for i in infos.values()) > 1  # continuous batching happened
        finals = [srv.mask_bits(sids[i]) for i in range(mix.nsess)]

    def check(i):
        """Oracle after each of session i's requests (a session's requests
        are served in the order they were queued, which is id order)."""
        n = mix.sizes[i]
        cur = mix.base(i).copy()
        flows = []
        seq = sorted(per[i], key=lambda x: x[0])
        for k, (rid, src) in enumerate(seq):
            if isinstance(src, G.Workload):
                cur = src.copy()
            else:
                G.apply_stroke(cur, G.rasterise(src, n, n))
            if k == len(seq) - 1 or n <= 1024:
                flows.append((rid, O.bk_cut(*cur.arrays())[0]))
        f, m = O.bk_cut(*cur.arrays())
        return flows, f, packed(m), (seq[-1][0] if seq else None)

    with cf.ThreadPoolExecutor(8) as ex:
        results = list(ex.map(check, range(mix.nsess)))
    checked = 0
    for i, (flows, f, bits, last) in enumerate(results):
        for rid, fo in flows:
            assert infos[rid].flow == fo, (i, rid)
            checked += 1
        words, rid = finals[i]
        if last is not None:
            assert rid == last, (i, rid, last)
        assert (words == bits).all(), i
    assert checked >= len(infos) // 3
