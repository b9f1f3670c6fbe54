"""Colour GMMs on the device against the float64 oracle (gmm_oracle):
scores (float32, within one float32 ulp of the oracle's float64 score
rounded to float32), unaries (int32, equal), EM fits (parameters within a
relative 1e-7: the device adds its sums in another order), the GMM drop-in,
and cuts on GMM unaries against Boykov-Kolmogorov."""
import ctypes

import numpy as np
import pytest

import gcmi_lib as L
import gcmi_workloads as G
import gmm_oracle as M
import oracle as O
from GMM import GMM
from GraphCut import GraphCut
from GraphCutBatch import GraphCutBatch
from gcmi_lib import DeviceBuffer, check, lib

pytestmark = pytest.mark.gpu

FG = ([0.5, 0.3, 0.2], [[180, 150, 90], [200, 170, 110], [150, 120, 60]],
      [[400, 300, 500], [250, 250, 250], [900, 600, 400]])
BG = ([0.6, 0.4], [[60, 80, 100], [40, 60, 90]], [[500, 400, 600], [300, 300, 300]])


def device_scores(model, rgb):
    a = np.ascontiguousarray(rgb, np.uint8)
    n = a.shape[0] * a.shape[1] if a.ndim == 3 else a.shape[0]
    ch = a.shape[-1]
    d_in = DeviceBuffer(max(a.nbytes, 4)).upload(a)
    d_out = DeviceBuffer(max(n * 4, 4))
    g = L.gmm(*model)
    check(lib().gcmi_gmm_score(ctypes.byref(g), d_in.ptr, ch, n, d_out.ptr, None))
    check(lib().gcmi_stream_sync(None))
    return d_out.download(np.float32, n)


def assert_scores(got, ref64):
    ref = ref64.astype(np.float32)
    diff = np.abs(got.astype(np.float64) - ref.astype(np.float64))
    assert (diff <= np.spacing(np.abs(ref)).astype(np.float64)).all()
    assert (got == ref).mean() >= 1 - 1e-5


@pytest.mark.parametrize("ch", [3, 4])
def test_scores(ch):
    rng = np.random.default_rng(ch)
    img = rng.integers(0, 256, (333, 517, ch)).astype(np.uint8)
    for model in (FG, BG, ([1.0], [[0, 0, 0]], [[1, 1, 1]]),
                  (rng.dirichlet(np.ones(8)), rng.uniform(0, 255, (8, 3)),
                   rng.uniform(1, 3000, (8, 3)))):
        assert_scores(device_scores(model, img), M.score(*model, img))


def test_gmm_unaries_and_cut():
    W, H = 301, 211
    img = G.three_ellipses(W, H, seed=21).image
    src, snk = (a.reshape(H, W) for a in M.unaries(FG, BG, img, 10.0, 1000.0))
    host = G.caps_from_rgb(img)
    gc = GraphCut(None, None, (W, H))
    gc.upload_gmm(img, L.gmm(*FG), L.gmm(*BG), scale=10.0, cap=1000.0)
    cs, ct, right, down = gc.capacities()
    assert (cs == src).all() and (ct == snk).all()
    assert (right == host.right).all() and (down == host.down).all()
    fo, mo = O.bk_cut(src, snk, host.right, host.down)
    assert gc.cut() == fo
    assert (gc.labels() == mo).all()
    # with the reference's weight (lamda) and in a batch
    inv = G.caps_from_rgb(img, pair="inv", pair_scale=60.0)
    gc2 = GraphCut(None, None, (W, H), lamda=60)
    gc2.upload_gmm(img, L.gmm(*FG), L.gmm(*BG), scale=5.0, cap=2000.0)
    s5, t5 = (a.reshape(H, W) for a in M.unaries(FG, BG, img, 5.0, 2000.0))
    fo2, mo2 = O.bk_cut(s5, t5, inv.right, inv.down)
    assert gc2.cut() == fo2 and (gc2.labels() == mo2).all()
    b = GraphCutBatch([(W, H), (64, 48)])
    b.upload_gmm(0, img, L.gmm(*FG), L.gmm(*BG), 10.0, 1000.0)
    small = G.centred(64, 48, seed=3)
    b.upload(1, *small.arrays())
    flows = b.cut()
    assert flows[0] == fo and (b.labels(0) == mo).all()
    assert flows[1] == O.bk_cut(*small.arrays())[0]


def test_fit_matches_oracle():
    rng = np.random.default_rng(7)
    centres = np.array([[40, 60, 90], [200, 150, 80], [120, 120, 120], [20, 200, 30]])
    x = np.clip(np.concatenate([rng.normal(c, 12 + 4 * i, (25000, 3))
                                for i, c in enumerate(centres)]), 0, 255)
    x = np.rint(x).astype(np.uint8)
    rng.shuffle(x)
    rgba = np.concatenate([x, np.zeros((len(x), 1), np.uint8)], axis=1)
    for samples, ch in ((x, 3), (rgba, 4)):
        d = DeviceBuffer(samples.nbytes).upload(samples)
        g = L.Gmm()
        ll = ctypes.c_double()
        its = ctypes.c_int()
        check(lib().gcmi_gmm_fit(d.ptr, ch, len(x), 4, 15, 0.0, 30.0, 0.001, 0, ctypes.byref(g),
                                 ctypes.byref(ll), ctypes.byref(its)))
        w, mu, v, ll_o, it_o = M.fit(x, 4, 15, init_covar=30.0, min_covar=0.001)
        assert its.value == it_o == 15
        gw, gmu, gv = g.params()
        assert np.allclose(gw, w, rtol=1e-7)
        assert np.allclose(gmu, mu, rtol=1e-7)
        assert np.allclose(gv, v, rtol=1e-7)
        assert np.isclose(ll.value, ll_o, rtol=1e-9)
    # a warm start from the fitted model, and the tolerance stop
    check(lib().gcmi_gmm_fit(d.ptr, 4, len(x), 4, 100, 1e-6, 30.0, 0.001, 1, ctypes.byref(g),
                             ctypes.byref(ll), ctypes.byref(its)))
    assert its.value < 100


def test_gmm_dropin_fit_and_score():
    """GMM(ctx, iterations, components).fit(dSamples) on packed RGBA samples,
    then score(dPopulation, dScore), as the reference's test_GMM does."""
    img = G.three_ellipses(200, 160, seed=9).image
    rgba = np.concatenate([img, np.full(img.shape[:2] + (1,), 255, np.uint8)], axis=2)
    packed = np.ascontiguousarray(rgba).view(np.uint32).reshape(-1)
    d_samples = DeviceBuffer(packed[: 20 * 500].nbytes).upload(packed[: 20 * 500])
    gm = GMM(None, 65, 4)
    weights, means, covars = gm.fit(d_samples, retParams=True)
    w, mu, v, _, it = M.fit(rgba.reshape(-1, 4)[: 20 * 500], 4, 65, tol=0.01)
    assert gm.iterations == it
    assert np.allclose(weights, w, rtol=1e-6) and np.allclose(means, mu, rtol=1e-6)
    assert np.allclose(covars, v, rtol=1e-6)
    d_pop = DeviceBuffer(packed.nbytes).upload(packed)
    d_score = DeviceBuffer(packed.size * 4)
    gm.score(d_pop, d_score)
    assert_scores(d_score.download(np.float32, packed.size), M.score(weights, means, covars, rgba))


def test_gmm_argument_checks():
    g = L.gmm([1.0], [[0, 0, 0]], [[1, 1, 1]])
    g.var[0][1] = 0.0
    d = DeviceBuffer(64)
    assert lib().gcmi_gmm_score(ctypes.byref(g), d.ptr, 3, 4, d.ptr, None) == L.EINVAL
    g2 = L.Gmm()
    assert lib().gcmi_gmm_fit(d.ptr, 3, 2, 3, 5, 0.0, 30.0, 0.001, 0, ctypes.byref(g2),
                              None, None) == L.EINVAL  # fewer samples than components
