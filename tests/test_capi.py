"""The C-ABI library: loads, exports every declared symbol, validates
arguments; no compute call runs without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import gcmi_lib as L
from conftest import REPO


def declared_symbols():
    text = open(os.path.join(REPO, "include", "gcmi.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(gcmi_\w+)\s*\(", text, re.M)))


def test_header_declares_api():
    names = declared_symbols()
    assert "gcmi_cut_run" in names and "gcmi_prefix_sum" in names
    assert len(names) >= 30


def test_library_exports_every_symbol():
    lib = L.lib()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert set(L.SIGNATURES) >= set(declared_symbols())


def test_version_and_errors():
    assert L.lib().gcmi_version().decode().startswith("gcmi")
    h = ctypes.c_void_p()
    assert L.lib().gcmi_cut_create(64, 64, 48, ctypes.byref(h)) == L.EINVAL
    assert "tile" in L.last_error()
    assert L.lib().gcmi_cut_create(0, 64, 32, ctypes.byref(h)) == L.EINVAL
    assert L.lib().gcmi_cut_run(None, 0, 0, 0, None) == L.EINVAL


@pytest.mark.skipif(L.device_count() > 0, reason="a GPU is present")
def test_no_device_fails_loudly():
    h = ctypes.c_void_p()
    assert L.lib().gcmi_cut_create(64, 64, 32, ctypes.byref(h)) == L.ENODEVICE
    assert L.lib().gcmi_prefix_sum(None, None, 0, None) == L.ENODEVICE
    from GraphCut import GraphCut
    with pytest.raises(L.GcmiError):
        GraphCut(None, None, (64, 64))


def test_server_argument_checks():
    h = ctypes.c_void_p()
    assert L.lib().gcmi_server_create(48, 1 << 30, 4, 4, 0, ctypes.byref(h)) == L.EINVAL
    assert L.lib().gcmi_server_create(32, 1 << 30, 0, 4, 0, ctypes.byref(h)) == L.EINVAL
    assert L.lib().gcmi_server_create(32, 1 << 30, 4, 4, 3, ctypes.byref(h)) == L.EINVAL
    assert L.lib().gcmi_server_wait(None, 1, 0, None) == L.EINVAL
    assert L.lib().gcmi_server_destroy(None) == L.OK


@pytest.mark.skipif(L.device_count() > 0, reason="a GPU is present")
def test_server_without_device_fails_loudly():
    from GraphCutServer import GraphCutServer
    with pytest.raises(L.GcmiError) as e:
        GraphCutServer(hbm_budget=1 << 30, max_sessions=4)
    assert e.value.code == L.ENODEVICE


def test_request_struct_layout():
    """gcmi_request: int64, six int32, a double, four pointers (72 bytes)."""
    assert ctypes.sizeof(L.Request) == 72
    assert L.Request.width.offset == 32
    assert L.Request.p0.offset == 40


def test_stroke_polyline_rasteriser_cases():
    """The host rasteriser (the oracle of the device brush) on cases fixed
    by geometry alone."""
    import gcmi_workloads as G
    # horizontal segment on y = 3.5, width 1: rows 3 and 4 are exactly r away
    idx = G.rasterise_polyline(10, 8, [(2.0, 3.5), (6.0, 3.5)], 1)
    ys, xs = np.divmod(idx, 10)
    assert set(ys) == {3, 4} and set(xs) == set(range(2, 7))
    # width 0 on the diagonal: exactly the centres on it
    idx = G.rasterise_polyline(6, 6, [(0.0, 0.0), (5.0, 5.0)], 0)
    assert list(idx) == [0, 7, 14, 21, 28, 35]
    # one point paints nothing; a disc of radius 2 around a repeated point
    assert G.rasterise_polyline(9, 9, [(4.0, 4.0)], 5).size == 0
    idx = G.rasterise_polyline(9, 9, [(4.0, 4.0), (4.0, 4.0)], 4)
    ys, xs = np.divmod(idx, 9)
    assert ((xs - 4) ** 2 + (ys - 4) ** 2 <= 4).all() and idx.size == 13
    # outside the image
    assert G.rasterise_polyline(9, 9, [(-30.0, -30.0), (-10.0, -20.0)], 8).size == 0
    # the same cases through the bounding-box statement
    assert list(G.rasterise_polyline_box(6, 6, [(0.0, 0.0), (5.0, 5.0)], 0)) == \
        [0, 7, 14, 21, 28, 35]
    assert G.rasterise_polyline_box(9, 9, [(4.0, 4.0), (4.0, 4.0)], 4).size == 13


def test_band_rasteriser_equals_box_rasteriser():
    """The band rasteriser (tests only pixels near each segment) paints the
    same pixels as the bounding-box statement, on random, axis-parallel,
    integer-vertex and partly outside polylines."""
    import gcmi_workloads as G
    rng = np.random.default_rng(5)
    for k in range(300):
        W, H = int(rng.integers(1, 200)), int(rng.integers(1, 200))
        n = int(rng.integers(1, 6))
        pts = list(zip(rng.uniform(-40, W + 40, n), rng.uniform(-40, H + 40, n)))
        if k % 7 == 0:
            pts = [(x, pts[0][1]) for x, _ in pts]
        if k % 11 == 0:
            pts = [(pts[0][0], y) for _, y in pts]
        if k % 13 == 0:
            pts = [(float(int(x)), float(int(y))) for x, y in pts]
        width = float(rng.choice([0, 1, 2, 3.5, 8, 17, 32]))
        assert np.array_equal(G.rasterise_polyline(W, H, pts, width),
                              G.rasterise_polyline_box(W, H, pts, width)), (k, W, H, pts, width)


PRIM_BENCH = os.path.join(REPO, "cl-util_amd", "gcmi_prim_bench")


@pytest.mark.skipif(L.device_count() > 0, reason="a GPU is present")
def test_prim_bench_without_device():
    """The native primitive benchmark is built and exits with status 3 (no
    usable GPU) instead of measuring anything."""
    import subprocess
    assert os.path.exists(PRIM_BENCH)
    r = subprocess.run([PRIM_BENCH, "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3
    assert "no HIP device" in r.stderr


def test_rgb_model_layout_and_default():
    """struct gcmi_rgb_model as ctypes sees it; the default model is the
    synthetic workloads' (host-only calls, no device needed)."""
    assert ctypes.sizeof(L.RgbModel) == 104
    assert L.RgbModel.pair_kind.offset == 72 and L.RgbModel.pair_scale.offset == 80
    m = L.rgb_model()
    assert list(m.mu_bg) == [60.0, 80.0, 100.0] and list(m.mu_fg) == [180.0, 150.0, 90.0]
    assert (m.unary_scale, m.unary_div, m.unary_cap) == (10.0, 100.0, 100.0)
    assert m.pair_kind == L.PAIR_EXP and m.pair_scale == 50.0
    assert (m.inv_beta, m.inv_eps) == (0.1, 0.05)
    m2 = L.rgb_model(pair_kind=L.PAIR_INV, mu_fg=(1, 2, 3))
    assert m2.pair_kind == L.PAIR_INV and list(m2.mu_fg) == [1.0, 2.0, 3.0]
    with pytest.raises(TypeError):
        L.rgb_model(lamda=60)


def test_rgb_host_model_matches_generator():
    """The colour model's host statement reproduces the generator's
    capacities, and its inverse weight follows the reference's formula at
    closed-form points (d2 = 0: lamda / eps; d2 = 100: lamda / (1 + eps))."""
    import gcmi_workloads as G
    w = G.centred(64, 48, seed=2)
    c = G.caps_from_rgb(w.image)
    for a, b in zip(w.arrays(), c.arrays()):
        assert (a == b).all()
    img = np.zeros((1, 3, 3), np.uint8)
    img[0, 1] = (6, 8, 0)  # d2 = 100 to both neighbours
    inv = G.caps_from_rgb(img, pair="inv", pair_scale=60.0)
    assert inv.right[0, 0] == round(60.0 / 1.05) == 57 and inv.right[0, 1] == 57
    flat = G.caps_from_rgb(np.zeros((2, 2, 3), np.uint8), pair="inv", pair_scale=60.0)
    assert flat.right[0, 0] == 1200 and flat.down[0, 0] == 1200 and flat.right[0, 1] == 0
