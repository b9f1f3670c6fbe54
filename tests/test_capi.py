"""The C-ABI library: loads, exports every declared symbol, validates
arguments; no compute call runs without a GPU."""
import ctypes
import os
import re

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
    assert L.lib().gcmi_server_create(32, 1 << 30, 4, 4, 2, ctypes.byref(h)) == L.EINVAL
    assert L.lib().gcmi_server_wait(None, 1, 0, None) == L.EINVAL
    assert L.lib().gcmi_server_destroy(None) == L.OK


@pytest.mark.skipif(L.device_count() > 0, reason="a GPU is present")
def test_server_without_device_fails_loudly():
    from GraphCutServer import GraphCutServer
    with pytest.raises(L.GcmiError) as e:
        GraphCutServer(hbm_budget=1 << 30, max_sessions=4)
    assert e.value.code == L.ENODEVICE
