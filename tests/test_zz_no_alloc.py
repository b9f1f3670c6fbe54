"""gcmi_tiles_build_noalloc (the tile-list build whose multi-block scratch is
a fixed device array, for callers that cannot allocate) through the C ABI:
its list and length equal gcmi_tiles_build's on the same stamps, for two
calls in a row on multi-block sizes (the second smaller, so it reuses the
array the first left behind), a one-block size and an empty input; sizes
above its maximum are refused before any memory is touched."""
import numpy as np
import pytest

import gcmi_lib as L
from gcmi_lib import DeviceBuffer, check, lib
from IncrementalTileList import Operator

pytestmark = pytest.mark.gpu


def build(fn, d_tiles, n, op, operand):
    d_list = DeviceBuffer(max(n, 1) * 4)
    d_len = DeviceBuffer(4)
    d_len.upload(np.array([-7], np.int32))
    check(fn(d_tiles.ptr, n, int(op), int(operand), d_list.ptr, d_len.ptr, None))
    check(lib().gcmi_device_sync())
    k = int(d_len.download(np.int32, 1)[0])
    return k, (d_list.download(np.int32, k) if k > 0 else np.zeros(0, np.int32))


@pytest.mark.parametrize("cases", [
    [(1_000_003, Operator.GTE, 10, 1), (300_001, Operator.LT, 5, 2)],
    [(4097 * 17, Operator.EQUAL, 3, 3), (5000, Operator.GT, 18, 4), (0, Operator.GTE, 0, 5)],
])
def test_noalloc_build_equals_build(cases):
    for n, op, operand, seed in cases:
        stamps = np.random.default_rng(seed).integers(0, 20, max(n, 1)).astype(np.int32)
        d_tiles = DeviceBuffer(max(n, 1) * 4).upload(stamps)
        k0, l0 = build(lib().gcmi_tiles_build_noalloc, d_tiles, n, op, operand)
        k1, l1 = build(lib().gcmi_tiles_build, d_tiles, n, op, operand)
        assert k0 == k1, (n, k0, k1)
        assert np.array_equal(l0, l1), n
        # and both are the predicate's ascending indices
        ref = {Operator.GTE: stamps[:n] >= operand, Operator.LT: stamps[:n] < operand,
               Operator.EQUAL: stamps[:n] == operand, Operator.GT: stamps[:n] > operand}[op]
        assert np.array_equal(l0, np.flatnonzero(ref).astype(np.int32)), n


def test_noalloc_refuses_oversize():
    d = DeviceBuffer(64)
    big = 1 << 26
    assert lib().gcmi_tiles_build_noalloc(d.ptr, big + 1, 3, 0, d.ptr, d.ptr, None) == L.EINVAL
    assert "maximum" in L.last_error()
