"""PrefixSum / StreamCompact / IncrementalTileList on the GPU against the
oracle's closed forms, on the reference tests' own shapes
(test_PrefixSum.py:15-35, test_StreamCompact.py:18-41,
test_IncrementalTileList.py:18-40) and the edge sizes around the one-block
and multi-block paths."""
import numpy as np
import pytest

import oracle as O
from gcmi_lib import DeviceBuffer
from IncrementalTileList import IncrementalTileList, Logical, Operator
from PrefixSum import PrefixSum
from StreamCompact import StreamCompact

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 2, 511, 512, 513, 4096, 65535, 65536, 65537, 262144, 1000003]


def dev(a, n=None):
    n = a.size if n is None else n
    b = DeviceBuffer(max(n, 1) * 4)
    if a.size:
        b.upload(a)
    return b


@pytest.mark.parametrize("n", SIZES)
def test_prefix_sum(n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 20, n).astype(np.int32)
    d, tot = dev(a), DeviceBuffer(4)
    PrefixSum(None, None, max(n, 1)).scan(d, tot, n)
    ref, total = O.exclusive_scan(a)
    assert int(tot.download(np.int32, 1)[0]) == total
    if n:
        assert (d.download(np.int32, n) == ref).all()


def test_prefix_sum_reference_shape():
    # test_PrefixSum.py:15-35: 65,536 ints in [0,20); every element checked
    a = np.random.default_rng(1).integers(0, 20, 65536).astype(np.int32)
    ps = PrefixSum(None, None, 65536)
    d = ps.factory()
    d.upload(a)
    tot = DeviceBuffer(4)
    ps.scan(d, tot, 65536)
    ref, total = O.exclusive_scan(a)
    assert int(tot.download(np.int32, 1)[0]) == total == int(a.sum())
    assert (d.download(np.int32, 65536) == ref).all()


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("density", [0.0, 0.5, 1.0])
def test_stream_compact(n, density):
    rng = np.random.default_rng(n + int(density * 10))
    f = (rng.random(n) < density).astype(np.int32)
    if n > 3:
        f[rng.integers(0, n, 3)] = [7, -1, 2]  # any nonzero counts
    df, dl, dn = dev(f), DeviceBuffer(max(n, 1) * 4), DeviceBuffer(4)
    StreamCompact(None, None, max(n, 1)).compact(df, dl, dn, n)
    ref = O.compact(f)
    k = int(dn.download(np.int32, 1)[0])
    assert k == ref.size
    if k:
        assert (dl.download(np.int32, k) == ref).all()
    if n:
        assert (df.download(np.int32, n) == f).all()  # flags untouched


def test_stream_compact_capacity_error():
    sc = StreamCompact(None, None, 100)
    with pytest.raises(ValueError):
        sc.flagFactory(101)
    with pytest.raises(ValueError):
        sc.compact(DeviceBuffer(4), DeviceBuffer(4), DeviceBuffer(4), 101)


def test_tile_list_reference_case():
    # test_IncrementalTileList.py:18-40: 4096^2 at 16^2 tiles, stamps in
    # [0,20), build(GTE, 10), row-major indices
    tl = IncrementalTileList(None, None, (4096, 4096), (16, 16))
    assert tl.dim == (256, 256)
    hT = np.random.default_rng(2).integers(0, 20, (256, 256)).astype(np.int32)
    tl.set_tiles(hT)
    tl.build(Operator.GTE, 10)
    ys, xs = np.nonzero(hT >= 10)
    ref = ys * 256 + xs
    assert tl.length == ref.size
    assert (tl.list() == ref).all()


@pytest.mark.parametrize("dims", [((400, 300), (16, 16)), ((640, 480), (32, 32)),
                                  ((8192, 8192), (16, 16)), ((48, 16), (16, 16))])
@pytest.mark.parametrize("op", [Operator.EQUAL, Operator.GT, Operator.LT, Operator.GTE,
                                Operator.LTE, 7])
def test_tile_list_operators(dims, op):
    tl = IncrementalTileList(None, None, *dims)
    tw, th = tl.dim
    hT = np.random.default_rng(tw + op).integers(-2, 20, (th, tw)).astype(np.int32)
    tl.set_tiles(hT)
    tl.build(op, 10)
    ref = O.compact(O.tile_flag(hT.reshape(-1), op, 10))
    assert tl.length == ref.size and (tl.list() == ref).all()


@pytest.mark.parametrize("logical", [Logical.AND, Logical.OR, 5])
def test_tile_list_logical_and_incorporate(logical):
    tl = IncrementalTileList(None, None, (1024, 512), (16, 16))
    other = IncrementalTileList(None, None, (1024, 512), (16, 16))
    rng = np.random.default_rng(logical)
    a = rng.integers(0, 20, (32, 64)).astype(np.int32)
    b = rng.integers(0, 20, (32, 64)).astype(np.int32)
    tl.set_tiles(a)
    other.set_tiles(b)
    tl.buildLogical(other.d_tiles, Operator.GT, 5, Operator.LTE, 12, logical)
    ref = O.compact(O.tile_flag_logical(a.reshape(-1), b.reshape(-1), O.GT, 5, O.LTE, 12, logical))
    assert (tl.list() == ref).all()
    it = tl.increment()
    tl.incorporate(other.d_tiles, Operator.LTE, 7, Operator.GTE, 3, logical)
    refT = O.tile_increment_logical(a, b, O.LTE, 7, O.GTE, 3, logical, it)
    assert (tl.tiles() == refT).all()


def test_tile_list_reset_increment_default_build():
    tl = IncrementalTileList(None, None, (256, 256), (32, 32))
    assert tl.iteration == 0 and (tl.tiles() == -1).all()
    tl.build()  # EQUAL iteration (0): nothing
    assert tl.length == 0
    it = tl.increment()
    tl.incorporate(tl.d_tiles, Operator.EQUAL, -1, Operator.EQUAL, -1, Logical.AND)
    tl.build()
    assert tl.length == 64 and (tl.list() == np.arange(64)).all()
    assert it == 1
    tl.reset()
    assert tl.iteration == 0 and (tl.tiles() == -1).all()


def test_native_prim_bench():
    """The native benchmark (scan, compact, tile-list build over the 11 image
    sizes x tiles 16, 32 sweep) checks every result on the host."""
    import os
    import subprocess
    from conftest import REPO
    exe = os.path.join(REPO, "cl-util_amd", "gcmi_prim_bench")
    r = subprocess.run([exe, "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = r.stdout.strip().splitlines()
    assert rows[0].startswith("img dim,mp,tile dim,num tiles")
    assert len(rows) == 1 + 22
    assert all(row.endswith(",ok") for row in rows[1:])
