"""CPU oracle for the grid cut and its tile-list primitives.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke test and by the
CPU-baseline leg of bench.py, always as the checker.  The shipped library never
imports it.

Contents
  pr_cut      restatement of the reference's phase-1 push-relabel (C, int64);
              see gc_oracle.c for the reference lines it follows
  bk_cut      independent Boykov-Kolmogorov max-flow (C, int64)
  cut_value   value of a labelling
  brute_force exhaustive min cut over all labellings (<= 16 pixels)
  scipy_cut   scipy.sparse.csgraph.maximum_flow (dinic) on the same graph
  exclusive_scan / compact / tile_flag / tile_flag_logical /
  tile_increment_logical
              closed forms of PrefixSum.scan (prefixsum.cl:81-212,
              PrefixSum.pyx:83-129), StreamCompact.compact
              (streamcompact.cl:1-14, StreamCompact.pyx:66-82) and the
              IncrementalTileList kernels (incremental_tile_list.cl:40-154)
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        i32p = ctypes.POINTER(ctypes.c_int32)
        i64p = ctypes.POINTER(ctypes.c_int64)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.oracle_pr_cut.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p,
                                    i32p, i32p, ctypes.c_int, ctypes.c_int64,
                                    i64p, u8p, i64p]
        L.oracle_pr_cut.restype = ctypes.c_int
        L.oracle_bk_cut.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p,
                                    i32p, i32p, i64p, u8p]
        L.oracle_bk_cut.restype = ctypes.c_int
        L.oracle_cut_value.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p,
                                       i32p, i32p, u8p]
        L.oracle_cut_value.restype = ctypes.c_int64
        _lib = L
    return _lib


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _check(source, sink, right, down):
    H, W = np.shape(source)
    for a in (sink, right, down):
        if np.shape(a) != (H, W):
            raise ValueError("all inputs must share one [H][W] shape")
    return W, H


def pr_cut(source, sink, right, down, bfs=5, max_iters=0):
    """Reference-algorithm restatement.  Returns (flow, mask, iterations)."""
    W, H = _check(source, sink, right, down)
    arrs = [_i32(a) for a in (source, sink, right, down)]
    flow = ctypes.c_int64()
    iters = ctypes.c_int64()
    mask = np.zeros((H, W), np.uint8)
    rc = lib().oracle_pr_cut(W, H, *[p for _, p in arrs], int(bfs),
                             int(max_iters), ctypes.byref(flow),
                             mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                             ctypes.byref(iters))
    if rc == -1:
        raise MemoryError("oracle_pr_cut: allocation failed")
    if rc == -2:
        raise RuntimeError("oracle_pr_cut: iteration cap reached")
    return int(flow.value), mask, int(iters.value)


def bk_cut(source, sink, right, down):
    """Boykov-Kolmogorov.  Returns (flow, mask)."""
    W, H = _check(source, sink, right, down)
    arrs = [_i32(a) for a in (source, sink, right, down)]
    flow = ctypes.c_int64()
    mask = np.zeros((H, W), np.uint8)
    rc = lib().oracle_bk_cut(W, H, *[p for _, p in arrs], ctypes.byref(flow),
                             mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    if rc != 0:
        raise MemoryError("oracle_bk_cut: allocation failed")
    return int(flow.value), mask


def cut_value(source, sink, right, down, mask):
    W, H = _check(source, sink, right, down)
    arrs = [_i32(a) for a in (source, sink, right, down)]
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    return int(lib().oracle_cut_value(
        W, H, *[p for _, p in arrs],
        m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))


def _cut_values_all(source, sink, right, down):
    """Cut value of every labelling of a tiny grid, by bit pattern."""
    H, W = np.shape(source)
    n = W * H
    labels = ((np.arange(1 << n)[:, None] >> np.arange(n)[None, :]) & 1).astype(bool)
    cs = np.asarray(source, np.int64).reshape(-1)
    ct = np.asarray(sink, np.int64).reshape(-1)
    r = np.asarray(right, np.int64).reshape(-1)
    d = np.asarray(down, np.int64).reshape(-1)
    val = np.where(labels, ct[None, :], cs[None, :]).sum(1)
    for p in range(n):
        x, y = p % W, p // W
        if x < W - 1:
            val += np.where(labels[:, p] != labels[:, p + 1], r[p], 0)
        if y < H - 1:
            val += np.where(labels[:, p] != labels[:, p + W], d[p], 0)
    return labels, val


def brute_force(source, sink, right, down):
    """Exhaustive min cut: (min value, intersection of all minimum source
    sets) -- the minimal source set is that intersection."""
    H, W = np.shape(source)
    if W * H > 16:
        raise ValueError("brute force is for <= 16 pixels")
    labels, val = _cut_values_all(source, sink, right, down)
    best = val.min()
    inter = labels[val == best].all(0)
    return int(best), inter.reshape(H, W).astype(np.uint8)


def scipy_cut(source, sink, right, down, with_mask=False):
    """Max-flow value by scipy's Dinic on the same graph (int32 CSR); with
    with_mask, also the minimal source set (reachable from s in scipy's
    final residual graph)."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import breadth_first_order, maximum_flow
    H, W = np.shape(source)
    n = W * H
    s, t = n, n + 1
    cs = np.asarray(source, np.int64).reshape(-1)
    ct = np.asarray(sink, np.int64).reshape(-1)
    base = np.minimum(cs, ct)
    rows, cols, caps = [], [], []
    idx = np.arange(n)
    ex = cs - ct
    m = ex > 0
    rows.append(np.full(m.sum(), s)); cols.append(idx[m]); caps.append(ex[m])
    m = ex < 0
    rows.append(idx[m]); cols.append(np.full(m.sum(), t)); caps.append(-ex[m])
    r = np.asarray(right, np.int64).reshape(H, W)[:, :-1].reshape(-1)
    a = (np.arange(H)[:, None] * W + np.arange(W - 1)[None, :]).reshape(-1)
    rows += [a, a + 1]; cols += [a + 1, a]; caps += [r, r]
    d = np.asarray(down, np.int64).reshape(H, W)[:-1, :].reshape(-1)
    a = (np.arange(H - 1)[:, None] * W + np.arange(W)[None, :]).reshape(-1)
    rows += [a, a + W]; cols += [a + W, a]; caps += [d, d]
    R = np.concatenate(rows); C = np.concatenate(cols); V = np.concatenate(caps)
    keep = V > 0
    g = csr_matrix((V[keep].astype(np.int32), (R[keep], C[keep])),
                   shape=(n + 2, n + 2))
    g.sum_duplicates()
    res = maximum_flow(g, s, t, method="dinic")
    flow = int(res.flow_value) + int(base.sum())
    if not with_mask:
        return flow
    resid = (g.astype(np.int64) - res.flow.astype(np.int64)).tocsr()
    resid.data = (resid.data > 0).astype(np.int8)
    resid.eliminate_zeros()
    order = breadth_first_order(resid, s, directed=True, return_predecessors=False)
    mask = np.zeros(n + 2, np.uint8)
    mask[order] = 1
    return flow, mask[:n].reshape(H, W)


# ------------------------------------------------------------ primitives


def exclusive_scan(values):
    """PrefixSum.scan: in-place exclusive scan and the total."""
    v = np.asarray(values, np.int64)
    out = np.zeros_like(v)
    if v.size:
        out[1:] = np.cumsum(v)[:-1]
    return out.astype(np.int32), int(v.sum())


def compact(flags):
    """StreamCompact.compact: ascending indices of nonzero flags."""
    return np.flatnonzero(np.asarray(flags) != 0).astype(np.int32)


EQUAL, GT, LT, GTE, LTE = 0, 1, 2, 3, 4  # IncrementalTileList.pyx:26-31
AND, OR = 0, 1                           # IncrementalTileList.pyx:33-35


def _cmp(a, op, b):
    a = np.asarray(a)
    if op == EQUAL:
        return a == b
    if op == GT:
        return a > b
    if op == LT:
        return a < b
    if op == GTE:
        return a >= b
    if op == LTE:
        return a <= b
    return np.zeros(a.shape, bool)  # unknown operator flags nothing


def tile_flag(tiles, op, operand):
    """flag kernel, incremental_tile_list.cl:40-64."""
    return _cmp(tiles, op, operand).astype(np.int32)


def tile_flag_logical(t1, t2, op1, a, op2, b, logical):
    """flag_logical kernel, incremental_tile_list.cl:66-108."""
    f1, f2 = _cmp(t1, op1, a), _cmp(t2, op2, b)
    if logical == AND:
        f = f1 & f2
    elif logical == OR:
        f = f1 | f2
    else:
        f = f1
    return f.astype(np.int32)


def tile_increment_logical(t1, t2, op1, a, op2, b, logical, iteration):
    """increment_logical kernel, incremental_tile_list.cl:110-154: stamp
    `iteration` where the predicate holds, keep the old stamp elsewhere."""
    f = tile_flag_logical(t1, t2, op1, a, op2, b, logical).astype(bool)
    return np.where(f, np.int32(iteration), np.asarray(t1, np.int32))
