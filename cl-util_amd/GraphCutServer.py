"""GraphCutServer: resident grid-cut sessions behind a continuously batching
worker (libgcmi's gcmi_server_*).

Each session is one image whose capacities, residual graph and heights stay
in device memory between requests.  A request is a new image (cut from its
capacities) or a brush stroke (new unaries for some pixels, reflowed warm
from the session's residual graph).  The worker thread inside the library
solves every queued request it can -- one per session, up to `max_batch` --
in one batched run; requests that arrive meanwhile join the next run.  The
answer of every request is the same as a from-scratch GraphCut of the
session's capacities at that point: the max-flow value and the minimal
source set.

mode "batched" is the serving loop (strokes reflow warm); mode "cold"
batches the same way but cuts every session from scratch; mode "plain"
serves one request at a time, each cut from scratch, as the baseline.

No counterpart in the reference, which serves one user per GraphCut object
and re-uploads the excess image after each stroke (test_GraphCut.py:74-75).
"""
import ctypes

import numpy as np

from GraphCut import _i32, _points
from gcmi_lib import (EBUSY, OPT_DEVICE_ROUNDS, OPT_LOCAL_RELABEL, REQ_IMAGE, REQ_POLYLINE, REQ_RGB,
                      REQ_STROKE, GcmiError, Request, _i64p, check, host_ptr, lib, rgb_image)

QUEUED, DONE, FAILED = 0, 1, 2
MODES = {"batched": 0, "plain": 1, "cold": 2}
STAT_NAMES = ("solves", "served", "queued", "sessions", "tiles_used", "tiles_capacity",
              "device_bytes", "ns_apply", "ns_solve", "ns_out", "lanes", "ns_wait")


class RequestInfo:
    """Completion record of one request (times in ns of a steady clock)."""

    __slots__ = ("status", "flow", "t_arrive", "t_start", "t_done", "batch", "t_queued")

    def __init__(self, v):
        (self.status, self.flow, self.t_arrive, self.t_start, self.t_done, self.batch,
         self.t_queued) = (int(x) for x in v)

    @property
    def latency_ms(self):
        return (self.t_done - self.t_arrive) / 1e6

    @property
    def queue_ms(self):
        return (self.t_start - self.t_arrive) / 1e6


class GraphCutServer:
    def __init__(self, tile=32, hbm_budget=200 * 10**9, max_sessions=256, max_batch=64,
                 mode="batched", sweeps=0, bfs=0, local_relabel=False, device_rounds=0,
                 lanes=0):
        """lanes: pools solved concurrently, each with its share of the
        sessions (0: 3 for "batched" and "cold", 1 for "plain")."""
        self.tile = int(tile[0] if isinstance(tile, (tuple, list)) else tile)
        if mode not in MODES:
            raise ValueError("mode must be one of %s" % sorted(MODES))
        self.mode = mode
        h = ctypes.c_void_p()
        check(lib().gcmi_server_create_lanes(self.tile, int(hbm_budget), int(max_sessions),
                                             int(max_batch), MODES[mode], int(lanes),
                                             ctypes.byref(h)))
        self._h = h.value
        self.dims = {}
        if sweeps or bfs:
            check(lib().gcmi_server_set_params(self._h, int(sweeps), int(bfs)))
        if local_relabel:
            check(lib().gcmi_server_set_option(self._h, OPT_LOCAL_RELABEL, 1))
        if device_rounds:
            check(lib().gcmi_server_set_option(self._h, OPT_DEVICE_ROUNDS, int(device_rounds)))

    def close_server(self):
        if getattr(self, "_h", None):
            lib().gcmi_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close_server()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close_server()

    # ------------------------------------------------------------ sessions
    def open(self, width, height):
        """Admits a width x height session; GcmiError(EBUDGET) if the pool
        has no slot or room for it."""
        sid = ctypes.c_int()
        check(lib().gcmi_server_open(self._h, int(width), int(height), ctypes.byref(sid)))
        self.dims[sid.value] = (int(width), int(height))
        return sid.value

    def close(self, session):
        check(lib().gcmi_server_close(self._h, int(session)))
        self.dims.pop(session, None)

    def _shape(self, session):
        w, h = self.dims[session]
        return (h, w)

    # ------------------------------------------------------------ requests
    def submit_image(self, session, source, sink, right, down):
        s = self._shape(session)
        a = [_i32(x, s) for x in (source, sink, right, down)]
        rid = ctypes.c_int64()
        check(lib().gcmi_server_submit_image(self._h, int(session), *[host_ptr(x) for x in a],
                                             ctypes.byref(rid)))
        return rid.value

    def set_rgb_model(self, model):
        """Colour model (gcmi_lib.rgb_model) of later submit_rgb requests."""
        check(lib().gcmi_server_set_rgb_model(self._h, ctypes.byref(model)))

    def submit_rgb(self, session, image):
        """Queues a new image given as uint8 [H][W][3 or 4] colour; its
        capacities are built on the device (GraphCut.upload_rgb)."""
        a, ch = rgb_image(image)
        if a.shape[:2] != self._shape(session):
            raise ValueError("expected an image of shape %s" % (self._shape(session),))
        rid = ctypes.c_int64()
        check(lib().gcmi_server_submit_rgb(self._h, int(session), host_ptr(a), ch,
                                           ctypes.byref(rid)))
        return rid.value

    def submit_stroke(self, session, index, source, sink):
        idx = np.ascontiguousarray(index, dtype=np.int32).reshape(-1)
        n = idx.size
        src = np.ascontiguousarray(np.broadcast_to(source, (n,)), dtype=np.int32)
        snk = np.ascontiguousarray(np.broadcast_to(sink, (n,)), dtype=np.int32)
        rid = ctypes.c_int64()
        check(lib().gcmi_server_submit_stroke(self._h, int(session), n, host_ptr(idx),
                                              host_ptr(src), host_ptr(snk), ctypes.byref(rid)))
        return rid.value

    def submit_polyline(self, session, points, width, source, sink):
        """Queues a brush stroke rasterised on the device (GraphCut.paint)."""
        xy = _points(points)
        rid = ctypes.c_int64()
        check(lib().gcmi_server_submit_polyline(self._h, int(session), xy.shape[0],
                                                host_ptr(xy), float(width), int(source),
                                                int(sink), ctypes.byref(rid)))
        return rid.value

    def replay(self, trace):
        """Open-loop load generator: trace is a list of (t_ns, session, kind,
        payload) with t_ns on time.monotonic_ns()'s clock (<= 0: at once),
        kind "image" (payload: source, sink, right, down), "rgb" (a uint8
        colour image), "stroke" (index, source, sink) or "polyline" (points,
        width, source, sink).
        Submits each request at its time from the calling thread (the GIL
        is released meanwhile) and returns the request ids."""
        n = len(trace)
        reqs = (Request * n)()
        keep = []
        for i, (t_ns, session, kind, payload) in enumerate(trace):
            q = reqs[i]
            q.t_ns = int(t_ns)
            q.session = int(session)
            if kind == "image":
                shape = self._shape(session)
                a = [_i32(x, shape) for x in payload]
                q.kind = REQ_IMAGE
                q.p0, q.p1, q.p2, q.p3 = (x.ctypes.data for x in a)
            elif kind == "stroke":
                idx = np.ascontiguousarray(payload[0], dtype=np.int32).reshape(-1)
                m = idx.size
                s = np.ascontiguousarray(np.broadcast_to(payload[1], (m,)), dtype=np.int32)
                t = np.ascontiguousarray(np.broadcast_to(payload[2], (m,)), dtype=np.int32)
                a = [idx, s, t]
                q.kind = REQ_STROKE
                q.n = m
                q.p0, q.p1, q.p2 = (x.ctypes.data for x in a)
            elif kind == "rgb":
                a, ch = rgb_image(payload)
                if a.shape[:2] != self._shape(session):
                    raise ValueError("request %d: image shape %s" % (i, a.shape))
                q.kind = REQ_RGB
                q.n = ch
                q.p0 = a.ctypes.data
                a = [a]
            elif kind == "polyline":
                xy = _points(payload[0])
                a = [xy]
                q.kind = REQ_POLYLINE
                q.n = xy.shape[0]
                q.width = float(payload[1])
                q.source, q.sink = int(payload[2]), int(payload[3])
                q.p0 = xy.ctypes.data
            else:
                raise ValueError("unknown request kind %r" % (kind,))
            keep.append(a)
        ids = np.zeros(n, np.int64)
        check(lib().gcmi_server_replay(self._h, n, reqs, ids.ctypes.data_as(_i64p)))
        return ids

    def wait(self, request, timeout=None):
        """RequestInfo of a request; raises GcmiError(EBUSY) if it is not
        done within `timeout` seconds (None: no limit)."""
        info = (ctypes.c_int64 * 7)()
        ms = -1 if timeout is None else int(timeout * 1000)
        rc = lib().gcmi_server_wait(self._h, int(request), ms, info)
        if rc == EBUSY:
            raise GcmiError(rc, "request %d still in flight" % request)
        check(rc)
        return RequestInfo(info)

    def poll(self, request):
        """RequestInfo, status QUEUED while in flight (no wait)."""
        info = (ctypes.c_int64 * 7)()
        rc = lib().gcmi_server_wait(self._h, int(request), 0, info)
        if rc not in (0, EBUSY):
            check(rc)
        return RequestInfo(info)

    def forget(self, request):
        check(lib().gcmi_server_forget(self._h, int(request)))

    def mask_bits(self, session):
        """(packed uint32 words, request id) of the session's last result."""
        w, h = self.dims[session]
        out = np.empty((w * h + 31) // 32, np.uint32)
        rid = ctypes.c_int64()
        check(lib().gcmi_server_mask(self._h, int(session), host_ptr(out), ctypes.byref(rid)))
        return out, rid.value

    def mask(self, session):
        """(uint8 [H][W] labels, request id): 1 = minimal source set."""
        w, h = self.dims[session]
        bits, rid = self.mask_bits(session)
        lab = np.unpackbits(bits.view(np.uint8), bitorder="little")[: w * h]
        return lab.reshape(h, w), rid

    def stats(self):
        out = (ctypes.c_int64 * len(STAT_NAMES))()
        check(lib().gcmi_server_stats(self._h, out, len(STAT_NAMES)))
        return dict(zip(STAT_NAMES, (int(x) for x in out)))
