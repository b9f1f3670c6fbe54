"""GraphCutBatch: many independent grid cuts solved together.

Every kernel launch of a batched run covers the tiles of all its images, so
the launches, host synchronisations and tail rounds of `len(dims)` cuts are
shared.  Each image keeps its own capacities, flow value and labels (the
minimal source set), exactly as a GraphCut would give.  There is no
counterpart in the reference, which cuts one image per GraphCut object
(GraphCut.pyx:80-418).
"""
import ctypes

import numpy as np

from GraphCut import STAT_NAMES, _edit_arrays, _i32, _points
from gcmi_lib import (KERNEL_CLASSES, OPT_DEVICE_ROUNDS, OPT_GRAPH, OPT_LOCAL_RELABEL, check,
                      host_ptr, lib, ptr, rgb_image, rgb_model)


class GraphCutBatch:
    def __init__(self, dims, tile=32, bfs=0, sweeps=0):
        """dims: sequence of (width, height)."""
        self.dims = [(int(w), int(h)) for w, h in dims]
        if not self.dims:
            raise ValueError("empty batch")
        self.tile = int(tile[0] if isinstance(tile, (tuple, list)) else tile)
        self.bfs = int(bfs)
        self.sweeps = int(sweeps)
        ws = np.array([d[0] for d in self.dims], np.int32)
        hs = np.array([d[1] for d in self.dims], np.int32)
        h = ctypes.c_void_p()
        check(lib().gcmi_batch_create(len(self.dims), host_ptr(ws), host_ptr(hs), self.tile,
                                      ctypes.byref(h)))
        self._h = h.value
        self.flows = None

    def __len__(self):
        return len(self.dims)

    def close(self):
        if getattr(self, "_h", None):
            lib().gcmi_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, i):
        return (self.dims[i][1], self.dims[i][0])

    def upload(self, i, source, sink, right, down):
        s = self._shape(i)
        a = [_i32(x, s) for x in (source, sink, right, down)]
        check(lib().gcmi_batch_upload(self._h, int(i), *[host_ptr(x) for x in a]))

    def upload_rgb(self, i, image, model=None):
        """Image i's capacities built on the device from a colour image (as
        GraphCut.upload_rgb)."""
        a, ch = rgb_image(image)
        if a.shape[:2] != self._shape(i):
            raise ValueError("expected an image of shape %s" % (self._shape(i),))
        m = rgb_model() if model is None else model
        check(lib().gcmi_batch_upload_rgb(self._h, int(i), host_ptr(a), ch, ctypes.byref(m)))

    def upload_gmm(self, i, image, fg, bg, scale=10.0, cap=1000.0, pairwise=None):
        """Image i's capacities with GMM unaries (as GraphCut.upload_gmm;
        pairwise default: the workloads' model)."""
        a, ch = rgb_image(image)
        if a.shape[:2] != self._shape(i):
            raise ValueError("expected an image of shape %s" % (self._shape(i),))
        m = rgb_model() if pairwise is None else pairwise
        f = getattr(fg, "model", fg)
        b = getattr(bg, "model", bg)
        check(lib().gcmi_batch_upload_rgb_gmm(self._h, int(i), host_ptr(a), ch, ctypes.byref(m),
                                              ctypes.byref(f), ctypes.byref(b), float(scale),
                                              float(cap)))

    def upload_device(self, i, d_source, d_sink, d_right, d_down):
        check(lib().gcmi_batch_upload_device(self._h, int(i), ptr(d_source), ptr(d_sink),
                                             ptr(d_right), ptr(d_down)))

    def edit(self, i, index, source, sink):
        idx, s, t = _edit_arrays(index, source, sink)
        check(lib().gcmi_batch_edit(self._h, int(i), idx.size, host_ptr(idx), host_ptr(s),
                                    host_ptr(t)))

    def paint(self, i, points, width, source, sink):
        """Brush stroke on image i (as GraphCut.paint)."""
        xy = _points(points)
        check(lib().gcmi_batch_paint(self._h, int(i), xy.shape[0], host_ptr(xy), float(width),
                                     int(source), int(sink)))

    def cut(self, which=None, bfs=None, sweeps=None, max_rounds=0):
        """Solve all images (or the listed ones); returns their flow values."""
        bfs = self.bfs if bfs is None else int(bfs)
        sweeps = self.sweeps if sweeps is None else int(sweeps)
        if which is None:
            out = np.zeros(len(self.dims), np.int64)
            check(lib().gcmi_batch_run(self._h, sweeps, bfs, int(max_rounds),
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        else:
            w = np.ascontiguousarray(which, np.int32)
            out = np.zeros(w.size, np.int64)
            check(lib().gcmi_batch_run_subset(self._h, w.size, host_ptr(w), sweeps, bfs,
                                              int(max_rounds),
                                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        self.flows = [int(v) for v in out]
        return self.flows

    def labels(self, i):
        out = np.empty(self._shape(i), np.uint8)
        check(lib().gcmi_batch_labels(self._h, int(i), host_ptr(out)))
        return out

    def stats(self):
        a = (ctypes.c_int64 * len(STAT_NAMES))()
        check(lib().gcmi_batch_stats(self._h, a, len(STAT_NAMES)))
        return dict(zip(STAT_NAMES, [int(v) for v in a]))

    def set_local_relabel(self, enable=True):
        """As GraphCut.set_local_relabel."""
        check(lib().gcmi_batch_set_option(self._h, OPT_LOCAL_RELABEL, 1 if enable else 0))

    def set_device_rounds(self, rounds=8):
        """rounds > 0: device-driven runs -- fixed launch shapes with `rounds`
        BFS / flood rounds per relabel and one host synchronisation at the
        end (the host loop finishes a run that needs more); 0: host-driven.
        The cut is the same either way."""
        check(lib().gcmi_batch_set_option(self._h, OPT_DEVICE_ROUNDS, int(rounds)))

    def set_graph(self, rounds=8):
        """rounds > 0: device-driven runs with `rounds` rounds per relabel,
        recorded once as a hipGraph and replayed for later runs of the same
        images; 0: host-driven runs, no recording."""
        if int(rounds) > 0:
            check(lib().gcmi_batch_set_option(self._h, OPT_DEVICE_ROUNDS, int(rounds)))
            check(lib().gcmi_batch_set_option(self._h, OPT_GRAPH, 1))
        elif int(rounds) == 0:
            check(lib().gcmi_batch_set_option(self._h, OPT_GRAPH, 0))
            check(lib().gcmi_batch_set_option(self._h, OPT_DEVICE_ROUNDS, 0))
        else:
            check(lib().gcmi_batch_set_option(self._h, OPT_DEVICE_ROUNDS, int(rounds)))

    def profile(self, enable=True):
        check(lib().gcmi_batch_profile(self._h, 1 if enable else 0))

    def kernel_times(self):
        n = len(KERNEL_CLASSES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int64 * n)()
        check(lib().gcmi_batch_kernel_times(self._h, ms, cnt, n))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_CLASSES)}
