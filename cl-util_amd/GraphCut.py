"""GraphCut: tiled push-relabel max-flow / min-cut on a 4-connected grid.

Drop-in for the reference's GraphCut (GraphCut.pyx:29-437) with int32
capacities instead of an RGBA image.  Reference -> here:

  GraphCut(context, devices, dImg, lamda)  GraphCut(context, devices, dim, tile=(32,32))
      dImg / lamda built the pairwise       upload(source, sink, right, down) or
      weights on the device                 upload_pairwise(right, down)
  cut(dExcess, bfs=5) -> None               cut(dExcess=None, bfs=None) -> flow value
      dExcess = source - sink, mutated      dExcess (int32 [H][W], source - sink) is
                                            optional and is not mutated
  dHeight (device buffer)                   dHeight: heights as numpy [PH][PW]
  (no labels)                               labels(): uint8 [H][W], 1 = source side
                                            (the minimal source set)

Any width/height is accepted (the reference requires multiples of 32,
GraphCut.pyx:85-86); the grid is padded to whole tiles internally.
"""
import ctypes

import numpy as np

from gcmi_lib import (OPT_DEVICE_ROUNDS, OPT_GRAPH, OPT_LOCAL_RELABEL, DeviceBuffer, check, host_ptr,
                      lib, ptr, rgb_image, rgb_model)

TILE_DEFAULT = (32, 32)
BFS_DEFAULT = 0       # push rounds per global relabel (0: library default, 1)
SWEEPS_DEFAULT = 0    # in-tile sweeps per tile residency (0: library default, 4)

STAT_NAMES = ("rounds", "residencies", "global_relabels", "bfs_passes",
              "flood_passes", "host_syncs", "alg_bytes", "tiles", "relabel_tiles",
              "bfs_round_passes")


def _i32(a, shape):
    a = np.ascontiguousarray(a, dtype=np.int32)
    if a.shape != shape:
        raise ValueError("expected shape %s, got %s" % (shape, a.shape))
    return a


def _points(points):
    xy = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    return xy


def _edit_arrays(index, source, sink):
    idx = np.ascontiguousarray(index, dtype=np.int32).reshape(-1)
    s = np.ascontiguousarray(source, dtype=np.int32).reshape(-1)
    t = np.ascontiguousarray(sink, dtype=np.int32).reshape(-1)
    if not (idx.size == s.size == t.size):
        raise ValueError("index, source and sink must have the same length")
    return idx, s, t


class GraphCut:
    lWorksize = (16, 16)

    def __init__(self, context=None, devices=None, dim=None, tile=TILE_DEFAULT,
                 bfs=BFS_DEFAULT, sweeps=SWEEPS_DEFAULT):
        if dim is None:
            raise ValueError("dim=(width, height) is required")
        if hasattr(dim, "dim"):  # a Buffer2D, as the reference's dImg
            dim = dim.dim
        self.dim = (int(dim[0]), int(dim[1]))
        if isinstance(tile, int):
            tile = (tile, tile)
        if tile[0] != tile[1]:
            raise ValueError("tiles are square (16, 32 or 64)")
        self.tile = (int(tile[0]), int(tile[1]))
        self.bfs = int(bfs)
        self.sweeps = int(sweeps)
        h = ctypes.c_void_p()
        check(lib().gcmi_cut_create(self.dim[0], self.dim[1], self.tile[0], ctypes.byref(h)))
        self._h = h.value
        self._right = None
        self._down = None
        self._device_pairwise = False
        self.flow = None

    @property
    def shape(self):
        return (self.dim[1], self.dim[0])

    def close(self):
        if getattr(self, "_h", None):
            lib().gcmi_cut_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- input
    def upload(self, source, sink, right, down):
        """int32 [H][W] capacities: source (s->p), sink (p->t), right
        (w(p, p+1)), down (w(p, p+W)); copied to the device."""
        s = self.shape
        a = [_i32(x, s) for x in (source, sink, right, down)]
        check(lib().gcmi_cut_upload(self._h, *[host_ptr(x) for x in a]))
        self._right, self._down = a[2], a[3]
        self._device_pairwise = False

    def upload_rgb(self, image, model=None):
        """Builds the capacities on the device from a uint8 [H][W][3] (or
        RGBA [H][W][4]) image: colour-distance unaries and pairwise weights
        (the reference builds its weights from the RGBA image on the device,
        graphcut.cl:1018-1131).  model: an RgbModel (gcmi_lib.rgb_model);
        default: the synthetic workloads' model."""
        a, ch = rgb_image(image)
        if a.shape[:2] != self.shape:
            raise ValueError("expected an image of shape %s" % (self.shape,))
        m = rgb_model() if model is None else model
        check(lib().gcmi_cut_upload_rgb(self._h, host_ptr(a), ch, ctypes.byref(m)))
        self._right = self._down = None
        self._device_pairwise = True

    def capacities(self):
        """(source, sink, right, down) int32 [H][W] as stored on the device."""
        out = [np.empty(self.shape, np.int32) for _ in range(4)]
        check(lib().gcmi_cut_capacities(self._h, *[host_ptr(a) for a in out]))
        return tuple(out)

    def upload_pairwise(self, right, down):
        """Keep pairwise weights for later cut(dExcess) calls."""
        s = self.shape
        self._right, self._down = _i32(right, s), _i32(down, s)

    def edit(self, index, source, sink):
        """Warm-start edit: new source/sink capacities for distinct pixels
        (flat indices y*W + x).  The next cut() reflows from the current
        residual graph; its result equals a from-scratch cut of the edited
        capacities."""
        idx, s, t = _edit_arrays(index, source, sink)
        check(lib().gcmi_cut_edit(self._h, idx.size, host_ptr(idx), host_ptr(s), host_ptr(t)))

    def paint(self, points, width, source, sink):
        """Warm-start edit by a brush stroke rasterised on the device: every
        pixel whose centre lies within width/2 of the polyline `points`
        ([n][2] (x, y), n <= 16) gets capacities (source, sink)."""
        xy = _points(points)
        check(lib().gcmi_cut_paint(self._h, xy.shape[0], host_ptr(xy), float(width),
                                   int(source), int(sink)))

    def upload_device(self, d_source, d_sink, d_right, d_down):
        check(lib().gcmi_cut_upload_device(self._h, ptr(d_source), ptr(d_sink), ptr(d_right),
                                           ptr(d_down)))

    # ------------------------------------------------------------------ cut
    def cut(self, dExcess=None, bfs=None, sweeps=None, max_rounds=0):
        """Run push-relabel to termination; returns the max-flow value.

        dExcess (optional): int32 [H][W] signed excess, source - sink as in
        test_GraphCut.py:44-47; it sets source = max(e, 0), sink = max(-e, 0)
        with the pairwise weights of the last upload / upload_pairwise."""
        if dExcess is not None and self._right is None and self._device_pairwise:
            # pairwise weights built on the device: new unaries for every
            # pixel, cut from scratch
            e = np.asarray(dExcess.get() if hasattr(dExcess, "get") else dExcess)
            e = _i32(e, self.shape).reshape(-1)
            self.forget()
            self.edit(np.arange(e.size, dtype=np.int32), np.maximum(e, 0),
                      np.maximum(-e.astype(np.int64), 0).astype(np.int32))
            dExcess = None
        if dExcess is not None:
            if self._right is None:
                raise ValueError("upload_pairwise() before cut(dExcess)")
            e = np.asarray(dExcess)
            if hasattr(dExcess, "get"):
                e = dExcess.get()
            e = _i32(e, self.shape)
            self.upload(np.maximum(e, 0), np.maximum(-e.astype(np.int64), 0).astype(np.int32),
                        self._right, self._down)
        bfs = self.bfs if bfs is None else int(bfs)
        sweeps = self.sweeps if sweeps is None else int(sweeps)
        flow = ctypes.c_int64()
        check(lib().gcmi_cut_run(self._h, sweeps, bfs, int(max_rounds), ctypes.byref(flow)))
        self.flow = int(flow.value)
        return self.flow

    def labels(self):
        """uint8 [H][W]: 1 for the minimal source (foreground) set."""
        out = np.empty(self.shape, np.uint8)
        check(lib().gcmi_cut_labels(self._h, host_ptr(out)))
        return out

    def labels_packed(self):
        """uint32 words, bit p % 32 of word p // 32 = label of pixel p."""
        out = np.empty((self.shape[0] * self.shape[1] + 31) // 32, np.uint32)
        check(lib().gcmi_cut_labels_packed(self._h, host_ptr(out)))
        return out

    def forget(self):
        """Drop the solved state: the next cut() starts from the stored
        capacities (every edit included) instead of reflowing warm."""
        check(lib().gcmi_cut_forget(self._h))

    def labels_device(self, d_out):
        check(lib().gcmi_cut_labels_device(self._h, ptr(d_out)))

    def stats(self):
        a = (ctypes.c_int64 * len(STAT_NAMES))()
        check(lib().gcmi_cut_stats(self._h, a, len(STAT_NAMES)))
        return dict(zip(STAT_NAMES, [int(v) for v in a]))

    def set_local_relabel(self, enable=True):
        """Reflows of edits, and every run after its first relabel, relabel
        only the changed tiles for their first relabels; the cut is the same
        either way."""
        check(lib().gcmi_cut_set_option(self._h, OPT_LOCAL_RELABEL, 1 if enable else 0))

    def set_device_rounds(self, rounds=8):
        """rounds > 0: device-driven runs -- fixed launch shapes with `rounds`
        BFS / flood rounds per relabel and one host synchronisation at the
        end (the host loop finishes a run that needs more); 0: host-driven.
        The cut is the same either way."""
        check(lib().gcmi_cut_set_option(self._h, OPT_DEVICE_ROUNDS, int(rounds)))

    def set_graph(self, rounds=8):
        """rounds > 0: device-driven runs with `rounds` rounds per relabel,
        recorded once as a hipGraph and replayed for later runs of the same
        images; 0: host-driven runs, no recording."""
        if int(rounds) > 0:
            check(lib().gcmi_cut_set_option(self._h, OPT_DEVICE_ROUNDS, int(rounds)))
            check(lib().gcmi_cut_set_option(self._h, OPT_GRAPH, 1))
        elif int(rounds) == 0:
            check(lib().gcmi_cut_set_option(self._h, OPT_GRAPH, 0))
            check(lib().gcmi_cut_set_option(self._h, OPT_DEVICE_ROUNDS, 0))
        else:
            check(lib().gcmi_cut_set_option(self._h, OPT_DEVICE_ROUNDS, int(rounds)))

    def profile(self, enable=True):
        """Time every later run's kernels with HIP events on the cut's stream."""
        check(lib().gcmi_cut_profile(self._h, 1 if enable else 0))

    def kernel_times(self):
        """{class: (total ms, launches)} of the last profiled run."""
        from gcmi_lib import KERNEL_CLASSES
        n = len(KERNEL_CLASSES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int64 * n)()
        check(lib().gcmi_cut_kernel_times(self._h, ms, cnt, n))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_CLASSES)}

    @property
    def dHeight(self):
        """Heights (BFS distances) of the padded grid, [PH][PW] int32."""
        g = (ctypes.c_int * 3)()
        check(lib().gcmi_cut_geometry(self._h, g))
        out = np.empty((g[1], g[0]), np.int32)
        check(lib().gcmi_cut_heights(self._h, host_ptr(out)))
        return out

    @property
    def maxHeight(self):
        g = (ctypes.c_int * 3)()
        check(lib().gcmi_cut_geometry(self._h, g))
        return g[0] * g[1]

    def reset(self):
        """Kept for API parity (GraphCut.pyx:420-421): every cut() starts
        from the uploaded capacities, so there is nothing to reset."""
        return None
