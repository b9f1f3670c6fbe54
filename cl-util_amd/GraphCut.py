"""GraphCut: tiled push-relabel max-flow / min-cut on a 4-connected grid.

Drop-in for the reference's GraphCut (GraphCut.pyx:29-437).  Reference ->
here:

  GraphCut(context, devices, dImg, lamda=60)
      dImg: an RGBA image on the device   GraphCut(context, devices, dImg, lamda)
      lamda: the pairwise weight's scale     dImg: an image (uint8 [H][W][3|4],
                                             packed uint32 RGBA [H][W], or a
                                             Buffer2D of either); its weights
                                             lamda / (0.1 sqrt(d2) + 0.05) are
                                             built on the device, int32
                                           GraphCut(context, devices, (W, H), tile=...)
                                             then upload(source, sink, right, down)
  cut(dExcess, bfs=5) -> None               cut(dExcess=None, bfs=None) -> flow value
      dExcess = source - sink, mutated      dExcess (int32 [H][W], source - sink) is
                                            optional and is not mutated
  push / relabel / startBfs /              the same hooks, driving the solve step by
  isCompleted(dExcess)                     step (isCompleted() -> bool)
  tilelistLoad / Bfs / Edges / Border      snapshots of the nearest per-tile sets
                                           (active, changed, holding excess,
                                           pending inflow) as IncrementalTileLists
  dHeight (device buffer)                   dHeight: heights as numpy [PH][PW]
  (no labels)                               labels(): uint8 [H][W], 1 = source side
                                            (the minimal source set)

Any width/height is accepted (the reference requires multiples of 32,
GraphCut.pyx:85-86); the grid is padded to whole tiles internally.
"""
import ctypes

import numpy as np

from gcmi_lib import (OPT_DEVICE_ROUNDS, OPT_GRAPH, OPT_LOCAL_RELABEL, PAIR_INV, TILES_ACTIVE,
                      TILES_BORDER, TILES_CHANGED, TILES_EXCESS, DeviceBuffer, check, host_ptr,
                      lib, ptr, rgb_image, rgb_model)

TILE_DEFAULT = (32, 32)
LAMDA_DEFAULT = 60    # the reference's pairwise scale (GraphCut.pyx:80)
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


def _image_of(dImg):
    """The uint8 [H][W][3|4] image a dImg argument holds, or None for a
    plain (width, height)."""
    a = dImg
    if hasattr(dImg, "get") and hasattr(dImg, "dim"):  # a Buffer2D
        if np.dtype(dImg.dtype) not in (np.dtype(np.uint32), np.dtype(np.int32)):
            return None
        a = dImg.get()
    if isinstance(a, np.ndarray):
        if a.ndim == 2 and a.dtype in (np.uint32, np.int32):  # packed RGBA, R in the low byte
            return np.ascontiguousarray(a).view(np.uint8).reshape(a.shape[0], a.shape[1], 4)
        if a.ndim == 3 and a.dtype == np.uint8 and a.shape[2] in (3, 4):
            return a
        raise ValueError("dImg: expected uint8 [H][W][3|4] or packed uint32 RGBA [H][W]")
    return None


class GraphCut:
    lWorksize = (16, 16)

    def __init__(self, context=None, devices=None, dim=None, lamda=None, tile=TILE_DEFAULT,
                 bfs=BFS_DEFAULT, sweeps=SWEEPS_DEFAULT):
        """dim: (width, height), or the reference's dImg -- an image whose
        pairwise weights lamda / (0.1 sqrt(d2) + 0.05) (graphcut.cl:1018-1031)
        are built on the device (lamda defaults to the reference's 60), the
        unaries then coming from cut(dExcess).  With a plain (width, height),
        lamda (if given) is the scale of later upload_rgb() calls."""
        if dim is None:
            raise ValueError("dim=(width, height) or an image is required")
        image = _image_of(dim)
        if image is not None:
            dim = (image.shape[1], image.shape[0])
            if lamda is None:
                lamda = LAMDA_DEFAULT
        elif hasattr(dim, "dim"):  # a Buffer2D of another type: its shape only
            dim = dim.dim
        self.lamda = None if lamda is None else float(lamda)
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
        self._stepping = False
        self.flow = None
        if image is not None:
            self.upload_rgb(image)

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
        self._stepping = False

    def upload_rgb(self, image, model=None):
        """Builds the capacities on the device from a uint8 [H][W][3] (or
        RGBA [H][W][4]) image: colour-distance unaries and pairwise weights
        (the reference builds its weights from the RGBA image on the device,
        graphcut.cl:1018-1131).  model: an RgbModel (gcmi_lib.rgb_model);
        default: the synthetic workloads' model."""
        a, ch = rgb_image(image)
        if a.shape[:2] != self.shape:
            raise ValueError("expected an image of shape %s" % (self.shape,))
        if model is None:
            model = rgb_model() if self.lamda is None else \
                rgb_model(pair_kind=PAIR_INV, pair_scale=self.lamda)
        check(lib().gcmi_cut_upload_rgb(self._h, host_ptr(a), ch, ctypes.byref(model)))
        self._right = self._down = None
        self._device_pairwise = True
        self._stepping = False

    def upload_gmm(self, image, fg, bg, scale=10.0, cap=1000.0, pairwise=None):
        """Capacities built on the device from a colour image with unaries
        from two colour GMMs (GMM objects or gcmi_lib.Gmm): source =
        rint(scale * min(max(score_bg, 0), cap)), sink the same under fg --
        the reference feeds GMM scores to the cut as its unaries
        (test_GraphCut.py:44-47).  Pairwise weights as upload_rgb builds
        them (pairwise: an RgbModel; default: the reference's weight with
        this cut's lamda, or the workloads' model without one)."""
        a, ch = rgb_image(image)
        if a.shape[:2] != self.shape:
            raise ValueError("expected an image of shape %s" % (self.shape,))
        if pairwise is None:
            pairwise = rgb_model() if self.lamda is None else \
                rgb_model(pair_kind=PAIR_INV, pair_scale=self.lamda)
        f = getattr(fg, "model", fg)
        b = getattr(bg, "model", bg)
        check(lib().gcmi_cut_upload_rgb_gmm(self._h, host_ptr(a), ch, ctypes.byref(pairwise),
                                            ctypes.byref(f), ctypes.byref(b), float(scale),
                                            float(cap)))
        self._right = self._down = None
        self._device_pairwise = True
        self._stepping = False

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
        self._stepping = False

    def paint(self, points, width, source, sink):
        """Warm-start edit by a brush stroke rasterised on the device: every
        pixel whose centre lies within width/2 of the polyline `points`
        ([n][2] (x, y), n <= 16) gets capacities (source, sink)."""
        xy = _points(points)
        check(lib().gcmi_cut_paint(self._h, xy.shape[0], host_ptr(xy), float(width),
                                   int(source), int(sink)))
        self._stepping = False

    def upload_device(self, d_source, d_sink, d_right, d_down):
        check(lib().gcmi_cut_upload_device(self._h, ptr(d_source), ptr(d_sink), ptr(d_right),
                                           ptr(d_down)))

    # ------------------------------------------------------------------ cut
    def _set_excess(self, dExcess):
        """New unaries from a signed excess image (source - sink)."""
        if self._right is None and self._device_pairwise:
            # pairwise weights built on the device: new unaries for every
            # pixel, cut from scratch
            e = np.asarray(dExcess.get() if hasattr(dExcess, "get") else dExcess)
            e = _i32(e, self.shape).reshape(-1)
            self.forget()
            self.edit(np.arange(e.size, dtype=np.int32), np.maximum(e, 0),
                      np.maximum(-e.astype(np.int64), 0).astype(np.int32))
            return
        if self._right is None:
            raise ValueError("upload_pairwise() before cut(dExcess)")
        e = np.asarray(dExcess)
        if hasattr(dExcess, "get"):
            e = dExcess.get()
        e = _i32(e, self.shape)
        self.upload(np.maximum(e, 0), np.maximum(-e.astype(np.int64), 0).astype(np.int32),
                    self._right, self._down)

    def cut(self, dExcess=None, bfs=None, sweeps=None, max_rounds=0):
        """Run push-relabel to termination; returns the max-flow value.

        dExcess (optional): int32 [H][W] signed excess, source - sink as in
        test_GraphCut.py:44-47; it sets source = max(e, 0), sink = max(-e, 0)
        with the pairwise weights of the last upload / upload_pairwise (or
        those built on the device from the image)."""
        if dExcess is not None:
            self._set_excess(dExcess)
        self._stepping = False
        bfs = self.bfs if bfs is None else int(bfs)
        sweeps = self.sweeps if sweeps is None else int(sweeps)
        flow = ctypes.c_int64()
        check(lib().gcmi_cut_run(self._h, sweeps, bfs, int(max_rounds), ctypes.byref(flow)))
        self.flow = int(flow.value)
        return self.flow

    # ---------------------------------------------------------- step hooks
    # The reference's hooks (GraphCut.pyx:187-346), for callers that drive
    # the solve themselves, e.g. its own cut loop: startBfs, then push and
    # relabel, with startBfs + isCompleted every few rounds.  dExcess, if
    # given to the first hook of a solve, sets the unaries as cut(dExcess)
    # does; later hooks of the same solve ignore it (the excess is part of
    # the cut's own state).  The finished solve's flow and labels equal
    # cut()'s.
    def _step(self, dExcess):
        if dExcess is not None and not self._stepping:
            self._set_excess(dExcess)
        self._stepping = True

    def push(self, dExcess=None):
        """One push round over the tiles holding active pixels."""
        self._step(dExcess)
        check(lib().gcmi_cut_push(self._h, int(self.sweeps)))

    def relabel(self, dExcess=None):
        """A relabel of the tiles that changed since their distances were
        computed (every tile before this solve's first relabel from
        scratch); returns the number of tiles left active."""
        self._step(dExcess)
        n = ctypes.c_int32()
        check(lib().gcmi_cut_relabel(self._h, ctypes.byref(n)))
        return n.value

    def startBfs(self, dExcess=None):
        """A global relabel over every tile; returns the tiles left active."""
        self._step(dExcess)
        n = ctypes.c_int32()
        check(lib().gcmi_cut_start_bfs(self._h, ctypes.byref(n)))
        return n.value

    def isCompleted(self, dExcess=None):
        """True right after a relabel that left no active tile (the solve is
        then finished: flow, labels()) or when the last solve is complete."""
        done = ctypes.c_int()
        flow = ctypes.c_int64()
        check(lib().gcmi_cut_is_completed(self._h, ctypes.byref(done), ctypes.byref(flow)))
        if done.value:
            self.flow = int(flow.value)
            self._stepping = False
        return bool(done.value)

    def _tile_list(self, which):
        from IncrementalTileList import IncrementalTileList, Operator
        pw, ph, t = self._geometry()
        tl = IncrementalTileList(None, None, (pw, ph), (t, t))
        # (its stamps were initialised on the library's stream; the copy
        # runs on the cut's)
        check(lib().gcmi_stream_sync(None))
        check(lib().gcmi_cut_tile_flags(self._h, which, ptr(tl.d_tiles)))
        tl.iteration = 1
        tl.build(Operator.GT, 0)
        return tl

    @property
    def tilelistLoad(self):
        """Tiles holding active pixels (pushed next), as an
        IncrementalTileList snapshot (GraphCut.pyx:142)."""
        return self._tile_list(TILES_ACTIVE)

    @property
    def tilelistBorder(self):
        """Tiles with inflow pending from a neighbour (GraphCut.pyx:148)."""
        return self._tile_list(TILES_BORDER)

    @property
    def tilelistBfs(self):
        """Tiles the next relabel recomputes: changed since their distances
        were computed (GraphCut.pyx:144)."""
        return self._tile_list(TILES_CHANGED)

    @property
    def tilelistEdges(self):
        """Tiles holding positive excess, where the minimal source set grows
        from (the nearest counterpart of GraphCut.pyx:146)."""
        return self._tile_list(TILES_EXCESS)

    def labels(self):
        """uint8 [H][W]: 1 for the minimal source (foreground) set."""
        out = np.empty(self.shape, np.uint8)
        check(lib().gcmi_cut_labels(self._h, host_ptr(out)))
        return out

    def labels_packed(self, out=None):
        """uint32 words, bit p % 32 of word p // 32 = label of pixel p.  `out`
        (optional): a C-contiguous uint32 array of that many words to copy
        into, e.g. page-locked (gcmi_lib.pinned_empty), which the copy
        reaches faster than ordinary host memory."""
        n = (self.shape[0] * self.shape[1] + 31) // 32
        if out is None:
            out = np.empty(n, np.uint32)
        elif out.dtype != np.uint32 or out.size != n or not out.flags.c_contiguous:
            raise ValueError("out: %d C-contiguous uint32 words" % n)
        check(lib().gcmi_cut_labels_packed(self._h, host_ptr(out)))
        return out

    def forget(self):
        """Drop the solved state: the next cut() starts from the stored
        capacities (every edit included) instead of reflowing warm."""
        check(lib().gcmi_cut_forget(self._h))
        self._stepping = False

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

    def _geometry(self):
        g = (ctypes.c_int * 3)()
        check(lib().gcmi_cut_geometry(self._h, g))
        return g[0], g[1], g[2]

    @property
    def dHeight(self):
        """Heights (BFS distances) of the padded grid, [PH][PW] int32."""
        g = self._geometry()
        out = np.empty((g[1], g[0]), np.int32)
        check(lib().gcmi_cut_heights(self._h, host_ptr(out)))
        return out

    @property
    def maxHeight(self):
        g = (ctypes.c_int * 3)()
        check(lib().gcmi_cut_geometry(self._h, g))
        return g[0] * g[1]

    def execute(self, queue=None, input="height"):
        """A state field as an image, the counterpart of the reference's
        display filter (GraphCut.execute / filter_transpose,
        GraphCut.pyx:423-437): int32 [PH][PW] row-major over the padded grid
        of "height", "excess", "up", "down", "left" or "right" (residual
        capacities toward that neighbour)."""
        from gcmi_lib import FIELDS
        if input not in FIELDS:
            raise ValueError("input: one of %s" % sorted(FIELDS))
        g = self._geometry()
        out = np.empty((g[1], g[0]), np.int32)
        check(lib().gcmi_cut_field(self._h, FIELDS[input], host_ptr(out)))
        return out

    def reset(self):
        """Kept for API parity (GraphCut.pyx:420-421): every cut() starts
        from the uploaded capacities, so there is nothing to reset."""
        return None
