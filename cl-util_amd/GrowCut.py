"""GrowCut: cellular-automaton segmentation on the tile-list runtime.

Drop-in for the reference's GrowCut (GrowCut.py:26-153): the same
constructor (context, devices, img, neighbourhood, weight), label(d_points,
n_points, label) and evolve(iterations), with isComplete.  The image is a
uint8 [H][W][3|4] array (or packed uint32 RGBA, or a Buffer2D of either);
the weight is one of the reference's functions, named by its WEIGHT_*
attribute.  libgcmi runs the automaton in float32 on 16x16 tiles restricted
to the tiles the last step changed, and reads the list's length back every
few steps rather than after each.
"""
import ctypes
import sys

import numpy as np

from gcmi_lib import check, host_ptr, lib

LINEAR, POW2, POW3, POW1_5, SQRT = 0, 1, 2, 3, 4


def _image(img):
    a = img.get() if hasattr(img, "get") and hasattr(img, "dim") else img
    a = np.asarray(a)
    if a.ndim == 2 and a.dtype in (np.uint32, np.int32):
        a = np.ascontiguousarray(a).view(np.uint8).reshape(a.shape[0], a.shape[1], 4)
    a = np.ascontiguousarray(a, np.uint8)
    if a.ndim != 3 or a.shape[2] not in (3, 4):
        raise ValueError("img: uint8 [H][W][3|4] or packed uint32 RGBA [H][W]")
    return a


class GrowCut:
    class NEIGHBOURHOOD:
        VON_NEUMANN = 0

    # the reference's weight functions g(X) (GrowCut.py:30-34), by name
    WEIGHT_DEFAULT = LINEAR
    WEIGHT_POW2 = POW2
    WEIGHT_POW3 = POW3
    WEIGHT_POW1_5 = POW1_5
    WEIGHT_POW_SQRT = SQRT

    TILEW = TILEH = 16

    def __init__(self, context=None, devices=None, img=None, neighbourhood=0, weight=None):
        if neighbourhood != GrowCut.NEIGHBOURHOOD.VON_NEUMANN:
            raise NotImplementedError("only the von Neumann neighbourhood")
        a = _image(img)
        self.dim = (a.shape[1], a.shape[0])
        self.weight = LINEAR if weight is None else int(weight)
        h = ctypes.c_void_p()
        check(lib().gcmi_growcut_create(self.dim[0], self.dim[1], host_ptr(a), a.shape[2],
                                        self.weight, ctypes.byref(h)))
        self._h = h.value
        self.isComplete = False
        self.iterations = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().gcmi_growcut_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def label(self, d_points, n_points=None, label=1):
        """Seed pixels (flat indices y * W + x; a host array or anything
        with .get()) with a label and strength 1."""
        pts = d_points.get() if hasattr(d_points, "get") else d_points
        pts = np.ascontiguousarray(pts, np.int32).reshape(-1)
        if n_points is not None:
            pts = pts[: int(n_points)]
        check(lib().gcmi_growcut_label(self._h, pts.size, host_ptr(pts), int(label)))
        self.isComplete = False

    def evolve(self, iterations=sys.maxsize):
        """Up to `iterations` steps (until nothing changes); isComplete once
        a step found nothing left to update."""
        n = int(min(iterations, 2**31 - 1))
        done = ctypes.c_int()
        steps = ctypes.c_int()
        check(lib().gcmi_growcut_evolve(self._h, n, ctypes.byref(done), ctypes.byref(steps)))
        self.isComplete = bool(done.value)
        self.iterations += steps.value

    def labels(self):
        out = np.empty((self.dim[1], self.dim[0]), np.uint8)
        check(lib().gcmi_growcut_read(self._h, host_ptr(out), None, None, None))
        return out

    def strength(self):
        out = np.empty((self.dim[1], self.dim[0]), np.float32)
        check(lib().gcmi_growcut_read(self._h, None, host_ptr(out), None, None))
        return out

    def tiles(self):
        """Tile stamps [tilesH][tilesW] and the current iteration."""
        tw = (self.dim[0] + 15) // 16
        th = (self.dim[1] + 15) // 16
        out = np.empty((th, tw), np.int32)
        it = ctypes.c_int32()
        check(lib().gcmi_growcut_read(self._h, None, None, host_ptr(out), ctypes.byref(it)))
        return out, it.value
