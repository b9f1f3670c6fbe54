"""IncrementalTileList: per-tile iteration stamps and the compacted list of
tiles whose stamp satisfies a predicate.

Drop-in for IncrementalTileList.pyx:37-208 (same constructor, enums, members
and methods; the OpenCL context/devices arguments are accepted and ignored).
`build` evaluates the predicate inside libgcmi's compaction kernel, so the
reference's separate flag launch, copy and scan (IncrementalTileList.pyx:
119-136, StreamCompact.pyx:77-81) are one launch; `length` still reads the
device count back lazily, as the reference does (:106-112).
"""
import numpy as np

from Buffer2D import Buffer2D
from gcmi_lib import DeviceBuffer, check, lib, ptr

szInt = 4


class Operator:
    EQUAL = 0
    GT = 1
    LT = 2
    GTE = 3
    LTE = 4


class Logical:
    AND = 0
    OR = 1


class IncrementalTileList:
    def __init__(self, context=None, devices=None, global_dim=None, tiles_dim=None):
        if global_dim is None or tiles_dim is None:
            raise ValueError("global_dim and tiles_dim are required")
        self.dim = (int(global_dim[0]) // int(tiles_dim[0]),
                    int(global_dim[1]) // int(tiles_dim[1]))
        self.global_dim = (int(global_dim[0]), int(global_dim[1]))
        self.tiles_dim = (int(tiles_dim[0]), int(tiles_dim[1]))
        n = self.dim[0] * self.dim[1]
        self._n = n
        self.h_length = np.zeros(1, np.int32)
        self.d_length = DeviceBuffer(szInt)
        self.d_list = DeviceBuffer(max(n, 1) * szInt)
        self.d_tiles = Buffer2D.fromBuffer(DeviceBuffer(max(n, 1) * szInt), self.dim, np.int32)
        self.d_flags = Buffer2D.fromBuffer(DeviceBuffer(max(n, 1) * szInt), self.dim, np.int32)
        self.is_dirty = True
        self.init_iteration = -1
        self.iteration = 0
        self.reset()

    @property
    def length(self):
        if self.is_dirty:
            self.h_length = self.d_length.download(np.int32, 1)
            self.is_dirty = False
        return int(self.h_length[0])

    def increment(self):
        self.iteration += 1
        return self.iteration

    def build(self, operator=None, operand=None):
        if operator is None:
            operator = Operator.EQUAL
        if operand is None:
            operand = self.iteration
        check(lib().gcmi_tiles_build(ptr(self.d_tiles), self._n, int(operator), int(operand),
                                     ptr(self.d_list), ptr(self.d_length), None))
        self.is_dirty = True

    def incorporate(self, d_tiles2, operator1, operand1, operator2, operand2, logical):
        """Stamp the current iteration where (tiles OP1 a) LOGICAL (tiles2 OP2 b)."""
        check(lib().gcmi_tiles_increment_logical(
            ptr(self.d_tiles), ptr(d_tiles2), self._n, int(operator1), int(operator2),
            int(operand1), int(operand2), int(logical), int(self.iteration), None))

    def buildLogical(self, d_tiles2, operator1, operand1, operator2, operand2, logical):
        check(lib().gcmi_tiles_build_logical(
            ptr(self.d_tiles), ptr(d_tiles2), self._n, int(operator1), int(operator2),
            int(operand1), int(operand2), int(logical), ptr(self.d_list), ptr(self.d_length),
            None))
        self.is_dirty = True

    def reset(self):
        self.iteration = self.init_iteration + 1
        check(lib().gcmi_tiles_fill(ptr(self.d_tiles), self._n, int(self.init_iteration), None))

    def execute(self, queue=None, input=None):
        """The visualisation filter (IncrementalTileList.pyx:191-208): a
        Buffer2D of uint32 [H][W] pixels, opaque red (0xFF0000FF) where the
        pixel's tile in `input` (default: this list) holds input's
        iteration, opaque black elsewhere."""
        src = self if input is None else input
        out = Buffer2D(None, None, src.global_dim, np.uint32)
        check(lib().gcmi_tiles_filter(ptr(src.d_tiles), src.dim[0], src.dim[1], src.tiles_dim[0],
                                      src.tiles_dim[1], src.global_dim[0], src.global_dim[1],
                                      int(src.iteration), ptr(out), None))
        check(lib().gcmi_stream_sync(None))
        return out

    # host conveniences (the reference reads these with cl.enqueue_copy)
    def tiles(self):
        """Stamps as a [tilesH][tilesW] int32 array."""
        return self.d_tiles.download(np.int32, self._n).reshape(self.dim[1], self.dim[0])

    def set_tiles(self, stamps):
        a = np.ascontiguousarray(stamps, np.int32).reshape(-1)
        if a.size != self._n:
            raise ValueError("expected %d stamps" % self._n)
        self.d_tiles.upload(a)

    def list(self):
        """The current list, d_list[:length], as int32."""
        n = self.length
        return self.d_list.download(np.int32, n) if n else np.zeros(0, np.int32)
