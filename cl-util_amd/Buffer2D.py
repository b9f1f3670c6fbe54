"""Buffer2D: a device buffer with a 2-D (width, height) shape and dtype.

Mirrors Buffer2D.py:8-34 of the reference (a cl.Buffer subclass with .dim and
.dtype).  The OpenCL context and flags arguments are accepted and ignored;
rows are not padded (the reference pads width to 32 elements for coalescing;
the kernels here use their own tile-blocked layout instead).
"""
import numpy as np

from gcmi_lib import DeviceBuffer


class Buffer2D(DeviceBuffer):
    def __init__(self, context=None, flags=None, dim=None, dtype=None, hostbuf=None):
        if hostbuf is not None:
            hostbuf = np.ascontiguousarray(hostbuf)
            dim = (hostbuf.shape[1], hostbuf.shape[0])
            dtype = hostbuf.dtype
        elif dim is None or dtype is None:
            raise ValueError("Dimensions and datatype required")
        self.dim = (int(dim[0]), int(dim[1]))
        self.dtype = np.dtype(dtype)
        DeviceBuffer.__init__(self, self.dtype.itemsize * self.dim[0] * self.dim[1])
        if hostbuf is not None:
            self.upload(hostbuf)

    @staticmethod
    def fromBuffer(buffer, dim, dtype):
        buffer.dim = (int(dim[0]), int(dim[1]))
        buffer.dtype = np.dtype(dtype)
        return buffer

    def get(self):
        """Host copy as [height][width]."""
        return self.download(self.dtype, self.dim[0] * self.dim[1]).reshape(
            self.dim[1], self.dim[0])
