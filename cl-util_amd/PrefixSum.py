"""PrefixSum: in-place exclusive scan on the device.

Drop-in for PrefixSum.pyx:25-129 (same constructor, factory and scan; the
OpenCL context/devices arguments are accepted and ignored).  The reference's
recursive Blelloch scan with one blocking wait per level is replaced by
libgcmi's single-launch block scan (multi-block above 16,384 elements).
"""
import ctypes

from gcmi_lib import DeviceBuffer, check, lib, ptr

LEN_WORKGROUP = 256
ELEMENTS_PER_THREAD = 2
ELEMENTS_PER_WORKGROUP = ELEMENTS_PER_THREAD * LEN_WORKGROUP
szInt = 4


def pow2gt(x):
    """Smallest power of two >= x (clutil.pyx:15-20)."""
    i = 1
    while i < x:
        i <<= 1
    return i


def roundUp(size, multiple):
    r = size % multiple
    return size if r == 0 else size + multiple - r


class PrefixSum:
    def __init__(self, context=None, devices=None, capacity=0, profile=False):
        """profile: each scan waits for its launches and adds their device
        time (ns, HIP events) to `elapsed`, as the reference does when built
        with PROFILE_GPU (PrefixSum.pyx:100-101)."""
        self.capacity = roundUp(int(capacity), ELEMENTS_PER_WORKGROUP)
        self.elapsed = 0
        self.profile = bool(profile)

    def factory(self, length=None):
        if length is None:
            length = self.capacity
        elif length > self.capacity:
            raise ValueError("length > self.capacity: {0}, {1}".format(length, self.capacity))
        return DeviceBuffer(pow2gt(length) * szInt)

    def scan(self, dArray, dTotal, length=None):
        """In-place exclusive scan of dArray[0:length]; total -> dTotal[0]."""
        if length is None:
            length = dArray.size // szInt
        if length > ELEMENTS_PER_WORKGROUP and length > self.capacity:
            raise ValueError("length > self.capacity: {0}, {1}".format(length, self.capacity))
        if self.profile:
            ns = ctypes.c_int64()
            check(lib().gcmi_prefix_sum_timed(ptr(dArray), ptr(dTotal), int(length), None,
                                              ctypes.byref(ns)))
            self.elapsed += ns.value
        else:
            check(lib().gcmi_prefix_sum(ptr(dArray), ptr(dTotal), int(length), None))
