"""StreamCompact: ascending indices of nonzero flags.

Drop-in for StreamCompact.pyx:22-82 (same constructor, flagFactory,
listFactory and compact).  The reference copies the flags into the list,
scans in place and scatters in place (a read/write race between work-items,
streamcompact.cl:13); libgcmi reads the flags once and writes the list
through a wave ballot, so the flags are left untouched.
"""
from gcmi_lib import DeviceBuffer, check, lib, ptr
from PrefixSum import PrefixSum

szInt = 4


class StreamCompact:
    def __init__(self, context=None, devices=None, capacity=0):
        self.capacity = int(capacity)
        self.prefix_sum = PrefixSum(context, devices, capacity)

    def flagFactory(self, length=None):
        if length is None:
            length = self.capacity
        elif length > self.capacity:
            raise ValueError("length > self.capacity: {0}, {1}".format(length, self.capacity))
        return DeviceBuffer(length * szInt)

    def listFactory(self, length=None):
        if length is None:
            length = self.capacity
        elif length > self.capacity:
            raise ValueError("length > self.capacity: {0}, {1}".format(length, self.capacity))
        return self.prefix_sum.factory(length)

    def compact(self, dFlags, dList, dLength, length=None):
        if length is None:
            length = dFlags.size // szInt
        elif length > self.capacity:
            raise ValueError("length > self.capacity: {0}, {1}".format(length, self.capacity))
        check(lib().gcmi_stream_compact(ptr(dFlags), ptr(dList), ptr(dLength), int(length), None))
