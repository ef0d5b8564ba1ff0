"""Device buffers for the host side: a minimal array handle over hipMalloc'd
memory, so the package needs no other GPU framework.  Any other allocator
(e.g. a torch tensor's ``data_ptr()``) can be passed to the C ABI instead.
"""
import ctypes as C

import numpy as np

from . import _lib


class DeviceArray:
    """A contiguous device allocation with a NumPy shape and dtype."""

    def __init__(self, shape, dtype=np.float64, ptr=None, owner=None):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._owner = owner
        if ptr is None:
            p = C.c_void_p()
            _lib.call("f2d_malloc", C.byref(p), max(self.nbytes, 1))
            self.ptr = p.value
            self._own = True
        else:
            self.ptr = ptr
            self._own = False

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype if dtype is not None else a.dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @classmethod
    def zeros(cls, shape, dtype=np.float64):
        return cls.from_host(np.zeros(shape, dtype=dtype))

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch: %d vs %d bytes" % (a.nbytes, self.nbytes))
        _lib.call("f2d_memcpy_h2d", self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes)

    def get(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _lib.call("f2d_memcpy_d2h", out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes)
        return out

    def view(self, index):
        """Device view of the sub-array ``index`` along the first axis."""
        sub = self.shape[1:]
        step = int(np.prod(sub)) * self.dtype.itemsize
        return DeviceArray(sub, self.dtype, ptr=self.ptr + index * step, owner=self)

    def copy_from(self, other):
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _lib.call("f2d_memcpy_d2d", self.ptr, other.ptr, self.nbytes)

    def free(self):
        if self._own and self.ptr:
            _lib.load().f2d_free(self.ptr)
        self.ptr = None
        self._own = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __repr__(self):
        return "DeviceArray(shape=%s, dtype=%s)" % (self.shape, self.dtype.name)


def ptr(a):
    """Device address of a DeviceArray (or an integer address, or None)."""
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return int(a)
