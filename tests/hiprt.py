"""Device buffers for the tests of the device-pointer entry points, through
the HIP runtime that libnbc.so itself links (loaded after it, so the process
holds one runtime instance)."""
import ctypes as C

import numpy as np

import nbc

_rt = None


def rt():
    global _rt
    if _rt is None:
        nbc.lib()  # maps libnbc.so and its libamdhip64 first
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipSetDevice.argtypes = [C.c_int]
        for f in (L.hipMalloc, L.hipFree, L.hipMemcpy, L.hipDeviceSynchronize, L.hipSetDevice):
            f.restype = C.c_int
        _rt = L
    return _rt


H2D, D2H = 1, 2


class DeviceArray:
    """A device allocation with the shape and dtype of a host array."""

    def __init__(self, shape, dtype):
        self.shape, self.dtype = tuple(np.atleast_1d(shape)), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        assert rt().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)) == 0, "hipMalloc failed"

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        assert rt().hipMemcpy(d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return d

    def host(self):
        a = np.empty(self.shape, self.dtype)
        assert rt().hipDeviceSynchronize() == 0
        assert rt().hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, D2H) == 0
        return a

    def free(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
