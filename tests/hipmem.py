"""Device buffers for GPU tests through the HIP runtime the library itself
links (/opt/rocm's libamdhip64): a second HIP runtime in the process, such
as torch's bundled one, cannot initialise next to it."""
import ctypes as C

import numpy as np

import nigiri_amd as na

_H2D, _D2H = 1, 2


def _hip():
    na.lib()  # loads the library and with it its HIP runtime
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DeviceArray:
    """A device allocation of n elements of a numpy dtype."""

    def __init__(self, n, dtype):
        self.h = _hip()
        self.n, self.dtype = n, np.dtype(dtype)
        self.ptr = C.c_void_p()
        if self.h.hipMalloc(C.byref(self.ptr), max(1, n * self.dtype.itemsize)) != 0:
            raise RuntimeError("hipMalloc failed")

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        if self.h.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, _H2D) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    def download(self):
        out = np.empty(self.n, self.dtype)
        self.h.hipDeviceSynchronize()
        if self.h.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, _D2H) != 0:
            raise RuntimeError("hipMemcpy D2H failed")
        return out

    def free(self):
        if self.ptr:
            self.h.hipFree(self.ptr)
            self.ptr = C.c_void_p()
