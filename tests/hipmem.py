"""Device buffers for GPU tests that call entry points taking device
pointers.  They are made through the HIP runtime libtg.so is linked against:
dlopen of its soname returns the copy the process already holds, so the test
and the library share one runtime.  (PyTorch ships a runtime of its own; when
the library's is loaded first, torch's second copy finds no device.)"""
import ctypes as C

import numpy as np

from discrete_adjoint_amd import _lib

_H2D, _D2H = 1, 2
_rt = None


def runtime():
    global _rt
    if _rt is None:
        _lib.load()
        rt = C.CDLL("libamdhip64.so.7")
        rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        rt.hipMalloc.restype = C.c_int
        rt.hipFree.argtypes = [C.c_void_p]
        rt.hipFree.restype = C.c_int
        rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        rt.hipMemcpy.restype = C.c_int
        rt.hipDeviceSynchronize.restype = C.c_int
        _rt = rt
    return _rt


class DeviceArray:
    """n elements of dtype in device memory (hipMalloc); freed by free() or
    on leaving a with block."""

    def __init__(self, n, dtype=np.float64):
        self.n, self.dtype = int(n), np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        rc = runtime().hipMalloc(C.byref(p), max(self.nbytes, 1))
        if rc != 0:
            raise RuntimeError(f"hipMalloc({self.nbytes}) returned {rc}")
        self.ptr = p.value

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        rc = runtime().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, _H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D returned {rc}")

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        rc = runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H returned {rc}")
        return out

    def free(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


def synchronize():
    rc = runtime().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize returned {rc}")
