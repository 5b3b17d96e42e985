"""Device buffers for the GPU tests, from the same HIP runtime libed_hip.so links
(libamdhip64.so.7).  torch ships its own HIP runtime; mixing both in one process
fails to initialise the second, so the tests never touch the GPU through torch."""
import ctypes
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        import ed_solver_amd as ed
        ed.lib()   # loads libamdhip64.so.7 as a dependency
        for cand in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so.7"):
            try:
                _hip = ctypes.CDLL(cand)
                break
            except OSError:
                continue
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DevBuf:
    def __init__(self, nbytes: int):
        self.ptr = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.ptr), max(1, nbytes)) == 0
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DevBuf":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        assert hip().hipMemcpy(b.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0
        return b

    def to_array(self, dtype, n) -> np.ndarray:
        out = np.empty(n, dtype=dtype)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, 2) == 0
        return out

    @property
    def addr(self) -> int:
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()
