"""Device buffers for GPU tests, from the HIP runtime the library itself
loads (plain hipMalloc / hipMemcpy through ctypes)."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
    return _hip


class DeviceBuffer:
    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), C.c_size_t(max(self.nbytes, 1))) == 0

    @classmethod
    def of(cls, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def download(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert out.nbytes <= self.nbytes
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes), 2) == 0
        return out

    @property
    def value(self) -> int:
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        self.free()
