"""Device memory for tests that hand device pointers to the C ABI, from
the HIP runtime libsonic_amd.so itself links (torch carries a runtime of
its own, which cannot share the device with it in one process)."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from sonicsolver_amd import load
        load()   # the library first, so its runtime is the one found
        lib = ctypes.CDLL("libamdhip64.so.7")
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipDeviceSynchronize.argtypes = []
        _hip = lib
    return _hip


class DeviceArray:
    """A float64 device array of `count` values."""

    def __init__(self, count: int):
        self.count = int(count)
        p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(p), 8 * self.count) == 0
        self.ptr = p.value

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == self.count
        assert hip().hipMemcpy(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p),
                               8 * self.count, 1) == 0
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.count)
        assert hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                               8 * self.count, 2) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()
