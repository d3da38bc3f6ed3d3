"""Minimal HIP runtime helpers for GPU tests (device buffers, copies), through
ctypes on the HIP runtime that libcavityflow_hip.so already loaded."""
import ctypes

import numpy as np

_L = None


def rt():
    global _L
    if _L is None:
        _L = ctypes.CDLL("libamdhip64.so.7")
        for fn in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize",
                   "hipStreamCreate", "hipStreamDestroy", "hipStreamSynchronize"):
            getattr(_L, fn).restype = ctypes.c_int
    return _L


class DeviceBuffer:
    """Raw device allocation; passes as a device buffer (has data_ptr())."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        assert rt().hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.nbytes)) == 0

    def data_ptr(self):
        return self.ptr.value

    def upload(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        assert rt().hipMemcpy(self.ptr, arr.ctypes.data_as(ctypes.c_void_p),
                              ctypes.c_size_t(arr.nbytes), 1) == 0

    def download(self, dtype, count) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        assert rt().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                              ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        if self.ptr.value:
            rt().hipDeviceSynchronize()
            rt().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def stream_create():
    """A new (blocking, default-priority) HIP stream handle as an int."""
    st = ctypes.c_void_p()
    assert rt().hipStreamCreate(ctypes.byref(st)) == 0
    return st.value


def stream_destroy(st):
    rt().hipStreamSynchronize(ctypes.c_void_p(st))
    rt().hipStreamDestroy(ctypes.c_void_p(st))
