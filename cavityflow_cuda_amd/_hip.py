"""The few HIP runtime calls the host drivers need beyond the C ABI: streams
and events for ordering work across two streams (ctypes on the HIP runtime
that libcavityflow_hip.so itself loads). Streams are created with the
default priority."""
from __future__ import annotations

import ctypes

_L = None

HIP_STREAM_NON_BLOCKING = 0x01
HIP_EVENT_DISABLE_TIMING = 0x02


def rt() -> ctypes.CDLL:
    global _L
    if _L is None:
        from ._lib import lib
        lib()   # loads the HIP runtime as a dependency of the library
        L = ctypes.CDLL("libamdhip64.so.7")
        for fn in ("hipStreamCreateWithFlags", "hipStreamDestroy", "hipStreamSynchronize",
                   "hipEventCreateWithFlags", "hipEventDestroy", "hipEventRecord",
                   "hipStreamWaitEvent", "hipSetDevice", "hipGetErrorString"):
            getattr(L, fn).restype = ctypes.c_int if fn != "hipGetErrorString" else ctypes.c_char_p
        _L = L
    return _L


def _ok(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what}: {rt().hipGetErrorString(rc).decode()} ({rc})")


class Stream:
    """A non-blocking HIP stream (no implicit ordering with the legacy
    default stream)."""

    def __init__(self, device: int = 0):
        _ok(rt().hipSetDevice(int(device)), "hipSetDevice")
        h = ctypes.c_void_p()
        _ok(rt().hipStreamCreateWithFlags(ctypes.byref(h), HIP_STREAM_NON_BLOCKING),
            "hipStreamCreateWithFlags")
        self.handle = h.value

    def synchronize(self):
        _ok(rt().hipStreamSynchronize(ctypes.c_void_p(self.handle)), "hipStreamSynchronize")

    def close(self):
        if self.handle:
            rt().hipStreamSynchronize(ctypes.c_void_p(self.handle))
            rt().hipStreamDestroy(ctypes.c_void_p(self.handle))
            self.handle = None


class Event:
    """An ordering-only event (timing disabled)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _ok(rt().hipEventCreateWithFlags(ctypes.byref(h), HIP_EVENT_DISABLE_TIMING),
            "hipEventCreateWithFlags")
        self.handle = h.value

    def record(self, stream):
        _ok(rt().hipEventRecord(ctypes.c_void_p(self.handle), ctypes.c_void_p(stream)),
            "hipEventRecord")

    def wait_on(self, stream):
        """Work queued on `stream` after this call waits for the event."""
        _ok(rt().hipStreamWaitEvent(ctypes.c_void_p(stream), ctypes.c_void_p(self.handle), 0),
            "hipStreamWaitEvent")

    def close(self):
        if self.handle:
            rt().hipEventDestroy(ctypes.c_void_p(self.handle))
            self.handle = None
