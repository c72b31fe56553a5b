"""Device buffers, streams and events over the C ABI (no torch needed).

The stand-in for the reference's TypedBuffer / DeviceStopwatch plumbing
(typedbuffer.h:19-146, pk.h:1192-1223).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call, lib


def device_count() -> int:
    n = C.c_int(0)
    try:
        call("palz_device_count", C.byref(n))
    except Exception:
        return 0
    return n.value


class DeviceBuffer:
    """An owning device allocation with numpy round trips."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        call("palz_malloc", C.byref(p), C.c_size_t(max(self.nbytes, 1)))
        self.ptr = p.value or 0

    @classmethod
    def from_array(cls, a: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a, stream)
        return b

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes, "upload larger than the buffer"
        call("palz_memcpy_h2d", C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p),
             C.c_size_t(a.nbytes), _sp(stream))

    def download(self, dtype, count: int, stream=None, offset_bytes: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert offset_bytes + out.nbytes <= self.nbytes, "download past the buffer"
        call("palz_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset_bytes),
             C.c_size_t(out.nbytes), _sp(stream))
        return out

    def zero(self, stream=None) -> None:
        call("palz_memset", C.c_void_p(self.ptr), 0, C.c_size_t(self.nbytes), _sp(stream))

    def free(self) -> None:
        if self.ptr:
            lib().palz_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download_ptr(ptr: int, dtype, count: int, stream=None) -> np.ndarray:
    out = np.empty(count, dtype=dtype)
    if out.nbytes:
        call("palz_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr),
             C.c_size_t(out.nbytes), _sp(stream))
    return out


class Stream:
    def __init__(self):
        p = C.c_void_p()
        call("palz_stream_create", C.byref(p))
        self.handle = p.value

    def sync(self) -> None:
        call("palz_stream_sync", C.c_void_p(self.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().palz_stream_destroy(C.c_void_p(self.handle))
                self.handle = None
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        call("palz_event_create", C.byref(p))
        self.handle = p.value

    def record(self, stream=None) -> None:
        call("palz_event_record", C.c_void_p(self.handle), _sp(stream))

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float(0.0)
        call("palz_event_elapsed_ms", C.byref(ms), C.c_void_p(self.handle), C.c_void_p(later.handle))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                lib().palz_event_destroy(C.c_void_p(self.handle))
                self.handle = None
        except Exception:
            pass


def sync() -> None:
    call("palz_device_sync")


def _sp(stream) -> C.c_void_p:
    if stream is None:
        return C.c_void_p(None)
    if isinstance(stream, Stream):
        return C.c_void_p(stream.handle)
    return C.c_void_p(stream)
