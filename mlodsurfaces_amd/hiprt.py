"""Minimal ctypes access to the HIP runtime (streams, events, device memory) for the bench
and the serving loop, so that timing uses HIP events on the stream the kernels run on."""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os

_hip = None


def hip():
    global _hip
    if _hip is None:
        cands = [os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"),
                 ctypes.util.find_library("amdhip64") or "libamdhip64.so"]
        for c in cands:
            try:
                _hip = C.CDLL(c)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        h = _hip
        h.hipGetErrorString.restype = C.c_char_p
        h.hipGetErrorString.argtypes = [C.c_int]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        for f in ("hipStreamCreate", "hipEventCreate"):
            getattr(h, f).argtypes = [C.POINTER(C.c_void_p)]
        for f in ("hipStreamSynchronize", "hipStreamDestroy", "hipEventSynchronize", "hipEventDestroy", "hipFree"):
            getattr(h, f).argtypes = [C.c_void_p]
        h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return _hip


def check(err, what="hip"):
    if err != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(err).decode()}")


def device_synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    def __init__(self):
        self.h = C.c_void_p()
        check(hip().hipStreamCreate(C.byref(self.h)), "hipStreamCreate")

    @property
    def ptr(self):
        return self.h

    def synchronize(self):
        check(hip().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.h:
                hip().hipStreamDestroy(self.h)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        check(hip().hipEventCreate(C.byref(self.h)), "hipEventCreate")

    def record(self, stream=None):
        check(hip().hipEventRecord(self.h, stream.ptr if isinstance(stream, Stream) else stream), "hipEventRecord")

    def synchronize(self):
        check(hip().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(hip().hipEventElapsedTime(C.byref(ms), self.h, end.h), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            if self.h:
                hip().hipEventDestroy(self.h)
        except Exception:
            pass


class DeviceBuffer:
    """hipMalloc'd buffer filled from a numpy array."""

    def __init__(self, arr):
        import numpy as np
        a = np.ascontiguousarray(arr)
        self.nbytes = a.nbytes
        self.ptr = C.c_void_p()
        check(hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 16)), "hipMalloc")
        check(hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, 1), "hipMemcpy H2D")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def mem_info():
    f, t = C.c_size_t(), C.c_size_t()
    check(hip().hipMemGetInfo(C.byref(f), C.byref(t)), "hipMemGetInfo")
    return f.value, t.value
