"""Minimal HIP runtime plumbing for the tests and the benchmark (ctypes).

It binds the same libamdhip64 instance that libcusort_hip.so is linked
against, so device pointers, streams and events are those of the library's
runtime (loading PyTorch's bundled HIP runtime into the same process would
give a second, separate runtime).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_hip = None

H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        import cusort_hip
        cusort_hip.lib()                      # loads libamdhip64.so.7 as a dependency
        for name in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so.7", "libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("HIP runtime (libamdhip64) not found")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipGetErrorString.argtypes = [ctypes.c_int]
        _hip.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        _hip.hipFree.argtypes = [vp]
        _hip.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        _hip.hipMemcpyAsync.argtypes = [vp, vp, sz, ctypes.c_int, vp]
        _hip.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        _hip.hipStreamDestroy.argtypes = [vp]
        _hip.hipStreamSynchronize.argtypes = [vp]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        _hip.hipEventDestroy.argtypes = [vp]
        _hip.hipEventRecord.argtypes = [vp, vp]
        _hip.hipEventSynchronize.argtypes = [vp]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip.hipSetDevice.argtypes = [ctypes.c_int]
        _hip.hipMemGetInfo.argtypes = [ctypes.POINTER(sz), ctypes.POINTER(sz)]
    return _hip


def ok(err: int, what: str) -> None:
    if err != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(err).decode()} ({err})")


def device_count() -> int:
    n = ctypes.c_int(0)
    err = hip().hipGetDeviceCount(ctypes.byref(n))
    return int(n.value) if err == 0 else 0


def set_device(i: int = 0) -> None:
    ok(hip().hipSetDevice(i), "hipSetDevice")


def synchronize() -> None:
    ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info() -> tuple[int, int]:
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    ok(hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return int(free.value), int(total.value)


class Buffer:
    """Device allocation of `nbytes` bytes (freed when collected)."""

    def __init__(self, nbytes: int):
        self.nbytes = max(1, int(nbytes))
        p = ctypes.c_void_p()
        ok(hip().hipMalloc(ctypes.byref(p), self.nbytes), f"hipMalloc({self.nbytes})")
        self.ptr = int(p.value)

    def __del__(self):
        if getattr(self, "ptr", 0) and _hip is not None:
            _hip.hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(arr)
        assert offset + a.nbytes <= self.nbytes
        ok(hip().hipMemcpy(ctypes.c_void_p(self.ptr + offset), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D),
           "hipMemcpy H2D")

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        if out.nbytes:
            ok(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr + offset), out.nbytes,
                               D2H), "hipMemcpy D2H")
        return out

    def fill(self, byte: int = 0) -> None:
        ok(hip().hipMemset(ctypes.c_void_p(self.ptr), byte, self.nbytes), "hipMemset")


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        ok(hip().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")   # default priority
        self.handle = int(s.value)

    def synchronize(self) -> None:
        ok(hip().hipStreamSynchronize(ctypes.c_void_p(self.handle)), "hipStreamSynchronize")

    def __del__(self):
        if getattr(self, "handle", 0) and _hip is not None:
            _hip.hipStreamDestroy(ctypes.c_void_p(self.handle))
            self.handle = 0


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        ok(hip().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = int(e.value)

    def record(self, stream: "Stream | None" = None) -> None:
        s = ctypes.c_void_p(stream.handle if stream else 0)
        ok(hip().hipEventRecord(ctypes.c_void_p(self.handle), s), "hipEventRecord")

    def synchronize(self) -> None:
        ok(hip().hipEventSynchronize(ctypes.c_void_p(self.handle)), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        ok(hip().hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(self.handle), ctypes.c_void_p(end.handle)),
           "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        if getattr(self, "handle", 0) and _hip is not None:
            _hip.hipEventDestroy(ctypes.c_void_p(self.handle))
            self.handle = 0
