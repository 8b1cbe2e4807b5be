"""Minimal HIP runtime plumbing over ctypes: device buffers, streams, events.

Binds the libamdhip64 that libmgad.so itself links, so a process that never
imports torch uses exactly one HIP runtime.  Only what the drivers, tests and
benchmark need: hipMalloc/hipFree, host<->device copies, streams, events.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi

_hip = None

H2D, D2H, D2D = 1, 2, 3


def lib():
    global _hip
    if _hip is None:
        _capi.lib()  # loads libmgad.so, which pulls in libamdhip64.so.7
        L = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        L.hipMemcpyAsync.argtypes = [vp, vp, sz, ctypes.c_int, vp]
        L.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        L.hipStreamDestroy.argtypes = [vp]
        L.hipStreamSynchronize.argtypes = [vp]
        L.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        L.hipEventDestroy.argtypes = [vp]
        L.hipEventRecord.argtypes = [vp, vp]
        L.hipEventSynchronize.argtypes = [vp]
        L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        L.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.hipGetErrorString.restype = ctypes.c_char_p
        L.hipGetErrorString.argtypes = [ctypes.c_int]
        L.hipMemGetInfo.argtypes = [ctypes.POINTER(sz), ctypes.POINTER(sz)]
        _hip = L
    return _hip


class HipError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().hipGetErrorString(rc).decode()
        raise HipError(f"{what}: HIP error {rc} ({msg})")


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = lib().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize() -> None:
    check(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A HIP stream created with default priority."""

    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self) -> None:
        check(lib().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __del__(self):
        if _hip is not None and self.handle:
            _hip.hipStreamDestroy(self.handle)
            self.handle = ctypes.c_void_p()


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib().hipEventCreate(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream: Stream | None = None) -> None:
        check(lib().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self) -> None:
        check(lib().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        check(lib().hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        if _hip is not None and self.handle:
            _hip.hipEventDestroy(self.handle)
            self.handle = ctypes.c_void_p()


class DeviceArray:
    """A hipMalloc'd buffer of float64 (or raw bytes) with host copies."""

    def __init__(self, shape, dtype=np.float64, nbytes: int | None = None):
        self.shape = tuple(shape) if not np.isscalar(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(nbytes) if nbytes is not None else int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        check(lib().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @property
    def size(self) -> int:
        return int(np.prod(self.shape))

    def upload(self, a: np.ndarray, stream: Stream | None = None) -> None:
        """Copy from the host; with a stream, ordered on that stream (and
        finished on return), so no null-stream operation is issued."""
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        src = a.ctypes.data_as(ctypes.c_void_p)
        if stream is None:
            check(lib().hipMemcpy(self.ptr, src, self.nbytes, H2D), "hipMemcpy H2D")
        else:
            check(lib().hipMemcpyAsync(self.ptr, src, self.nbytes, H2D, stream.handle), "hipMemcpyAsync H2D")
            stream.synchronize()

    def download(self, stream: Stream | None = None) -> np.ndarray:
        """Copy to the host after all queued device work has finished, or,
        with a stream, after that stream's work (a device-wide synchronisation
        and null-stream copies are not permitted while another host thread
        captures a graph, so callers that share the device with such threads
        pass their stream)."""
        out = np.empty(self.shape, self.dtype)
        dst = out.ctypes.data_as(ctypes.c_void_p)
        if stream is None:
            synchronize()
            check(lib().hipMemcpy(dst, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        else:
            check(lib().hipMemcpyAsync(dst, self.ptr, self.nbytes, D2H, stream.handle), "hipMemcpyAsync D2H")
            stream.synchronize()
        return out

    def copy_from(self, other: "DeviceArray", stream: Stream | None = None) -> None:
        assert other.nbytes == self.nbytes
        check(lib().hipMemcpyAsync(self.ptr, other.ptr, self.nbytes, D2D, stream.handle if stream else None),
              "hipMemcpyAsync D2D")

    def zero(self) -> None:
        check(lib().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self) -> None:
        if _hip is not None and self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        self.free()
