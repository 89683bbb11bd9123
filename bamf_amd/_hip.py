"""Minimal HIP runtime access through ctypes: device buffers, streams, events.

Device memory for the host-side mirrors and tests is plain hipMalloc'd
memory; results travel back with hipMemcpy.  Nothing here computes.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_hip = None


def _load():
    global _hip
    if _hip is not None:
        return _hip
    names = ["libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"]
    rocm = os.environ.get("ROCM_PATH")
    if rocm:
        names.insert(0, os.path.join(rocm, "lib", "libamdhip64.so"))
    last = None
    for n in names:
        try:
            _hip = ctypes.CDLL(n)
            break
        except OSError as e:  # pragma: no cover - depends on the image
            last = e
    if _hip is None:
        raise RuntimeError(f"cannot load the HIP runtime (libamdhip64.so): {last}")
    h = _hip
    h.hipGetErrorString.restype = ctypes.c_char_p
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipEventDestroy.argtypes = [ctypes.c_void_p]
    h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    h.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
    h.hipSetDevice.argtypes = [ctypes.c_int]
    return h


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = _load().hipGetErrorString(rc).decode()
        raise RuntimeError(f"{what} failed: {msg} ({rc})")


H2D, D2H, D2D = 1, 2, 3


def device_count() -> int:
    try:
        h = _load()
    except RuntimeError:
        return 0
    n = ctypes.c_int(0)
    if h.hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


def synchronize() -> None:
    check(_load().hipDeviceSynchronize(), "hipDeviceSynchronize")


def copy_d2d(dst: "DeviceBuffer", src: "DeviceBuffer", nbytes: int, stream: "Stream | None" = None) -> None:
    """Asynchronous device-to-device copy on `stream`."""
    if nbytes > min(dst.nbytes, src.nbytes):
        raise ValueError("copy larger than its buffers")
    check(_load().hipMemcpyAsync(dst.vp, src.vp, nbytes, D2D, stream.vp if stream else None), "hipMemcpyAsync")


class DeviceBuffer:
    """An owned hipMalloc allocation."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(_load().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    @classmethod
    def zeros(cls, nbytes: int) -> "DeviceBuffer":
        b = cls(nbytes)
        check(_load().hipMemset(ctypes.c_void_p(b.ptr), 0, max(b.nbytes, 1)), "hipMemset")
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(_load().hipMemcpy(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D),
              "hipMemcpy H2D")

    def download(self, dtype, count: int | None = None, offset: int = 0) -> np.ndarray:
        """count elements of dtype starting at element `offset`."""
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize - offset
        if (offset + count) * dtype.itemsize > self.nbytes:
            raise ValueError("download beyond the buffer")
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(_load().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.c_void_p(self.ptr + offset * dtype.itemsize), out.nbytes, D2H),
                  "hipMemcpy D2H")
        return out

    def free(self) -> None:
        if self.ptr:
            _load().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def vp(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.ptr)


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(_load().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self) -> None:
        check(_load().hipStreamSynchronize(ctypes.c_void_p(self.handle)), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.handle:
                _load().hipStreamDestroy(ctypes.c_void_p(self.handle))
        except Exception:
            pass

    @property
    def vp(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.handle)


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(_load().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream: Stream | None = None) -> None:
        check(_load().hipEventRecord(ctypes.c_void_p(self.handle), stream.vp if stream else None), "hipEventRecord")

    def synchronize(self) -> None:
        check(_load().hipEventSynchronize(ctypes.c_void_p(self.handle)), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        check(_load().hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(self.handle), ctypes.c_void_p(end.handle)),
              "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                _load().hipEventDestroy(ctypes.c_void_p(self.handle))
        except Exception:
            pass
