"""Minimal HIP runtime helpers over the same libamdhip64 that libcujson_amd.so
links (device buffers, copies, events).

The process must use a single HIP runtime: torch wheels bundle their own
libamdhip64/libhsa-runtime64, and a second runtime in the same process does
not see the GPU once the first has opened it.  Device memory for the
device-resident entry point is therefore allocated here, not through torch.
"""
from __future__ import annotations

import ctypes

import numpy as np

_hip = None

H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        import cujson_amd
        cujson_amd.lib()  # loads libamdhip64.so.7 as a dependency first
        L = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        L.hipMemcpyAsync.argtypes = [vp, vp, sz, ctypes.c_int, vp]
        L.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamSynchronize.argtypes = [vp]
        L.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        L.hipEventDestroy.argtypes = [vp]
        L.hipEventRecord.argtypes = [vp, vp]
        L.hipEventSynchronize.argtypes = [vp]
        L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        L.hipHostMalloc.argtypes = [ctypes.POINTER(vp), sz, ctypes.c_uint]
        L.hipHostFree.argtypes = [vp]
        L.hipGetErrorString.argtypes = [ctypes.c_int]
        L.hipGetErrorString.restype = ctypes.c_char_p
        _hip = L
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


class DeviceBuffer:
    def __init__(self, nbytes: int):
        self.nbytes = nbytes
        p = ctypes.c_void_p()
        _check(hip().hipMalloc(ctypes.byref(p), max(nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    def upload(self, arr: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(arr)
        assert offset + a.nbytes <= self.nbytes
        _check(hip().hipMemcpy(ctypes.c_void_p(self.ptr + offset), a.ctypes.data, a.nbytes, H2D),
               "hipMemcpy H2D")

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        _check(hip().hipMemcpy(out.ctypes.data, ctypes.c_void_p(self.ptr + offset), out.nbytes,
                               D2H), "hipMemcpy D2H")
        return out

    def copy_from(self, src: "DeviceBuffer", nbytes: int):
        """Device-to-device copy (a runtime copy kernel)."""
        assert nbytes <= self.nbytes and nbytes <= src.nbytes
        _check(hip().hipMemcpy(ctypes.c_void_p(self.ptr), ctypes.c_void_p(src.ptr), nbytes, D2D),
               "hipMemcpy D2D")

    def fill(self, byte: int):
        _check(hip().hipMemset(ctypes.c_void_p(self.ptr), byte, self.nbytes), "hipMemset")

    def free(self):
        if self.ptr:
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _check(hip().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.h = e

    def record(self, stream: int):
        _check(hip().hipEventRecord(self.h, ctypes.c_void_p(stream)), "hipEventRecord")

    def synchronize(self):
        _check(hip().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, later: "Event") -> float:
        f = ctypes.c_float()
        _check(hip().hipEventElapsedTime(ctypes.byref(f), self.h, later.h), "hipEventElapsedTime")
        return f.value


def synchronize():
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class PinnedHostBuffer:
    """Pinned host memory (hipHostMalloc) viewed as a numpy uint8 array."""

    def __init__(self, nbytes: int):
        p = ctypes.c_void_p()
        _check(hip().hipHostMalloc(ctypes.byref(p), max(nbytes, 1), 0), "hipHostMalloc")
        self.ptr = p.value
        self.nbytes = nbytes
        self.array = np.ctypeslib.as_array((ctypes.c_uint8 * max(nbytes, 1)).from_address(self.ptr))

    def free(self):
        if self.ptr:
            self.array = None
            hip().hipHostFree(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
