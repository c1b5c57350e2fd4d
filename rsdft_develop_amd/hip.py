"""Minimal ctypes view of the HIP runtime the library itself links.

Device buffers and HIP events for callers of the device-pointer entry points
(``rsg_hpsi_f64`` / ``rsg_hpsi_f32``).  The runtime is resolved through the
already-loaded ``librsdft_gpu.so``, so buffers, events and the library's
stream all belong to one HIP runtime instance in the process.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import gpu

_rt = None
_H2D, _D2H, _D2D = 1, 2, 3


def runtime():
    global _rt
    if _rt is None:
        gpu.load_library()  # pulls in /opt/rocm/lib/libamdhip64.so.7
        rt = ctypes.CDLL("libamdhip64.so.7")
        vp, P = ctypes.c_void_p, ctypes.POINTER
        rt.hipMalloc.argtypes = [P(vp), ctypes.c_size_t]
        rt.hipFree.argtypes = [vp]
        rt.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        rt.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
        rt.hipDeviceSynchronize.argtypes = []
        rt.hipEventCreate.argtypes = [P(vp)]
        rt.hipEventRecord.argtypes = [vp, vp]
        rt.hipEventSynchronize.argtypes = [vp]
        rt.hipEventElapsedTime.argtypes = [P(ctypes.c_float), vp, vp]
        rt.hipEventDestroy.argtypes = [vp]
        rt.hipGetErrorString.restype = ctypes.c_char_p
        rt.hipGetErrorString.argtypes = [ctypes.c_int]
        _rt = rt
    return _rt


def _check(rc: int, what: str):
    if rc != 0:
        msg = runtime().hipGetErrorString(rc)
        raise RuntimeError(f"{what}: {msg.decode() if msg else rc}")


def copy_d2d(dst: int, src: int, nbytes: int):
    """Synchronous device-to-device copy between raw device addresses."""
    _check(runtime().hipMemcpy(dst, src, int(nbytes), _D2D), "hipMemcpy D2D")


class DeviceArray:
    """A column-major (Fortran-ordered) device array of float32/float64."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(runtime().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = int(p.value)

    @classmethod
    def from_numpy(cls, a: np.ndarray, dtype=None) -> "DeviceArray":
        a = np.asfortranarray(a, dtype=dtype or a.dtype)
        d = cls(a.shape, a.dtype)
        _check(runtime().hipMemcpy(d.ptr, a.ctypes.data, d.nbytes, _H2D), "hipMemcpy H2D")
        return d

    def fill_bytes(self, value: int):
        _check(runtime().hipMemset(self.ptr, int(value), self.nbytes), "hipMemset")

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        _check(runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _check(runtime().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.e = e.value

    def record(self, stream: int):
        _check(runtime().hipEventRecord(self.e, stream), "hipEventRecord")

    def synchronize(self):
        _check(runtime().hipEventSynchronize(self.e), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        _check(runtime().hipEventElapsedTime(ctypes.byref(ms), self.e, end.e), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            runtime().hipEventDestroy(self.e)
        except Exception:
            pass
