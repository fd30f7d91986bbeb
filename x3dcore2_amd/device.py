"""Minimal device-memory helper over the HIP runtime (hipMalloc / hipMemcpy / hipFree), for
callers that keep films and rays resident in HBM without pulling in another framework."""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_hip = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        _hip = C.CDLL(os.path.join(rocm, "lib", "libamdhip64.so"))
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetErrorString.restype = C.c_char_p
    return _hip


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {hip().hipGetErrorString(rc).decode()}")


def copy_device(dst: "DeviceArray", src: "DeviceArray", nbytes: int):
    """Device-to-device copy of nbytes (hipMemcpyDeviceToDevice)."""
    _check(hip().hipMemcpy(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), int(nbytes), 3), "hipMemcpy D2D")


def synchronize():
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A device allocation of `nbytes` bytes; `.ptr` is the device address."""

    H2D, D2H = 1, 2

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        _check(hip().hipMalloc(C.byref(p), max(int(nbytes), 16)), "hipMalloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def free(self):
        if self.ptr:
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        _check(hip().hipMemcpy(C.c_void_p(d.ptr), C.c_void_p(a.ctypes.data), a.nbytes, cls.H2D), "hipMemcpy H2D")
        return d

    def to_host(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(hip().hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr), out.nbytes, self.D2H),
               "hipMemcpy D2H")
        return out
