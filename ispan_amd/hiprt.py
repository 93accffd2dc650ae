"""Caller-side device memory, streams and device release through the same HIP
runtime library that libispan_hip.so is linked against (a second HIP runtime
in the process, such as the one bundled with torch, cannot share the device
with it).  Used by the GPU tests (caller-owned buffers and streams) and by
bench.py (copy calibration, device release before exit)."""
from __future__ import annotations

import ctypes as C

import numpy as np

import ispan_amd as I

_H2D, _D2H, _D2D = 1, 2, 3
_rt = None


def rt():
    global _rt
    if _rt is None:
        I.lib()  # loads libamdhip64.so.7 as a dependency; the CDLL below is that same copy
        r = C.CDLL("libamdhip64.so.7")
        for n, args in (("hipMalloc", [C.POINTER(C.c_void_p), C.c_size_t]), ("hipFree", [C.c_void_p]),
                        ("hipMemcpy", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
                        ("hipStreamCreate", [C.POINTER(C.c_void_p)]), ("hipStreamDestroy", [C.c_void_p]),
                        ("hipStreamSynchronize", [C.c_void_p]), ("hipDeviceSynchronize", []),
                        ("hipDeviceReset", []),
                        ("hipMemGetInfo", [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)])):
            getattr(r, n).argtypes = args
            getattr(r, n).restype = C.c_int
        _rt = r
    return _rt


def _ok(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A caller-owned device buffer holding a copy of a numpy array."""

    def __init__(self, host: np.ndarray):
        self.dtype, self.shape = host.dtype, host.shape
        self.nbytes = host.nbytes
        p = C.c_void_p()
        _ok(rt().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        self.upload(host)

    def upload(self, host: np.ndarray):
        h = np.ascontiguousarray(host, self.dtype)
        assert h.nbytes == self.nbytes
        _ok(rt().hipMemcpy(self.ptr, h.ctypes.data, self.nbytes, _H2D), "hipMemcpy H2D")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _ok(rt().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            _ok(rt().hipFree(self.ptr), "hipFree")
            self.ptr = None


class Stream:
    def __init__(self):
        s = C.c_void_p()
        _ok(rt().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        _ok(rt().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            _ok(rt().hipStreamDestroy(self.handle), "hipStreamDestroy")
            self.handle = None


def device_copy(dst: int, src: int, nbytes: int):
    """Device-to-device copy (the runtime's own copy kernel)."""
    _ok(rt().hipMemcpy(dst, src, nbytes, _D2D), "hipMemcpy D2D")


def upload(dst: int, host: np.ndarray):
    h = np.ascontiguousarray(host)
    _ok(rt().hipMemcpy(dst, h.ctypes.data, h.nbytes, _H2D), "hipMemcpy H2D")


def malloc(nbytes: int) -> int:
    p = C.c_void_p()
    _ok(rt().hipMalloc(C.byref(p), max(nbytes, 1)), "hipMalloc")
    return p.value


def free(ptr: int):
    _ok(rt().hipFree(ptr), "hipFree")


def mem_info() -> tuple[int, int]:
    """(free, total) device memory in bytes (hipMemGetInfo)."""
    f, t = C.c_size_t(), C.c_size_t()
    _ok(rt().hipMemGetInfo(C.byref(f), C.byref(t)), "hipMemGetInfo")
    return f.value, t.value


def synchronize():
    _ok(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


def release_device():
    """Tear the device state of this process down now (every Solver must be
    closed).  A process that ends right after it leaves nothing of the HIP
    runtime for exit-time handlers; profilers that write their results at
    exit need this after cooperative launches."""
    _ok(rt().hipDeviceReset(), "hipDeviceReset")
