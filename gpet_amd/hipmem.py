"""Device and pinned host memory through the HIP runtime that libgpet.so is
linked against (ctypes), for callers of the entry points that take device
pointers and for pinned output buffers.

The runtime is found among the libraries the process has already mapped
after libgpet.so is loaded, so that memory comes from the same runtime
instance the library uses (a second HIP runtime in one process, such as the
one bundled with a framework, does not share its device state).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_rt = None


def runtime():
    """The HIP runtime library of libgpet.so (loaded on first use)."""
    global _rt
    if _rt is not None:
        return _rt
    from .engine import LIB_PATH

    C.CDLL(LIB_PATH)
    path = None
    with open("/proc/self/maps") as fh:
        for line in fh:
            p = line.split()[-1]
            if os.path.basename(p).startswith("libamdhip64.so"):
                path = p
                break
    if path is None:
        raise OSError("libamdhip64 is not loaded by %s" % LIB_PATH)
    rt = C.CDLL(path)
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
    rt.hipHostFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    rt.hipDeviceSynchronize.argtypes = []
    for f in ("hipMalloc", "hipFree", "hipHostMalloc", "hipHostFree", "hipMemcpy", "hipMemset",
              "hipDeviceSynchronize"):
        getattr(rt, f).restype = C.c_int
    _rt = rt
    return rt


def _ok(e, what):
    if e != 0:
        raise RuntimeError("%s failed (hipError %d)" % (what, e))


class DeviceBuffer:
    """nbytes of device memory; .ptr is the device address."""

    def __init__(self, nbytes: int):
        self.rt = runtime()
        p = C.c_void_p()
        _ok(self.rt.hipMalloc(C.byref(p), max(int(nbytes), 1)), "hipMalloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _ok(self.rt.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")

    def download(self, nbytes: int) -> np.ndarray:
        out = np.empty(int(nbytes), np.uint8)
        if nbytes:
            _ok(self.rt.hipMemcpy(out.ctypes.data, self.ptr, int(nbytes), 2), "hipMemcpy D2H")
        return out

    def fill(self, byte: int):
        _ok(self.rt.hipMemset(self.ptr, int(byte), self.nbytes), "hipMemset")

    def free(self):
        if self.ptr:
            self.rt.hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedBuffer:
    """nbytes of pinned (page-locked) host memory, as a numpy uint8 array in
    .array; .ptr is its address."""

    def __init__(self, nbytes: int):
        self.rt = runtime()
        p = C.c_void_p()
        _ok(self.rt.hipHostMalloc(C.byref(p), max(int(nbytes), 1), 0), "hipHostMalloc")
        self.ptr = p.value
        self.array = np.ctypeslib.as_array((C.c_uint8 * max(int(nbytes), 1)).from_address(self.ptr))

    def free(self):
        if self.ptr:
            self.array = None
            self.rt.hipHostFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _ok(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")
