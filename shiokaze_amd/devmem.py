"""Device memory and streams for callers of shkz_project_device, through the same HIP runtime that
libshkz_mi355x.so links (ctypes).  Plumbing only: the benchmark and the GPU tests use it to hold
inputs resident in HBM without loading a second HIP runtime into the process (a framework that
bundles its own runtime has its own streams, which the library's runtime cannot order against)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as _shkz_lib

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        _shkz_lib()  # loads libamdhip64 as a dependency
        h = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
        h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        h.hipDeviceSynchronize.argtypes = []
        h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        _hip = h
    return _hip


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    def __init__(self, host: np.ndarray | None = None, shape=None, dtype=None):
        if host is not None:
            shape, dtype = host.shape, host.dtype
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        if host is not None:
            self.upload(host)

    def upload(self, host: np.ndarray, stream=None):
        host = np.ascontiguousarray(host, self.dtype)
        if stream is None:
            check(hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D), "hipMemcpy")
        else:
            check(hip().hipMemcpyAsync(self.ptr, host.ctypes.data, self.nbytes, H2D, stream), "hipMemcpyAsync")

    def copy_from(self, other: "DeviceArray", stream):
        check(hip().hipMemcpyAsync(self.ptr, other.ptr, self.nbytes, D2D, stream), "hipMemcpyAsync")

    def zero(self):
        check(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def data_ptr(self):
        return self.ptr


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(hip().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        check(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        hip().hipStreamDestroy(self.handle)


def device_synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
