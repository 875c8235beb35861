"""Device buffers and HIP events through the same HIP runtime the library uses.

libsimplefem_amd.so links /opt/rocm's libamdhip64.so.7.  A PyTorch wheel
bundles its own copy of the HIP runtime; two runtimes in one process cannot
both own the GPU, so the tests and the benchmark allocate device memory and
time with events here, via ctypes on the runtime already loaded by the
library, and never import torch.  (A caller that wants torch tensors must
import torch first: the library then binds to torch's copy of the runtime,
which carries the same soname.)
"""
from __future__ import annotations

import ctypes
from ctypes import c_float, c_int, c_size_t, c_void_p

import numpy as np

from . import capi

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        capi.load()  # the runtime the library is linked against
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(c_void_p), c_size_t]
        _hip.hipFree.argtypes = [c_void_p]
        _hip.hipMemcpy.argtypes = [c_void_p, c_void_p, c_size_t, c_int]
        _hip.hipMemset.argtypes = [c_void_p, c_int, c_size_t]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(c_int)]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(c_void_p)]
        _hip.hipEventRecord.argtypes = [c_void_p, c_void_p]
        _hip.hipEventSynchronize.argtypes = [c_void_p]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(c_float), c_void_p, c_void_p]
        _hip.hipEventDestroy.argtypes = [c_void_p]
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(c_void_p)]
        _hip.hipStreamDestroy.argtypes = [c_void_p]
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipGetErrorString.argtypes = [c_int]
    return _hip


def check(rc):
    if rc != 0:
        raise RuntimeError("HIP error %d: %s" % (rc, hip().hipGetErrorString(rc).decode()))


def device_count() -> int:
    n = c_int(0)
    rc = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    check(hip().hipDeviceSynchronize())


class DeviceArray:
    """A float64 (or int32) vector in HBM, zero-filled unless zero=False.

    The solvers take x as their initial guess, as the reference's do, so a
    vector that starts as whatever hipMalloc returned (another solve's data
    after a free, or NaN) would change the solve it is handed to."""

    def __init__(self, n: int, dtype=np.float64, zero: bool = True):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = c_void_p()
        check(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)))
        if zero:
            self.zero()

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape[0], a.dtype, zero=False)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        check(hip().hipMemcpy(self.ptr, a.ctypes.data_as(c_void_p), self.nbytes, H2D))

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        check(hip().hipMemcpy(out.ctypes.data_as(c_void_p), self.ptr, self.nbytes, D2H))
        return out

    def zero(self):
        check(hip().hipMemset(self.ptr, 0, self.nbytes))

    def copy_from(self, other: "DeviceArray"):
        """Device-to-device copy of other's first nbytes (hipMemcpy)."""
        if other.nbytes < self.nbytes:
            raise ValueError("source is smaller than the destination")
        check(hip().hipMemcpy(self.ptr, other.ptr, self.nbytes, D2D))

    @property
    def address(self) -> int:
        return self.ptr.value

    def free(self):
        if self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.ev = c_void_p()
        check(hip().hipEventCreate(ctypes.byref(self.ev)))

    def record(self, stream=None):
        check(hip().hipEventRecord(self.ev, stream))

    def synchronize(self):
        check(hip().hipEventSynchronize(self.ev))

    def elapsed_ms(self, later: "Event") -> float:
        ms = c_float(0.0)
        check(hip().hipEventElapsedTime(ctypes.byref(ms), self.ev, later.ev))
        return ms.value

    def __del__(self):
        try:
            hip().hipEventDestroy(self.ev)
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = c_void_p()
        check(hip().hipStreamCreate(ctypes.byref(self.handle)))

    def __del__(self):
        try:
            hip().hipStreamDestroy(self.handle)
        except Exception:
            pass
