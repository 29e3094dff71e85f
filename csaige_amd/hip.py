"""Minimal HIP runtime access over ctypes: device buffers and events on a given stream.

Used to hand device pointers to the ``_device`` entry points and to time kernels with
HIP events recorded on the library's own streams.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_hip = None


def lib():
    global _hip
    if _hip is None:
        path = "/opt/rocm/lib/libamdhip64.so"
        _hip = C.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        _hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipEventDestroy.argtypes = [C.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipGetErrorString.argtypes = [C.c_int]
    return _hip


H2D, D2H, D2D = 1, 2, 3


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {lib().hipGetErrorString(rc).decode()}")


class DeviceBuffer:
    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        _chk(lib().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)), f"hipMalloc({nbytes})")

    @property
    def value(self):
        return self.ptr.value

    def upload(self, arr, offset=0):
        a = np.ascontiguousarray(arr)
        _chk(lib().hipMemcpy(C.c_void_p(self.ptr.value + offset), a.ctypes.data, a.nbytes, H2D), "H2D")

    def download(self, arr, offset=0):
        _chk(lib().hipMemcpy(arr.ctypes.data, C.c_void_p(self.ptr.value + offset), arr.nbytes, D2H), "D2H")
        return arr

    def copy_from(self, other, nbytes=None):
        nb = self.nbytes if nbytes is None else int(nbytes)
        _chk(lib().hipMemcpy(self.ptr, other.ptr, nb, D2D), "D2D")

    def zero(self):
        _chk(lib().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_synchronize():
    _chk(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Event:
    def __init__(self):
        self.e = C.c_void_p()
        _chk(lib().hipEventCreate(C.byref(self.e)), "hipEventCreate")

    def record(self, stream):
        _chk(lib().hipEventRecord(self.e, C.c_void_p(stream)), "hipEventRecord")

    def synchronize(self):
        _chk(lib().hipEventSynchronize(self.e), "hipEventSynchronize")

    def elapsed_ms(self, end):
        t = C.c_float()
        end.synchronize()
        _chk(lib().hipEventElapsedTime(C.byref(t), self.e, end.e), "hipEventElapsedTime")
        return t.value

    def __del__(self):
        try:
            lib().hipEventDestroy(self.e)
        except Exception:
            pass
