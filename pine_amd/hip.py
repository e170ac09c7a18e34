"""Minimal HIP runtime access for the Python bindings: device buffers, copies,
synchronisation and events, through the same libamdhip64 the native library
links (one runtime per process).  torch is deliberately not used: it bundles
its own HIP runtime, and two runtimes in one process do not share the device.
"""
import ctypes

import numpy as np

_hip = None

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3


def rt():
    global _hip
    if _hip is None:
        from . import lib
        lib()  # loads libpinehip.so, which pulls in its libamdhip64
        _hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        _hip.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        _hip.hipFree.argtypes = [vp]
        _hip.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        _hip.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        _hip.hipEventRecord.argtypes = [vp, vp]
        _hip.hipEventSynchronize.argtypes = [vp]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        _hip.hipEventDestroy.argtypes = [vp]
    return _hip


class HipError(RuntimeError):
    pass


def check(rc, what="hip"):
    if rc != 0:
        raise HipError(f"{what}: {rt().hipGetErrorString(rc).decode()} ({rc})")


def device_count():
    n = ctypes.c_int(0)
    try:
        rc = rt().hipGetDeviceCount(ctypes.byref(n))
    except OSError:
        return 0
    return n.value if rc == 0 else 0


def synchronize():
    check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A device allocation with a numpy dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        if self.nbytes:
            check(rt().hipMalloc(ctypes.byref(self.ptr), self.nbytes), "hipMalloc")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(rt().hipMemcpy(d.ptr, a.ctypes.data, d.nbytes, hipMemcpyHostToDevice),
                  "hipMemcpy H2D")
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(rt().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, hipMemcpyDeviceToHost),
                  "hipMemcpy D2H")
        return out

    def zero(self):
        if self.nbytes:
            check(rt().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")
        return self

    def __len__(self):
        return self.shape[0]

    def free(self):
        if self.ptr and self.ptr.value:
            rt().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.h = ctypes.c_void_p()
        check(rt().hipEventCreate(ctypes.byref(self.h)), "hipEventCreate")

    def record(self, stream=None):
        check(rt().hipEventRecord(self.h, stream), "hipEventRecord")

    def elapsed_ms(self, end):
        check(rt().hipEventSynchronize(end.h), "hipEventSynchronize")
        ms = ctypes.c_float()
        check(rt().hipEventElapsedTime(ctypes.byref(ms), self.h, end.h), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            rt().hipEventDestroy(self.h)
        except Exception:
            pass
