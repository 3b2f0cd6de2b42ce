"""Minimal ctypes access to the HIP runtime: device buffers and events.

Used by tests and bench.py to hand device pointers to pfdd_fft and to time
work on a context's stream with HIP events.
"""
import ctypes

import numpy as np

_hip = None


def rt():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        vp = ctypes.c_void_p
        _hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        _hip.hipFree.argtypes = [vp]
        _hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        _hip.hipEventRecord.argtypes = [vp, vp]
        _hip.hipEventSynchronize.argtypes = [vp]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        _hip.hipEventDestroy.argtypes = [vp]
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def _ck(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


H2D, D2H = 1, 2


def device_count():
    n = ctypes.c_int(0)
    if rt().hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


class DeviceBuffer:
    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _ck(rt().hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _ck(rt().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        _ck(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")
        return b

    def to_array(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _ck(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _ck(rt().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _ck(rt().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.e = e.value

    def record(self, stream):
        _ck(rt().hipEventRecord(self.e, stream), "hipEventRecord")

    def synchronize(self):
        _ck(rt().hipEventSynchronize(self.e), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _ck(rt().hipEventElapsedTime(ctypes.byref(ms), self.e, end.e), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            rt().hipEventDestroy(self.e)
        except Exception:
            pass
