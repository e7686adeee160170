"""ctypes access to the HIP runtime that libomega_h_amd.so links.

Device buffers and events for callers of the device-pointer entry points
(osh_measure_edges_metric, osh_measure_qualities) and for timing kernels on
a context's own stream.  The runtime is resolved by soname after the library
is loaded, so this is the same HIP runtime instance the library uses (a
second runtime in the process, such as the one a PyTorch wheel bundles, would
not see the library's streams).
"""
import ctypes as C

import numpy as np

from . import load

_rt = None

H2D = 1
D2H = 2
D2D = 3


def runtime():
    global _rt
    if _rt is None:
        load()
        rt = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        P = C.c_void_p
        for name, args in {
            "hipMalloc": [C.POINTER(P), C.c_size_t],
            "hipFree": [P],
            "hipMemcpy": [P, P, C.c_size_t, C.c_int],
            "hipEventCreate": [C.POINTER(P)],
            "hipEventDestroy": [P],
            "hipEventRecord": [P, P],
            "hipEventSynchronize": [P],
            "hipEventElapsedTime": [C.POINTER(C.c_float), P, P],
            "hipStreamSynchronize": [P],
        }.items():
            f = getattr(rt, name)
            f.argtypes = args
            f.restype = C.c_int
        _rt = rt
    return _rt


def _check(status, what):
    if status != 0:
        raise RuntimeError(f"{what} failed with HIP status {status}")


class DeviceArray:
    """A device copy of a 1-D numpy array."""

    def __init__(self, host=None, dtype=None, n=None):
        rt = runtime()
        if host is not None:
            host = np.ascontiguousarray(host)
            self.dtype, self.n = host.dtype, host.size
        else:
            self.dtype, self.n = np.dtype(dtype), int(n)
        self.nbytes = self.dtype.itemsize * self.n
        p = C.c_void_p()
        _check(rt.hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        if host is not None and self.nbytes:
            _check(rt.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D), "hipMemcpy")

    def copy_from(self, other):
        """Device-to-device copy of `other` (same byte size) into this array."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        if self.nbytes:
            _check(runtime().hipMemcpy(self.ptr, other.ptr, self.nbytes, D2D), "hipMemcpy")

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        if self.nbytes:
            _check(runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H),
                   "hipMemcpy")
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            runtime().hipFree(self.ptr)
            self.ptr = None


class Event:
    def __init__(self):
        rt = runtime()
        e = C.c_void_p()
        _check(rt.hipEventCreate(C.byref(e)), "hipEventCreate")
        self.h = e.value

    def record(self, stream):
        _check(runtime().hipEventRecord(self.h, stream), "hipEventRecord")

    def synchronize(self):
        _check(runtime().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = C.c_float()
        _check(runtime().hipEventElapsedTime(C.byref(ms), self.h, end.h), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        if getattr(self, "h", None):
            runtime().hipEventDestroy(self.h)
            self.h = None
