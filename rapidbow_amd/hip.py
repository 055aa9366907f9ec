"""Minimal ctypes access to the HIP runtime (device memory, streams, events).

The library links /opt/rocm/lib/libamdhip64.so; this module binds the same
runtime so host code, tests and the benchmark share one HIP runtime in the
process (no PyTorch, whose wheel carries its own copy).
"""
import ctypes
import os

import numpy as np

_HIP_PATHS = ["/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"]
_hip = None
for _p in _HIP_PATHS:
    try:
        _hip = ctypes.CDLL(_p, mode=ctypes.RTLD_GLOBAL)
        break
    except OSError:
        continue
if _hip is None:
    raise ImportError("libamdhip64.so not found")

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_hip.hipMalloc.argtypes = [ctypes.POINTER(_vp), _sz]
_hip.hipFree.argtypes = [_vp]
_hip.hipMemcpy.argtypes = [_vp, _vp, _sz, ctypes.c_int]
_hip.hipMemcpyAsync.argtypes = [_vp, _vp, _sz, ctypes.c_int, _vp]
_hip.hipMemset.argtypes = [_vp, ctypes.c_int, _sz]
_hip.hipStreamCreate.argtypes = [ctypes.POINTER(_vp)]
_hip.hipStreamDestroy.argtypes = [_vp]
_hip.hipStreamSynchronize.argtypes = [_vp]
_hip.hipEventCreate.argtypes = [ctypes.POINTER(_vp)]
_hip.hipEventDestroy.argtypes = [_vp]
_hip.hipEventRecord.argtypes = [_vp, _vp]
_hip.hipEventSynchronize.argtypes = [_vp]
_hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), _vp, _vp]
_hip.hipGetErrorString.restype = ctypes.c_char_p
_hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
_hip.hipMemGetInfo.argtypes = [ctypes.POINTER(_sz), ctypes.POINTER(_sz)]

H2D, D2H, D2D = 1, 2, 3


class HipError(RuntimeError):
    pass


def check(e, what=""):
    if e != 0:
        raise HipError(f"{what}: {_hip.hipGetErrorString(e).decode()} ({e})")


def device_count():
    n = ctypes.c_int(0)
    e = _hip.hipGetDeviceCount(ctypes.byref(n))
    return n.value if e == 0 else 0


def synchronize():
    check(_hip.hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    free, total = _sz(0), _sz(0)
    check(_hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceArray:
    """Owning device buffer with a numpy dtype/shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = _vp()
        check(_hip.hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(_hip.hipMemcpy(d.ptr, a.ctypes.data, d.nbytes, H2D), "hipMemcpy H2D")
        return d

    @classmethod
    def zeros(cls, shape, dtype):
        d = cls(shape, dtype)
        d.fill_zero()
        return d

    def fill_zero(self):
        if self.nbytes:
            check(_hip.hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def fill_byte(self, v):
        if self.nbytes:
            check(_hip.hipMemset(self.ptr, int(v), self.nbytes), "hipMemset")

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            check(_hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def numpy(self, count=None):
        out = np.empty(self.shape, self.dtype)
        if count is not None:
            out = out.reshape(-1)[:count]
        if out.nbytes:
            check(_hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def copy_from(self, src, nbytes=None, stream=None):
        """Asynchronous device-to-device copy of nbytes (default: all of src)."""
        n = src.nbytes if nbytes is None else int(nbytes)
        assert n <= self.nbytes and n <= src.nbytes
        check(_hip.hipMemcpyAsync(self.ptr, src.ptr, n, D2D, stream.handle if stream else None), "hipMemcpyAsync D2D")

    def view(self, first_row, nb_rows):
        """Non-owning view of rows [first_row, first_row + nb_rows) (leading dimension)."""
        row = self.nbytes // max(self.shape[0], 1) if self.shape else self.dtype.itemsize
        assert 0 <= first_row and first_row + nb_rows <= self.shape[0]
        v = DeviceArray.__new__(DeviceArray)
        v.shape = (int(nb_rows),) + self.shape[1:]
        v.dtype, v.nbytes, v.ptr, v._owner = self.dtype, int(nb_rows) * row, self.ptr + int(first_row) * row, self
        return v

    def free(self):
        if self.ptr and not hasattr(self, "_owner"):
            _hip.hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = _vp()
        check(_hip.hipStreamCreate(ctypes.byref(p)), "hipStreamCreate")
        self.handle = p.value

    def synchronize(self):
        check(_hip.hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.handle:
                _hip.hipStreamDestroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        p = _vp()
        check(_hip.hipEventCreate(ctypes.byref(p)), "hipEventCreate")
        self.handle = p.value

    def record(self, stream=None):
        check(_hip.hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self):
        check(_hip.hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float(0)
        check(_hip.hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                _hip.hipEventDestroy(self.handle)
        except Exception:
            pass
