"""Minimal ctypes access to the HIP runtime (libamdhip64): device memory,
copies, streams and events.  Plumbing for tests, the bench and Python callers
of the device-pointer entry points; no kernels live here."""
import ctypes
import os

import numpy as np

_ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
_hip = None

H2D, D2H, D2D = 1, 2, 3


class HipError(RuntimeError):
    pass


def lib():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", os.path.join(_ROCM, "lib", "libamdhip64.so")):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise HipError("libamdhip64.so not found")
        h = _hip
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        h.hipMemcpyAsync.argtypes = [vp, vp, sz, ctypes.c_int, vp]
        h.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        h.hipSetDevice.argtypes = [ctypes.c_int]
        h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipStreamDestroy.argtypes = [vp]
        h.hipStreamSynchronize.argtypes = [vp]
        h.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipEventDestroy.argtypes = [vp]
        h.hipEventRecord.argtypes = [vp, vp]
        h.hipEventSynchronize.argtypes = [vp]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipGetErrorString.argtypes = [ctypes.c_int]
        h.hipDeviceGetAttribute.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int]
    return _hip


def check(rc, what):
    if rc != 0:
        raise HipError(f"{what}: {lib().hipGetErrorString(rc).decode()} ({rc})")


def device_count():
    n = ctypes.c_int(0)
    rc = lib().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    check(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A device allocation holding a numpy-described array (shape, dtype)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if hasattr(shape, "__len__") else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.shape, arr.dtype)
        d.upload(arr)
        return d

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes
        check(lib().hipMemcpy(self.ptr, arr.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def at(self, elem_offset):
        """Device address of element `elem_offset` (flat)."""
        return self.ptr + int(elem_offset) * self.dtype.itemsize

    def free(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(lib().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        check(lib().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            lib().hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(lib().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None):
        check(lib().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self):
        check(lib().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        check(lib().hipEventElapsedTime(ctypes.byref(ms), self.handle, later.handle), "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        if self.handle:
            lib().hipEventDestroy(self.handle)
            self.handle = None


def cu_count(device=0):
    v = ctypes.c_int()
    # hipDeviceAttributeMultiprocessorCount
    HIP_ATTR_MP_COUNT = 63
    rc = lib().hipDeviceGetAttribute(ctypes.byref(v), HIP_ATTR_MP_COUNT, device)
    return v.value if rc == 0 else None
