"""Minimal ctypes access to the HIP runtime for device buffers.

Used by the host driver, bench.py and the tests to hand device pointers to
the w3d_*_dev entry points without pulling in a framework.
"""
import ctypes as C
import os

import numpy as np

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        for cand in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _HIP = C.CDLL(cand)
                break
            except OSError:
                continue
        if _HIP is None:
            raise RuntimeError("libamdhip64.so not found")
        _HIP.hipGetErrorString.restype = C.c_char_p
    return _HIP


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


H2D, D2H, D2D = 1, 2, 3


def device_count():
    n = C.c_int()
    rc = hip().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


class DeviceArray:
    """A hipMalloc'd buffer with numpy-shaped host copies."""

    def __init__(self, n, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = max(self.n * self.dtype.itemsize, 1)
        p = C.c_void_p()
        _chk(hip().hipMalloc(C.byref(p), C.c_size_t(self.nbytes)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        _chk(hip().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p),
                             C.c_size_t(a.nbytes), H2D), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.n, self.dtype)
        _chk(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                             C.c_size_t(out.nbytes), D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_d2d(dst, src, nbytes):
    """Device-to-device copy (a plain copy kernel on one device)."""
    _chk(hip().hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes),
                         D2D), "hipMemcpy D2D")


def synchronize():
    _chk(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def device_name(dev=0):
    buf = C.create_string_buffer(256)
    if hip().hipDeviceGetName(buf, 256, dev) == 0:
        return buf.value.decode()
    return "unknown"


def visible():
    return os.environ.get("HIP_VISIBLE_DEVICES", "")
