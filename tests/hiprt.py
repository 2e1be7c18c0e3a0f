"""Minimal ctypes access to the HIP runtime for tests that need device buffers.

It binds libamdhip64.so.7 by soname, so it shares the runtime instance that
liblabeling_amd.so already loaded in this process (mixing in torch's bundled
copy of the runtime after the library initialised HIP is what we avoid).
"""
import ctypes

import numpy as np

_hip = None
H2D, D2H, D2D = 1, 2, 3


def lib():
    global _hip
    if _hip is None:
        h = ctypes.CDLL("libamdhip64.so.7")
        vp, sz, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, sz, i32]
        h.hipMemset.argtypes = [vp, i32, sz]
        h.hipMemcpyAsync.argtypes = [vp, vp, sz, i32, vp]
        h.hipDeviceSynchronize.argtypes = []
        h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipStreamDestroy.argtypes = [vp]
        h.hipStreamSynchronize.argtypes = [vp]
        h.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipEventRecord.argtypes = [vp, vp]
        h.hipEventSynchronize.argtypes = [vp]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        h.hipEventDestroy.argtypes = [vp]
        _hip = h
    return _hip


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(lib().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @property
    def addr(self):
        return self.ptr.value

    def upload(self, arr, offset=0):
        arr = np.ascontiguousarray(arr)
        check(lib().hipMemcpy(self.addr + offset, arr.ctypes.data, arr.nbytes, H2D), "hipMemcpy H2D")

    def download(self, shape, dtype, offset=0):
        out = np.empty(shape, dtype=dtype)
        check(lib().hipMemcpy(out.ctypes.data, self.addr + offset, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self):
        # hipMemset on device memory may return before it has run (it is
        # queued on the null stream), and the library's own streams do not
        # wait for the null stream: finish it before anything else runs
        check(lib().hipMemset(self.addr, 0, self.nbytes), "hipMemset")
        check(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def to_device(arr):
    b = DeviceBuffer(np.asarray(arr).nbytes)
    b.upload(arr)
    return b


def copy_async(dst, src, nbytes, stream):
    """Device-to-device copy enqueued on `stream` (a Stream)."""
    check(lib().hipMemcpyAsync(dst, src, nbytes, D2D, stream.handle), "hipMemcpyAsync")


def synchronize():
    check(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self):
        check(lib().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            lib().hipStreamDestroy(self.handle)
            self.handle = ctypes.c_void_p()
