"""Device memory, streams and events through the HIP runtime (ctypes).

Plumbing for Python callers of libvz: device buffers for rays, hits, accum
and image, plus HIPOutputBuffer, the Python mirror of CUDAOutputBuffer
(Core/CUDAOutputBuffer.h:65-105) with the device and zero-copy types.
"""
import ctypes as C
import os

import numpy as np

_hip = None

H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        for cand in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(cand)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        vp = C.c_void_p
        _hip.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
        _hip.hipFree.argtypes = [vp]
        _hip.hipHostMalloc.argtypes = [C.POINTER(vp), C.c_size_t, C.c_uint]
        _hip.hipHostFree.argtypes = [vp]
        _hip.hipHostGetDevicePointer.argtypes = [C.POINTER(vp), vp, C.c_uint]
        _hip.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        _hip.hipMemcpyAsync.argtypes = [vp, vp, C.c_size_t, C.c_int, vp]
        _hip.hipMemset.argtypes = [vp, C.c_int, C.c_size_t]
        _hip.hipStreamCreate.argtypes = [C.POINTER(vp)]
        _hip.hipStreamDestroy.argtypes = [vp]
        _hip.hipStreamSynchronize.argtypes = [vp]
        _hip.hipEventCreate.argtypes = [C.POINTER(vp)]
        _hip.hipEventDestroy.argtypes = [vp]
        _hip.hipEventRecord.argtypes = [vp, vp]
        _hip.hipEventSynchronize.argtypes = [vp]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipGetErrorString.argtypes = [C.c_int]
        _hip.hipSetDevice.argtypes = [C.c_int]
    return _hip


def _ck(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


def synchronize():
    _ck(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def copy_d2d(dst, src, nbytes):
    """Device-to-device copy (the runtime's copy kernel), synchronous."""
    _ck(hip().hipMemcpy(dst.ptr, src.ptr, int(nbytes), D2D), "hipMemcpy D2D")


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        _ck(hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_array(cls, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        b.upload(arr)
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _ck(hip().hipMemcpy(self.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes, H2D), "hipMemcpy H2D")

    def download(self, dtype, shape=None):
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize
        out = np.empty(n, dtype)
        _ck(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, n * dtype.itemsize, D2H),
            "hipMemcpy D2H")
        return out.reshape(shape) if shape is not None else out

    def zero(self):
        """Clears the buffer and waits for the clear: the serving loop renders
        on its own non-blocking stream, which does not wait for the null one."""
        _ck(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")
        _ck(hip().hipStreamSynchronize(None), "hipStreamSynchronize")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def value(self):
        return self.ptr.value


class Stream:
    def __init__(self):
        self.h = C.c_void_p()
        _ck(hip().hipStreamCreate(C.byref(self.h)), "hipStreamCreate")

    def sync(self):
        _ck(hip().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def close(self):
        if self.h:
            hip().hipStreamDestroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        _ck(hip().hipEventCreate(C.byref(self.h)), "hipEventCreate")

    def record(self, stream=None):
        _ck(hip().hipEventRecord(self.h, stream.h if stream is not None else None), "hipEventRecord")

    def sync(self):
        _ck(hip().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = C.c_float()
        _ck(hip().hipEventElapsedTime(C.byref(ms), self.h, end.h), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            hip().hipEventDestroy(self.h)
        except Exception:
            pass


HIP_DEVICE, GL_INTEROP, ZERO_COPY, HIP_P2P = 0, 1, 2, 3


class HIPOutputBuffer:
    """Pixel buffer with the CUDAOutputBuffer surface (CUDAOutputBuffer.h:65-105).

    HIP_DEVICE: hipMalloc'd pixels; ZERO_COPY: mapped pinned host memory.
    GL_INTEROP has no meaning on a headless GPU and raises.
    """

    def __init__(self, kind, width, height, pixel_dtype=np.uint32):
        if kind in (GL_INTEROP,):
            raise NotImplementedError("GL interop output buffers are not supported (headless)")
        self.kind = kind
        self.dtype = np.dtype(pixel_dtype)
        self.width = self.height = 0
        self._dev = None
        self._host = C.c_void_p()
        self._devptr = C.c_void_p()
        self.resize(width, height)

    def resize(self, width, height):
        width, height = max(int(width), 1), max(int(height), 1)  # ensureMinimumSize
        if (width, height) == (self.width, self.height):
            return
        self._release()
        self.width, self.height = width, height
        nbytes = width * height * self.dtype.itemsize
        if self.kind in (HIP_DEVICE, HIP_P2P):
            self._dev = DeviceBuffer(nbytes)
        else:
            _ck(hip().hipHostMalloc(C.byref(self._host), nbytes, 0x2), "hipHostMalloc mapped")
            _ck(hip().hipHostGetDevicePointer(C.byref(self._devptr), self._host, 0),
                "hipHostGetDevicePointer")

    def map(self):
        return self._dev.value if self._dev is not None else self._devptr.value

    def unmap(self, stream=None):
        if stream is not None:
            stream.sync()
        else:
            synchronize()

    def get_host_pointer(self):
        if self._dev is not None:
            return self._dev.download(self.dtype).reshape(self.height, self.width)
        buf = (C.c_uint8 * (self.width * self.height * self.dtype.itemsize)).from_address(
            self._host.value)
        return np.frombuffer(buf, self.dtype).reshape(self.height, self.width).copy()

    def _release(self):
        if self._dev is not None:
            self._dev.free()
            self._dev = None
        if self._host:
            hip().hipHostFree(self._host)
            self._host = C.c_void_p()

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass
