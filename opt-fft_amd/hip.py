"""Minimal ctypes binding of the HIP runtime: device buffers, copies, streams, events.

The library links /opt/rocm's libamdhip64; binding the same soname here keeps a single HIP
runtime in the process (no second runtime from another package).
"""
import ctypes

import numpy as np

_hip = None


def lib():
    global _hip
    if _hip is None:
        h = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
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
        h.hipGetLastError.argtypes = []
        h.hipGetErrorString.argtypes = [ctypes.c_int]
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipMemGetInfo.argtypes = [ctypes.POINTER(sz), ctypes.POINTER(sz)]
        h.hipStreamBeginCapture.argtypes = [vp, ctypes.c_int]
        h.hipStreamEndCapture.argtypes = [vp, ctypes.POINTER(vp)]
        h.hipGraphInstantiate.argtypes = [ctypes.POINTER(vp), vp, vp, vp, sz]
        h.hipGraphLaunch.argtypes = [vp, vp]
        h.hipGraphExecDestroy.argtypes = [vp]
        h.hipGraphDestroy.argtypes = [vp]
        _hip = h
    return _hip


H2D, D2H, D2D = 1, 2, 3


class HipError(RuntimeError):
    pass


def check(rc, what):
    if rc != 0:
        msg = lib().hipGetErrorString(rc).decode()
        raise HipError(f"{what} failed: {msg} ({rc})")


def device_count():
    n = ctypes.c_int(0)
    rc = lib().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    check(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    check(lib().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceBuffer:
    """A hipMalloc allocation of `nbytes`, freed on close()/garbage collection."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(lib().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @property
    def addr(self):
        return self.ptr.value

    def upload(self, arr, stream=None):
        a = np.ascontiguousarray(arr)
        if a.nbytes > self.nbytes:
            raise ValueError(f"upload of {a.nbytes} bytes into a {self.nbytes}-byte buffer")
        if stream is None:
            check(lib().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        else:
            check(lib().hipMemcpyAsync(self.ptr, a.ctypes.data, a.nbytes, H2D, stream),
                  "hipMemcpyAsync H2D")
            check(lib().hipStreamSynchronize(stream), "hipStreamSynchronize")

    def download(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("download larger than the buffer")
        check(lib().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def fill_zero(self):
        check(lib().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def close(self):
        if self.ptr is not None and self.ptr.value:
            lib().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self):
        check(lib().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().hipStreamDestroy(self.handle)
        self.handle = None


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib().hipEventCreate(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream=None):
        s = stream.handle if isinstance(stream, Stream) else stream
        check(lib().hipEventRecord(self.handle, s), "hipEventRecord")

    def synchronize(self):
        check(lib().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, later):
        ms = ctypes.c_float(0.0)
        check(lib().hipEventElapsedTime(ctypes.byref(ms), self.handle, later.handle),
              "hipEventElapsedTime")
        return float(ms.value)

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().hipEventDestroy(self.handle)
        self.handle = None


class GraphExec:
    """The work that `enqueue(stream)` puts on `stream`, captured once into a graph and
    instantiated; launch() replays it."""

    def __init__(self, stream, enqueue):
        L = lib()
        self.graph, self.exec = ctypes.c_void_p(), ctypes.c_void_p()
        check(L.hipStreamBeginCapture(stream.handle, 1), "hipStreamBeginCapture")  # thread-local
        try:
            enqueue(stream)
        finally:
            check(L.hipStreamEndCapture(stream.handle, ctypes.byref(self.graph)),
                  "hipStreamEndCapture")
        check(L.hipGraphInstantiate(ctypes.byref(self.exec), self.graph, None, None, 0),
              "hipGraphInstantiate")

    def launch(self, stream):
        check(lib().hipGraphLaunch(self.exec, stream.handle), "hipGraphLaunch")

    def close(self):
        if self.exec is not None and self.exec.value:
            lib().hipGraphExecDestroy(self.exec)
        if self.graph is not None and self.graph.value:
            lib().hipGraphDestroy(self.graph)
        self.exec = self.graph = None
