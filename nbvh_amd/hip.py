"""Minimal HIP runtime binding (device buffers, streams, events, graphs) for callers of the C ABI.

It binds the libamdhip64 that libnbvh_hip.so itself links (the system ROCm runtime), so device
pointers and streams are shared with the library. torch is deliberately not used: its wheel
ships a second HIP/HSA runtime, and two runtimes in one process do not share a device.
"""
import ctypes
import os

import numpy as np

from . import _ffi

_hip = None

H2D, D2H, D2D = 1, 2, 3


def _runtime():
    global _hip
    if _hip is None:
        _ffi.lib()  # load libnbvh_hip.so first: its libamdhip64 is the one bound here
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line and "/torch/" not in line:
                    path = line.split()[-1]
                    break
        if path is None:
            path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        L = ctypes.CDLL(path)
        vp, sz, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        sig = {
            "hipGetDeviceCount": [ctypes.POINTER(i32)],
            "hipSetDevice": [i32],
            "hipMalloc": [ctypes.POINTER(vp), sz],
            "hipFree": [vp],
            "hipMemcpy": [vp, vp, sz, i32],
            "hipMemcpyAsync": [vp, vp, sz, i32, vp],
            "hipMemset": [vp, i32, sz],
            "hipDeviceSynchronize": [],
            "hipStreamCreate": [ctypes.POINTER(vp)],
            "hipStreamDestroy": [vp],
            "hipStreamSynchronize": [vp],
            "hipEventCreate": [ctypes.POINTER(vp)],
            "hipEventDestroy": [vp],
            "hipEventRecord": [vp, vp],
            "hipEventSynchronize": [vp],
            "hipEventElapsedTime": [ctypes.POINTER(ctypes.c_float), vp, vp],
            "hipGetErrorString": [i32],
            "hipDeviceGetName": [ctypes.c_char_p, i32, i32],
            "hipStreamBeginCapture": [vp, i32],
            "hipStreamEndCapture": [vp, ctypes.POINTER(vp)],
            "hipGraphInstantiate": [ctypes.POINTER(vp), vp, vp, vp, sz],
            "hipGraphLaunch": [vp, vp],
            "hipGraphExecDestroy": [vp],
            "hipGraphDestroy": [vp],
            "hipHostMalloc": [ctypes.POINTER(vp), sz, ctypes.c_uint],
            "hipHostFree": [vp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_char_p if name == "hipGetErrorString" else i32
        _hip = L
    return _hip


def check(err, what="hip call"):
    if err != 0:
        raise RuntimeError(f"{what}: {_runtime().hipGetErrorString(err).decode()} ({err})")


def device_count():
    n = ctypes.c_int(0)
    err = _runtime().hipGetDeviceCount(ctypes.byref(n))
    return n.value if err == 0 else 0


def device_name(dev=0):
    buf = ctypes.create_string_buffer(256)
    check(_runtime().hipDeviceGetName(buf, 256, dev), "hipDeviceGetName")
    return buf.value.decode()


def synchronize():
    check(_runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Buffer:
    """Device allocation of `nbytes`."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(_runtime().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if stream is None:
            check(_runtime().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        else:
            check(_runtime().hipMemcpyAsync(self.ptr, a.ctypes.data, a.nbytes, H2D, stream.handle), "H2D async")

    def download(self, dtype, count=None):
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(count, dtype)
        check(_runtime().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self):
        check(_runtime().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self):
        if self.ptr:
            _runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HostArray:
    """Pinned (page-locked) host memory from hipHostMalloc, seen as a numpy array of `count`
    elements of `dtype` (.array); copies from and to it run as asynchronous DMA."""

    def __init__(self, dtype, count):
        dtype = np.dtype(dtype)
        self.nbytes = int(count) * dtype.itemsize
        p = ctypes.c_void_p()
        check(_runtime().hipHostMalloc(ctypes.byref(p), max(self.nbytes, 1), 0), "hipHostMalloc")
        self.ptr = p.value
        buf = (ctypes.c_char * max(self.nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype, int(count))

    def free(self):
        if self.ptr:
            self.array = None
            _runtime().hipHostFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A HIP stream at the default priority."""

    def __init__(self):
        s = ctypes.c_void_p()
        check(_runtime().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        check(_runtime().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            _runtime().hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(_runtime().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None):
        check(_runtime().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self):
        check(_runtime().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        check(_runtime().hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return ms.value


class Graph:
    """Stream capture into a HIP graph: `with Graph.capture(stream) as g: <enqueue work>`, then
    g.launch(stream) replays the captured work."""

    CAPTURE_GLOBAL = 0

    def __init__(self):
        self.graph = None
        self.exec = None

    @classmethod
    def capture(cls, stream):
        g = cls()
        g._stream = stream
        return g

    def __enter__(self):
        check(_runtime().hipStreamBeginCapture(self._stream.handle, self.CAPTURE_GLOBAL), "hipStreamBeginCapture")
        return self

    def __exit__(self, exc_type, *rest):
        gr = ctypes.c_void_p()
        err = _runtime().hipStreamEndCapture(self._stream.handle, ctypes.byref(gr))
        if exc_type is not None:
            return False
        check(err, "hipStreamEndCapture")
        self.graph = gr.value
        ex = ctypes.c_void_p()
        check(_runtime().hipGraphInstantiate(ctypes.byref(ex), self.graph, None, None, 0), "hipGraphInstantiate")
        self.exec = ex.value
        return False

    def launch(self, stream):
        check(_runtime().hipGraphLaunch(self.exec, stream.handle), "hipGraphLaunch")

    def destroy(self):
        if self.exec:
            _runtime().hipGraphExecDestroy(self.exec)
            self.exec = None
        if self.graph:
            _runtime().hipGraphDestroy(self.graph)
            self.graph = None
