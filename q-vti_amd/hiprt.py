"""Minimal ctypes binding of the HIP runtime that the engine library links
(ROCm's libamdhip64), for callers that hand the engine device buffers (the
allocation-free transform) and for the benchmark's stream-copy probe.

torch bundles a HIP runtime of its own; two HIP runtimes in one process do
not share the device, so code that drives libqvti never touches torch's GPU
side and uses this binding instead.
"""
import ctypes
import os

import numpy as np

H2D, D2H, D2D = 1, 2, 3
_rt = None


def rt():
    """The HIP runtime libqvti.so is linked against (loaded with it)."""
    global _rt
    if _rt is None:
        import qvti

        qvti.lib()  # maps libamdhip64.so.7 of ROCm into the process
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so.7")
        L = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        sig = {
            "hipMalloc": [ctypes.POINTER(vp), sz], "hipFree": [vp],
            "hipMemcpy": [vp, vp, sz, ctypes.c_int], "hipMemcpyAsync": [vp, vp, sz, ctypes.c_int, vp],
            "hipMemset": [vp, ctypes.c_int, sz], "hipDeviceSynchronize": [],
            "hipStreamCreate": [ctypes.POINTER(vp)], "hipStreamDestroy": [vp], "hipStreamSynchronize": [vp],
            "hipEventCreate": [ctypes.POINTER(vp)], "hipEventDestroy": [vp], "hipEventRecord": [vp, vp],
            "hipEventSynchronize": [vp], "hipEventElapsedTime": [ctypes.POINTER(ctypes.c_float), vp, vp],
            "hipGetErrorString": [ctypes.c_int],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_char_p if name == "hipGetErrorString" else ctypes.c_int
        _rt = L
    return _rt


def check(err, what="HIP call"):
    if err != 0:
        raise RuntimeError("%s failed: %s" % (what, rt().hipGetErrorString(err).decode()))


class DeviceBuffer:
    """nbytes of device memory (hipMalloc), freed on close."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(rt().hipMalloc(ctypes.byref(self.ptr), self.nbytes), "hipMalloc(%d)" % self.nbytes)

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(rt().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")

    def download(self, a):
        assert a.flags["C_CONTIGUOUS"] and a.nbytes <= self.nbytes
        check(rt().hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr, a.nbytes, D2H), "hipMemcpy D2H")
        return a

    def close(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Stream:
    def __init__(self):
        self.s = ctypes.c_void_p()
        check(rt().hipStreamCreate(ctypes.byref(self.s)), "hipStreamCreate")

    def synchronize(self):
        check(rt().hipStreamSynchronize(self.s), "hipStreamSynchronize")

    def close(self):
        if self.s:
            rt().hipStreamDestroy(self.s)
            self.s = ctypes.c_void_p()


def stream_copy(nbytes=2 << 30, reps=10):
    """Device-to-device copy of nbytes (hipMemcpyAsync, the runtime's blit
    kernel) repeated reps times on one stream between two HIP events: returns
    (seconds per copy, bytes read + written per second)."""
    L = rt()
    a, b = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
    st = Stream()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    try:
        check(L.hipMemset(a.ptr, 1, nbytes), "hipMemset")
        check(L.hipEventCreate(ctypes.byref(e0)))
        check(L.hipEventCreate(ctypes.byref(e1)))
        check(L.hipMemcpyAsync(b.ptr, a.ptr, nbytes, D2D, st.s))  # warm-up
        check(L.hipEventRecord(e0, st.s))
        for _ in range(reps):
            check(L.hipMemcpyAsync(b.ptr, a.ptr, nbytes, D2D, st.s))
        check(L.hipEventRecord(e1, st.s))
        check(L.hipEventSynchronize(e1))
        ms = ctypes.c_float()
        check(L.hipEventElapsedTime(ctypes.byref(ms), e0, e1))
        s = ms.value * 1e-3 / reps
        return s, 2.0 * nbytes / s
    finally:
        if e0:
            L.hipEventDestroy(e0)
        if e1:
            L.hipEventDestroy(e1)
        st.close()
        a.close()
        b.close()
