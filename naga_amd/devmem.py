"""Device memory, copies and event timing through the HIP runtime that
libnaga_amd.so itself is linked against (libamdhip64.so.7, resolved to the
copy already loaded with the library), for callers of the device-pointer
entry points and for measurement.  The library is loaded first, so one HIP
runtime serves the process."""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None

H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        from naga_amd import load
        load()  # the library's runtime is the one resolved below
        lib = C.CDLL("libamdhip64.so.7")
        vp, sz = C.c_void_p, C.c_size_t
        for name, args in (("hipMalloc", [C.POINTER(vp), sz]), ("hipFree", [vp]),
                           ("hipMemcpy", [vp, vp, sz, C.c_int]),
                           ("hipMemcpyAsync", [vp, vp, sz, C.c_int, vp]),
                           ("hipDeviceSynchronize", []), ("hipEventCreate", [C.POINTER(vp)]),
                           ("hipEventDestroy", [vp]), ("hipEventRecord", [vp, vp]),
                           ("hipEventSynchronize", [vp]),
                           ("hipEventElapsedTime", [C.POINTER(C.c_float), vp, vp]),
                           ("hipGetErrorString", [C.c_int])):
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = C.c_char_p if name == "hipGetErrorString" else C.c_int
        _hip = lib
    return _hip


def check(status: int, what: str) -> None:
    if status != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(status).decode()}")


class DeviceArray:
    """A device allocation holding a copy of a host array (or nbytes of
    uninitialised memory); .ptr is the device address."""

    def __init__(self, host: np.ndarray | None = None, nbytes: int | None = None):
        self.nbytes = int(host.nbytes if host is not None else nbytes)
        p = C.c_void_p()
        check(hip().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        if host is not None:
            h = np.ascontiguousarray(host)
            check(hip().hipMemcpy(self.ptr, h.ctypes.data, self.nbytes, H2D), "hipMemcpy")

    def to_host(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy")
        return out

    def free(self) -> None:
        if self.ptr:
            check(hip().hipFree(self.ptr), "hipFree")
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def copy_rate(nbytes: int = 2 << 30, reps: int = 10) -> float:
    """Device-to-device copy rate of the runtime's copy over nbytes, in
    bytes read + written per second, timed with HIP events on the null
    stream."""
    a, b = DeviceArray(nbytes=nbytes), DeviceArray(nbytes=nbytes)
    h = hip()
    e0, e1 = C.c_void_p(), C.c_void_p()
    check(h.hipEventCreate(C.byref(e0)), "hipEventCreate")
    check(h.hipEventCreate(C.byref(e1)), "hipEventCreate")
    check(h.hipMemcpyAsync(b.ptr, a.ptr, nbytes, D2D, None), "hipMemcpyAsync")
    synchronize()
    check(h.hipEventRecord(e0, None), "hipEventRecord")
    for _ in range(reps):
        check(h.hipMemcpyAsync(b.ptr, a.ptr, nbytes, D2D, None), "hipMemcpyAsync")
    check(h.hipEventRecord(e1, None), "hipEventRecord")
    check(h.hipEventSynchronize(e1), "hipEventSynchronize")
    ms = C.c_float()
    check(h.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
    h.hipEventDestroy(e0)
    h.hipEventDestroy(e1)
    a.free()
    b.free()
    return 2.0 * nbytes * reps / (ms.value * 1e-3)


def copy_once(nbytes: int) -> None:
    """One device-to-device copy of nbytes (a known-size kernel for scaling
    HBM byte counters)."""
    a, b = DeviceArray(nbytes=nbytes), DeviceArray(nbytes=nbytes)
    synchronize()
    check(hip().hipMemcpyAsync(b.ptr, a.ptr, nbytes, D2D, None), "hipMemcpyAsync")
    synchronize()
    a.free()
    b.free()
