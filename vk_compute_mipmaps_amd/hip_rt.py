"""Minimal ctypes bindings of the HIP runtime used by the host mirror.

The product library links libamdhip64.so.7; these bindings open the same
runtime (same soname), so device pointers, streams and events created here are
the ones the library's entry points expect.  Nothing here initialises the GPU
until a function is called.
"""
from __future__ import annotations

import ctypes

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3

_rt = None


class HipError(RuntimeError):
    pass


def rt():
    global _rt
    if _rt is None:
        from . import pyramid  # load the product library first: it pulls in the runtime

        pyramid.lib()
        R = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        R.hipGetErrorString.restype = ctypes.c_char_p
        R.hipGetErrorString.argtypes = [ctypes.c_int]
        for name, args in {
            "hipMalloc": [ctypes.POINTER(vp), sz],
            "hipFree": [vp],
            "hipMemcpy": [vp, vp, sz, ctypes.c_int],
            "hipMemcpyAsync": [vp, vp, sz, ctypes.c_int, vp],
            "hipMemset": [vp, ctypes.c_int, sz],
            "hipMemsetAsync": [vp, ctypes.c_int, sz, vp],
            "hipMemsetD32Async": [vp, ctypes.c_uint32, sz, vp],
            "hipStreamCreate": [ctypes.POINTER(vp)],
            "hipStreamDestroy": [vp],
            "hipStreamSynchronize": [vp],
            "hipDeviceSynchronize": [],
            "hipGetDeviceCount": [ctypes.POINTER(ctypes.c_int)],
            "hipSetDevice": [ctypes.c_int],
            "hipEventCreate": [ctypes.POINTER(vp)],
            "hipEventRecord": [vp, vp],
            "hipEventSynchronize": [vp],
            "hipEventElapsedTime": [ctypes.POINTER(ctypes.c_float), vp, vp],
            "hipEventDestroy": [vp],
            "hipGetLastError": [],
            "hipHostMalloc": [ctypes.POINTER(vp), sz, ctypes.c_uint],
            "hipHostFree": [vp],
            "hipDeviceGetAttribute": [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int],
            "hipMemGetInfo": [ctypes.POINTER(sz), ctypes.POINTER(sz)],
        }.items():
            fn = getattr(R, name)
            fn.restype = ctypes.c_int
            fn.argtypes = args
        _rt = R
    return _rt


def check(code: int, what: str = "HIP call"):
    if code != 0:
        msg = rt().hipGetErrorString(code)
        raise HipError(f"{what} failed: {code} ({msg.decode() if msg else '?'})")


def device_count() -> int:
    n = ctypes.c_int(0)
    if rt().hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


def malloc(nbytes: int) -> int:
    p = ctypes.c_void_p()
    check(rt().hipMalloc(ctypes.byref(p), max(1, nbytes)), f"hipMalloc({nbytes})")
    return p.value


def free(ptr: int):
    if ptr:
        check(rt().hipFree(ctypes.c_void_p(ptr)), "hipFree")


def memcpy_htod(dst: int, src, nbytes: int):
    check(rt().hipMemcpy(ctypes.c_void_p(dst), src, nbytes, hipMemcpyHostToDevice), "hipMemcpy H2D")


def memcpy_dtoh(dst, src: int, nbytes: int):
    check(rt().hipMemcpy(dst, ctypes.c_void_p(src), nbytes, hipMemcpyDeviceToHost), "hipMemcpy D2H")


def memcpy_dtod_async(dst: int, src: int, nbytes: int, stream: int = 0):
    check(rt().hipMemcpyAsync(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, hipMemcpyDeviceToDevice,
                              ctypes.c_void_p(stream)), "hipMemcpyAsync D2D")


def memset_d32_async(dst: int, value: int, count: int, stream: int = 0):
    check(rt().hipMemsetD32Async(ctypes.c_void_p(dst), value & 0xFFFFFFFF, count, ctypes.c_void_p(stream)),
          "hipMemsetD32Async")


def stream_create() -> int:
    s = ctypes.c_void_p()
    check(rt().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
    return s.value


def stream_destroy(s: int):
    check(rt().hipStreamDestroy(ctypes.c_void_p(s)), "hipStreamDestroy")


def stream_sync(s: int = 0):
    check(rt().hipStreamSynchronize(ctypes.c_void_p(s)), "hipStreamSynchronize")


def device_sync():
    check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


def event_create() -> int:
    e = ctypes.c_void_p()
    check(rt().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
    return e.value


def event_record(e: int, stream: int = 0):
    check(rt().hipEventRecord(ctypes.c_void_p(e), ctypes.c_void_p(stream)), "hipEventRecord")


def event_sync(e: int):
    check(rt().hipEventSynchronize(ctypes.c_void_p(e)), "hipEventSynchronize")


def event_elapsed_ms(a: int, b: int) -> float:
    ms = ctypes.c_float()
    check(rt().hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(a), ctypes.c_void_p(b)),
          "hipEventElapsedTime")
    return ms.value


def event_destroy(e: int):
    check(rt().hipEventDestroy(ctypes.c_void_p(e)), "hipEventDestroy")


def mem_info():
    free_b, total_b = ctypes.c_size_t(), ctypes.c_size_t()
    check(rt().hipMemGetInfo(ctypes.byref(free_b), ctypes.byref(total_b)), "hipMemGetInfo")
    return free_b.value, total_b.value
