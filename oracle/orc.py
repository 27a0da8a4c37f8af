"""ctypes wrapper of the CPU oracle (oracle/orc.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "liborc.so")

PIPE_GENERAL = 0
PIPE_FAST = 1


class Step(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "pipeline", "inputLevel", "levelCount", "pushConstant", "groupCountX",
        "inputWidth", "inputHeight", "reserved")]

    def as_tuple(self):
        return (self.pipeline, self.inputLevel, self.levelCount, self.groupCountX)


def build() -> str:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u32, u64, p = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p
        L.orc_full_levels.restype = u32
        L.orc_full_levels.argtypes = [u32, u32]
        L.orc_level_dims.argtypes = [u32, u32, u32, ctypes.POINTER(u32), ctypes.POINTER(u32)]
        L.orc_packed_offsets.restype = u64
        L.orc_packed_offsets.argtypes = [u32, u32, u32, p]
        L.orc_plan.restype = ctypes.c_int
        L.orc_plan.argtypes = [u32, u32, u32, ctypes.c_int, u32, u32, p, u32]
        L.orc_srgb_decode_f64.restype = ctypes.c_double
        L.orc_srgb_decode_f64.argtypes = [u32]
        L.orc_srgb_decode_f32.restype = ctypes.c_float
        L.orc_srgb_decode_f32.argtypes = [u32]
        L.orc_alpha_decode_f32.restype = ctypes.c_float
        L.orc_alpha_decode_f32.argtypes = [u32]
        L.orc_srgb_encode_exact.restype = u32
        L.orc_srgb_encode_exact.argtypes = [ctypes.c_double]
        L.orc_alpha_encode_f32.restype = u32
        L.orc_alpha_encode_f32.argtypes = [ctypes.c_float]
        L.orc_alpha_encode_f64.restype = u32
        L.orc_alpha_encode_f64.argtypes = [ctypes.c_double]
        L.orc_srgb_threshold.restype = ctypes.c_float
        L.orc_srgb_threshold.argtypes = [u32]
        for name in ("orc_pyramid_srgba8_f32", "orc_pyramid_srgba8_f64"):
            getattr(L, name).restype = ctypes.c_int
            getattr(L, name).argtypes = [p, u32, u32, u32, p, u32]
        L.orc_pyramid_rgba32f_f32.restype = ctypes.c_int
        L.orc_pyramid_rgba32f_f32.argtypes = [p, u32, u32, u32, p, u32]
        L.orc_pyramid_linear_f64.argtypes = [p, u32, u32, u32]
        L.orc_half_to_double.restype = ctypes.c_double
        L.orc_half_to_double.argtypes = [ctypes.c_uint16]
        L.orc_half_from_double.restype = ctypes.c_uint16
        L.orc_half_from_double.argtypes = [ctypes.c_double]
        L.orc_pyramid_rgba16f.restype = ctypes.c_int
        L.orc_pyramid_rgba16f.argtypes = [p, u32, u32, u32, p, u32, ctypes.c_int]
        L.orc_premultiply_srgba8.argtypes = [p, u64, ctypes.c_int]
        L.orc_reference_cpu_srgba8.argtypes = [p, u32, u32]
        L.orc_reference_cpu_srgba8_mt.argtypes = [p, u32, u32, u32]
        L.orc_compare_srgba8.restype = u32
        L.orc_compare_srgba8.argtypes = [p, p, u32, u32, u32] + [ctypes.POINTER(u32)] * 4 + [
            ctypes.POINTER(u64), ctypes.POINTER(u64)]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def full_levels(w, h):
    return lib().orc_full_levels(w, h)


def level_dims(w, h, level):
    a, b = ctypes.c_uint32(), ctypes.c_uint32()
    lib().orc_level_dims(w, h, level, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def packed_offsets(w, h, levels=0):
    levels = levels or full_levels(w, h)
    offs = np.zeros(levels, dtype=np.uint64)
    total = lib().orc_packed_offsets(w, h, levels, _ptr(offs))
    return offs, total


def plan(w, h, mip_levels=0, fast=True, divisibility=4, max_fast=6):
    steps = (Step * 64)()
    n = lib().orc_plan(w, h, mip_levels, 1 if fast else 0, divisibility, max_fast,
                       ctypes.cast(steps, ctypes.c_void_p), 64)
    if n < 0:
        raise ValueError("plan violates the dispatcher contract")
    return [steps[i] for i in range(n)]


def _steps_array(steps):
    arr = (Step * max(1, len(steps)))()
    for i, s in enumerate(steps):
        arr[i] = s
    return arr


def pyramid_srgba8(chain: np.ndarray, w, h, steps, mip_levels=0, precision="f32") -> np.ndarray:
    """Run an emulator on a packed uint8 chain (level 0 filled); returns it."""
    out = np.ascontiguousarray(chain, dtype=np.uint8).copy()
    arr = _steps_array(steps)
    fn = lib().orc_pyramid_srgba8_f32 if precision == "f32" else lib().orc_pyramid_srgba8_f64
    rc = fn(_ptr(out), w, h, mip_levels, ctypes.cast(arr, ctypes.c_void_p), len(steps))
    if rc != 0:
        raise ValueError("emulator rejected the plan")
    return out


def pyramid_rgba32f(chain: np.ndarray, w, h, steps, mip_levels=0) -> np.ndarray:
    out = np.ascontiguousarray(chain, dtype=np.float32).copy()
    arr = _steps_array(steps)
    rc = lib().orc_pyramid_rgba32f_f32(_ptr(out), w, h, mip_levels, ctypes.cast(arr, ctypes.c_void_p),
                                       len(steps))
    if rc != 0:
        raise ValueError("emulator rejected the plan")
    return out


def pyramid_rgba16f(chain: np.ndarray, w, h, steps, mip_levels=0, packed=False) -> np.ndarray:
    """RGBA16F emulator on a packed chain of uint16 half bits; returns it."""
    out = np.ascontiguousarray(chain).view(np.uint16).copy()
    arr = _steps_array(steps)
    rc = lib().orc_pyramid_rgba16f(_ptr(out), w, h, mip_levels, ctypes.cast(arr, ctypes.c_void_p), len(steps),
                                   1 if packed else 0)
    if rc != 0:
        raise ValueError("emulator rejected the plan")
    return out


def pyramid_linear_f64(chain: np.ndarray, w, h, mip_levels=0) -> np.ndarray:
    out = np.ascontiguousarray(chain, dtype=np.float64).copy()
    lib().orc_pyramid_linear_f64(_ptr(out), w, h, mip_levels)
    return out


def premultiply_srgba8(texels: np.ndarray, exact=True) -> np.ndarray:
    """Premultiplied alpha of sRGBA8 texels (scoped_image.hpp:233-252); returns a copy."""
    out = np.ascontiguousarray(texels, dtype=np.uint8).copy()
    lib().orc_premultiply_srgba8(_ptr(out), out.size // 4, 1 if exact else 0)
    return out


def reference_cpu(chain: np.ndarray, w, h, threads=1) -> np.ndarray:
    out = np.ascontiguousarray(chain, dtype=np.uint8).copy()
    lib().orc_reference_cpu_srgba8_mt(_ptr(out), w, h, threads)
    return out


def compare(a: np.ndarray, b: np.ndarray, w, h, levels=0):
    """MipmapStorage::compare: (worst delta, (x, y, level, channel), diff texels, total texels)."""
    a = np.ascontiguousarray(a, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.uint8)
    u = [ctypes.c_uint32() for _ in range(4)]
    d, t = ctypes.c_uint64(), ctypes.c_uint64()
    worst = lib().orc_compare_srgba8(_ptr(a), _ptr(b), w, h, levels, *[ctypes.byref(x) for x in u],
                                     ctypes.byref(d), ctypes.byref(t))
    return worst, tuple(x.value for x in u), d.value, t.value
