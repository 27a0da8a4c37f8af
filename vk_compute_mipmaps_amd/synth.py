"""Seeded synthetic level-0 images (the benchmark workloads' inputs).

All values come from a counter-based hash, so any texel can be produced
independently and in any order:

    mix(z)               = splitmix64 finaliser of z + 0x9E3779B97F4A7C15
    h(seed, s, i, j)     = mix(mix(mix(seed ^ (s << 56)) ^ i) ^ j)      (uint64)

sRGBA8 (4 bytes/texel, R G B A in memory):
    for channel c in R, G, B:
        value noise with a 64-px period: lattice value L(i, j) = h(seed, c, i, j) >> 56,
        smoothstep weight s(t) = t^2 (3 - 2t) with t = (x mod 64) / 64, kept exact as the
        integer sx = m^2 (192 - 2m), m = x mod 64, over the denominator 64^3;
        v = round-half-up of the bilinear blend of the four lattice values
        plus uniform integer noise in [-16, 16]: (h(seed, 4 + c, x, y) >> 32) mod 33 - 16,
        clamped to [0, 255]
    alpha: h(seed, 7, x, y) >> 56, uniform on [0, 255], not premultiplied.

RGBA16F (8 bytes/texel, IEEE half R G B A):
    RGB: exp(U) with U uniform on [ln 1e-3, ln 1e3), U from h(seed, c, x, y) >> 11 * 2^-53;
    alpha: uniform on [0, 1) from h(seed, 7, x, y); all rounded to half.
"""
from __future__ import annotations

import math

import numpy as np

_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xBF58476D1CE4E5B9)
_M3 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = z + _M1
        z = (z ^ (z >> np.uint64(30))) * _M2
        z = (z ^ (z >> np.uint64(27))) * _M3
    return z ^ (z >> np.uint64(31))


def hash4(seed: int, stream: int, i: np.ndarray, j: np.ndarray) -> np.ndarray:
    base = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) ^ np.uint64(stream << 56)
    z = _mix(np.asarray(base, dtype=np.uint64))
    z = _mix(z ^ i.astype(np.uint64))
    return _mix(z ^ j.astype(np.uint64))


def srgba8(width: int, height: int, seed: int, rows=None) -> np.ndarray:
    """Level 0 as a (height, width, 4) uint8 array (rows: optional (y0, y1) band)."""
    y0, y1 = rows if rows is not None else (0, height)
    x = np.arange(width, dtype=np.uint64)[None, :]
    y = np.arange(y0, y1, dtype=np.uint64)[:, None]
    xs, ys = np.broadcast_to(x, (y1 - y0, width)), np.broadcast_to(y, (y1 - y0, width))
    out = np.empty((y1 - y0, width, 4), dtype=np.uint8)
    i, j = xs >> np.uint64(6), ys >> np.uint64(6)
    mx, my = xs & np.uint64(63), ys & np.uint64(63)
    D = np.uint64(1 << 18)
    sx = mx * mx * (np.uint64(192) - np.uint64(2) * mx)
    sy = my * my * (np.uint64(192) - np.uint64(2) * my)
    for c in range(3):
        L00 = hash4(seed, c, i, j) >> np.uint64(56)
        L10 = hash4(seed, c, i + np.uint64(1), j) >> np.uint64(56)
        L01 = hash4(seed, c, i, j + np.uint64(1)) >> np.uint64(56)
        L11 = hash4(seed, c, i + np.uint64(1), j + np.uint64(1)) >> np.uint64(56)
        top = L00 * (D - sx) + L10 * sx
        bot = L01 * (D - sx) + L11 * sx
        num = top * (D - sy) + bot * sy
        v = ((num + np.uint64(1 << 35)) >> np.uint64(36)).astype(np.int64)
        noise = ((hash4(seed, 4 + c, xs, ys) >> np.uint64(32)) % np.uint64(33)).astype(np.int64) - 16
        out[..., c] = np.clip(v + noise, 0, 255).astype(np.uint8)
    out[..., 3] = (hash4(seed, 7, xs, ys) >> np.uint64(56)).astype(np.uint8)
    return out


def rgba16f(width: int, height: int, seed: int) -> np.ndarray:
    """Level 0 as a (height, width, 4) float16 array."""
    x = np.arange(width, dtype=np.uint64)[None, :]
    y = np.arange(height, dtype=np.uint64)[:, None]
    xs, ys = np.broadcast_to(x, (height, width)), np.broadcast_to(y, (height, width))
    out = np.empty((height, width, 4), dtype=np.float16)
    lo, hi = math.log(1e-3), math.log(1e3)
    scale = 2.0 ** -53
    for c in range(3):
        u = (hash4(seed, c, xs, ys) >> np.uint64(11)).astype(np.float64) * scale
        out[..., c] = np.exp(lo + u * (hi - lo)).astype(np.float16)
    out[..., 3] = ((hash4(seed, 7, xs, ys) >> np.uint64(11)).astype(np.float64) * scale).astype(np.float16)
    return out


def mixed_sizes(count: int, seed: int, lo: int = 64, hi: int = 4096, odd_fraction: float = 0.4):
    """Texture sizes of the batched workload: each side log-uniform on [lo, hi];
    with probability odd_fraction both sides are made odd, otherwise even."""
    idx = np.arange(count, dtype=np.uint64)
    u1 = (hash4(seed, 0, idx, idx) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    u2 = (hash4(seed, 1, idx, idx) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    u3 = (hash4(seed, 2, idx, idx) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    w = np.exp(math.log(lo) + u1 * (math.log(hi) - math.log(lo)))
    h = np.exp(math.log(lo) + u2 * (math.log(hi) - math.log(lo)))
    sizes = []
    for k in range(count):
        a, b = int(round(w[k])), int(round(h[k]))
        a, b = max(lo, min(hi, a)), max(lo, min(hi, b))
        if u3[k] < odd_fraction:
            a, b = a | 1, b | 1
            a, b = min(a, hi - 1), min(b, hi - 1)
        else:
            a, b = a & ~1, b & ~1
        sizes.append((a, b))
    return sizes
