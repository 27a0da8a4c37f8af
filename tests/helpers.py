"""Shared helpers for the tests: synthetic chains, oracle runs, GPU runs."""
from __future__ import annotations

import ctypes

import numpy as np

from oracle import orc
from vk_compute_mipmaps_amd import pyramid as P
from vk_compute_mipmaps_amd import synth


def level0_srgba8(w, h, seed, opaque=False):
    l0 = synth.srgba8(w, h, seed)
    if opaque:
        l0[..., 3] = 255
    return l0


def host_chain(w, h, level0: np.ndarray, levels=0, bpt=4) -> np.ndarray:
    offs, total = P.packed_layout(w, h, levels, bpt)
    chain = np.zeros(total, dtype=np.uint8)
    raw = np.ascontiguousarray(level0).view(np.uint8).reshape(-1)
    chain[: raw.size] = raw
    return chain


def to_orc_steps(steps):
    out = []
    for s in steps:
        o = orc.Step()
        for name, _ in orc.Step._fields_:
            setattr(o, name, getattr(s, name))
        out.append(o)
    return out


def emulate_srgba8(chain, w, h, steps, mip_levels=0, precision="f32"):
    return orc.pyramid_srgba8(chain, w, h, to_orc_steps(steps), mip_levels, precision)


def gpu_run(w, h, level0, stream, fmt=P.FORMAT_SRGBA8, fast=True, mip_levels=0, levels=0,
            general_dispatcher=None, fast_dispatcher=None, arithmetic=P.ARITH_FP32):
    """Upload level 0 into a poisoned packed chain, dispatch, download."""
    from vk_compute_mipmaps_amd import hip_rt

    chain = P.DeviceMipChain(w, h, fmt, levels)
    try:
        chain.upload_level0(level0)
        chain.poison(stream=stream)
        chain.dispatch(stream=stream, fast=fast, mip_levels=mip_levels, arithmetic=arithmetic,
                       general_dispatcher=general_dispatcher, fast_dispatcher=fast_dispatcher)
        hip_rt.stream_sync(stream)
        return chain.download()
    finally:
        chain.free()


def level_slices(w, h, levels=0, bpt=4):
    offs, total = P.packed_layout(w, h, levels, bpt)
    out = []
    for l, o in enumerate(offs):
        lw, lh = P.level_dims(w, h, l)
        out.append((o, lw, lh))
    return out


def level_view(chain: np.ndarray, w, h, level, bpt=4):
    for l, (o, lw, lh) in enumerate(level_slices(w, h, bpt=bpt)):
        if l == level:
            return chain[o:o + lw * lh * bpt].reshape(lh, lw, bpt)
    raise IndexError(level)


def rgb_alpha_stats(a: np.ndarray, b: np.ndarray, w, h):
    """(max |a-b|, fraction of RGB channels equal, fraction of texels equal) over levels >= 1."""
    start = w * h * 4
    d = np.abs(a[start:].astype(np.int16) - b[start:].astype(np.int16)).reshape(-1, 4)
    if d.size == 0:
        return 0, 1.0, 1.0
    return int(d.max()), float((d[:, :3] == 0).mean()), float((d == 0).all(axis=1).mean())


def py_dispatcher(fn):
    """Wrap a Python (cmd, state) -> levels callable as nvpro_pyramid_dispatcher_t."""
    cb = P.DISPATCHER(fn)
    return cb


def fast_dispatcher_ex(divisibility, max_levels):
    L = P.lib()

    def fn(cmd, state):
        return L.nvproPyramidFastDispatcherEx(ctypes.c_void_p(cmd), state, divisibility, max_levels)

    return P.DISPATCHER(fn)
