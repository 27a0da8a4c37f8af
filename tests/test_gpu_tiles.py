"""The general pipeline's two tile geometries give the same bytes.

A two-level general step runs 31x7 tiles on persistent workgroups, or 31x3
thin tiles one per workgroup when all of them fit on the GPU at once.
NVPRO_PYRAMID_THIN=1 / =0 forces one geometry for every step, so each step of
these chains runs both ways (thin base steps of large images, full small steps
and tails' predecessors); all three runs must agree byte for byte, and the
default run must equal the oracle's FP32 emulator (sRGBA8) bit for bit.
"""
import os

import numpy as np
import pytest

from tests import helpers as H
from vk_compute_mipmaps_amd import hip_rt
from vk_compute_mipmaps_amd import pyramid as P
from vk_compute_mipmaps_amd import synth

pytestmark = pytest.mark.gpu

# Odd and even sides (kernel sizes 3x3, 2x3, 3x2, 2x2 and sides of 1), RGBA16F
# in both arithmetics and RGBA32F (the 16-byte geometries).
CASES = [(4095, 2047, P.FORMAT_SRGBA8, P.ARITH_FP32), (1542, 940, P.FORMAT_SRGBA8, P.ARITH_FP32),
         (1000, 777, P.FORMAT_SRGBA8, P.ARITH_FP32), (257, 1026, P.FORMAT_SRGBA8, P.ARITH_FP32),
         (1, 301, P.FORMAT_SRGBA8, P.ARITH_FP32), (1920, 1080, P.FORMAT_RGBA16F, P.ARITH_FP32),
         (1001, 513, P.FORMAT_RGBA16F, P.ARITH_PACKED_FP16), (645, 389, P.FORMAT_RGBA32F, P.ARITH_FP32)]


def _dispatch(w, h, fmt, arith, stream, mode):
    old = os.environ.get("NVPRO_PYRAMID_THIN")
    if mode is None:
        os.environ.pop("NVPRO_PYRAMID_THIN", None)
    else:
        os.environ["NVPRO_PYRAMID_THIN"] = mode
    c = P.DeviceMipChain(w, h, fmt)
    try:
        if fmt == P.FORMAT_RGBA32F:  # no device generator for this format
            rng = np.random.default_rng(11)
            c.upload_level0(rng.random((h, w, 4), dtype=np.float32) * np.float32(4.0))
        else:
            c.synthesize(11, stream)
        c.poison(stream=stream)
        c.dispatch(stream=stream, arithmetic=arith)
        hip_rt.stream_sync(stream)
        return c.download()
    finally:
        c.free()
        if old is None:
            os.environ.pop("NVPRO_PYRAMID_THIN", None)
        else:
            os.environ["NVPRO_PYRAMID_THIN"] = old


@pytest.mark.parametrize("w,h,fmt,arith", CASES)
def test_thin_and_full_tiles_agree(w, h, fmt, arith, stream):
    print(f"tiles {w}x{h} format {fmt} arithmetic {arith}", flush=True)
    auto = _dispatch(w, h, fmt, arith, stream, None)
    thin = _dispatch(w, h, fmt, arith, stream, "1")
    full = _dispatch(w, h, fmt, arith, stream, "0")
    assert np.array_equal(thin, auto), "thin tiles differ from the default choice"
    assert np.array_equal(full, auto), "full tiles differ from the default choice"
    if fmt == P.FORMAT_SRGBA8:
        level0 = synth.srgba8(w, h, 11)
        emu = H.emulate_srgba8(H.host_chain(w, h, level0), w, h, P.plan(w, h))
        assert np.array_equal(auto, emu), "differs from the oracle's FP32 emulator"
