"""GPU parity of the level-0 conversions (nvproPyramidPremultiplyAlpha).

sRGBA8: every (code, alpha) pair once, against the oracle's restatement of the
staging premultiply (include/scoped_image.hpp:233-252) with the library's exact
conversions: bit-exact.  RGBA16F: rgb * a in FP32 rounded to binary16, against
numpy's float32 product and its round-to-nearest-even conversion: bit-exact.
Then a premultiplied image goes through the chain like any other.
"""
import numpy as np
import pytest

from oracle import orc
from tests import helpers as H
from tests.test_io import _all_pairs
from vk_compute_mipmaps_amd import hip_rt
from vk_compute_mipmaps_amd import pyramid as P
from vk_compute_mipmaps_amd import synth

pytestmark = pytest.mark.gpu


def test_premultiply_srgba8_every_pair(stream):
    t = _all_pairs()
    c = P.DeviceMipChain(256, 256, P.FORMAT_SRGBA8)
    try:
        c.upload_level0(t)
        c.premultiply(stream)
        hip_rt.stream_sync(stream)
        got = c.download()[: t.size]
    finally:
        c.free()
    want = orc.premultiply_srgba8(t, exact=True)
    assert np.array_equal(got, want)


def test_premultiply_rgba16f(stream):
    w, h = 300, 77
    c = P.DeviceMipChain(w, h, P.FORMAT_RGBA16F)
    try:
        c.synthesize(5, stream)
        hip_rt.stream_sync(stream)
        before = c.download()[: w * h * 8].view(np.float16).reshape(-1, 4).copy()
        c.premultiply(stream)
        hip_rt.stream_sync(stream)
        got = c.download()[: w * h * 8].view(np.float16).reshape(-1, 4)
    finally:
        c.free()
    f = before.astype(np.float32)
    want = before.copy()
    want[:, :3] = (f[:, :3] * f[:, 3:4]).astype(np.float16)
    assert np.array_equal(got.view(np.uint16), want.view(np.uint16))


def test_premultiplied_chain_matches_oracle(stream):
    w, h = 512, 384
    l0 = orc.premultiply_srgba8(synth.srgba8(w, h, 31), exact=True)
    c = P.DeviceMipChain(w, h, P.FORMAT_SRGBA8)
    try:
        c.upload_level0(synth.srgba8(w, h, 31))
        c.premultiply(stream)
        c.dispatch(stream)
        hip_rt.stream_sync(stream)
        got = c.download()
    finally:
        c.free()
    emu = H.emulate_srgba8(H.host_chain(w, h, l0), w, h, P.plan(w, h))
    worst, where, _, _ = orc.compare(got, emu, w, h)
    assert worst == 0 and np.array_equal(got[: l0.size], l0.reshape(-1))
