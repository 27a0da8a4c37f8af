"""GPU parity of the float formats (RGBA32F, RGBA16F in FP32 and packed-FP16
arithmetic) against the oracle's emulators: bit-exact."""
import numpy as np
import pytest

from oracle import orc
from tests import helpers as H
from vk_compute_mipmaps_amd import pyramid as P
from vk_compute_mipmaps_amd import synth

pytestmark = pytest.mark.gpu

SIZES = [(1024, 1024), (1920, 1080), (4095, 2047), (2052, 2052), (257, 3), (64, 64), (1, 37)]


def _levels(w, h, bpt):
    offs, total = P.packed_layout(w, h, 0, bpt)
    return offs, total


@pytest.mark.parametrize("w,h", SIZES)
def test_rgba32f_exact(w, h, stream):
    rng = np.random.default_rng(w ^ h)
    l0 = rng.random((h, w, 4), dtype=np.float32) * np.float32(4.0)
    gpu = H.gpu_run(w, h, l0, stream, fmt=P.FORMAT_RGBA32F).view(np.float32)
    offs, total = _levels(w, h, 16)
    chain = np.zeros(total // 4, np.float32)
    chain[: l0.size] = l0.reshape(-1)
    emu = orc.pyramid_rgba32f(chain, w, h, H.to_orc_steps(P.plan(w, h, fmt=P.FORMAT_RGBA32F)))
    np.testing.assert_array_equal(gpu.view(np.uint32), emu.view(np.uint32))


@pytest.mark.parametrize("packed", [False, True])
@pytest.mark.parametrize("w,h", SIZES)
def test_rgba16f_exact(w, h, packed, stream):
    l0 = synth.rgba16f(w, h, 7)
    arith = P.ARITH_PACKED_FP16 if packed else P.ARITH_FP32
    gpu = H.gpu_run(w, h, l0, stream, fmt=P.FORMAT_RGBA16F, arithmetic=arith).view(np.uint16)
    offs, total = _levels(w, h, 8)
    chain = np.zeros(total // 2, np.uint16)
    chain[: l0.size] = l0.view(np.uint16).reshape(-1)
    emu = orc.pyramid_rgba16f(chain, w, h, H.to_orc_steps(P.plan(w, h, fmt=P.FORMAT_RGBA16F)), 0, packed)
    diff = np.nonzero(gpu != emu)[0]
    assert diff.size == 0, f"{diff.size} halves differ, first at element {diff[:5]}: gpu {gpu[diff[:5]]} emu {emu[diff[:5]]}"


@pytest.mark.parametrize("packed", [False, True])
def test_rgba16f_general_only_exact(packed, stream):
    w, h = 1024, 768
    l0 = synth.rgba16f(w, h, 9)
    arith = P.ARITH_PACKED_FP16 if packed else P.ARITH_FP32
    gpu = H.gpu_run(w, h, l0, stream, fmt=P.FORMAT_RGBA16F, fast=False, arithmetic=arith).view(np.uint16)
    offs, total = _levels(w, h, 8)
    chain = np.zeros(total // 2, np.uint16)
    chain[: l0.size] = l0.view(np.uint16).reshape(-1)
    emu = orc.pyramid_rgba16f(chain, w, h, H.to_orc_steps(P.plan(w, h, fast=False, fmt=P.FORMAT_RGBA16F)), 0,
                              packed)
    assert (gpu == emu).all()
