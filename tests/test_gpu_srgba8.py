"""GPU parity of the sRGBA8 path through the C-ABI (nvproPyramidDispatch).

The bar: every stored texel of levels >= 1 equals the FP32 oracle that
replays the reference schedule (same plan, same per-stage summation order,
exact conversions), and differs from the FP64 filter oracle by at most 1 LSB.
Inputs are the seeded synthetic images of the benchmark configurations.
"""
import numpy as np
import pytest

from oracle import orc
from tests import helpers as H
from vk_compute_mipmaps_amd import pyramid as P

pytestmark = pytest.mark.gpu


def _check_exact(w, h, seed, stream, fast=True, mip_levels=0, opaque=False):
    l0 = H.level0_srgba8(w, h, seed, opaque)
    gpu = H.gpu_run(w, h, l0, stream, fast=fast, mip_levels=mip_levels)
    steps = P.plan(w, h, mip_levels, fast=fast)
    host = H.host_chain(w, h, l0)
    emu = H.emulate_srgba8(host, w, h, steps, mip_levels)
    n_levels = mip_levels or P.full_level_count(w, h)
    offs, _ = P.packed_layout(w, h)
    end = offs[n_levels] if n_levels < len(offs) else gpu.size
    worst, where, ndiff, total = orc.compare(gpu[:end], emu[:end], w, h, n_levels)
    assert worst == 0, f"{w}x{h} seed {seed}: GPU differs from the FP32 oracle by {worst} at {where} ({ndiff}/{total} texels)"
    f64 = H.emulate_srgba8(host, w, h, steps, mip_levels, precision="f64")
    worst64, where64, _, _ = orc.compare(gpu[:end], f64[:end], w, h, n_levels)
    assert worst64 <= 1, f"{w}x{h}: {worst64} LSB from the FP64 oracle at {where64}"
    # The exact-match rate is a statistic: only meaningful over many texels.
    # The bar: >= 99.9 % of texels (all four channels) equal the FP64 oracle.
    if n_levels == P.full_level_count(w, h) and w * h >= 65536:
        _, rgb_exact, texel_exact = H.rgb_alpha_stats(gpu, f64, w, h)
        assert rgb_exact >= 0.999, f"{w}x{h}: only {rgb_exact:.5f} of RGB channels match FP64 exactly"
        assert texel_exact >= 0.999, f"{w}x{h}: only {texel_exact:.5f} of texels match FP64 exactly"
    # Nothing outside the requested levels was written.
    if end < gpu.size:
        assert (gpu[end:].view(np.uint32) == P.POISON_SRGBA8).all()
    return gpu


# configs: 1024^2 (seed 1), 4096^2 (seed 2), odd sizes (seeds 3-5)
@pytest.mark.parametrize("w,h,seed", [(1024, 1024, 1), (4096, 4096, 2), (4095, 2047, 3), (1920, 1080, 4),
                                      (2047, 4097, 5)])
def test_benchmark_configs_exact(w, h, seed, stream):
    _check_exact(w, h, seed, stream)


# Sizes of the reference's test images (docs/test_images.txt), each a distinct schedule.
@pytest.mark.parametrize("w,h", [(2560, 1440), (3840, 2160), (4094, 4094), (4095, 4095), (2048, 2048),
                                 (2052, 2052), (2047, 2047), (3095, 990), (1080, 4096)])
def test_reference_image_sizes_exact(w, h, stream):
    _check_exact(w, h, 11, stream)


@pytest.mark.parametrize("w,h", [(1024, 1024), (1920, 1080), (4096, 4096), (3095, 990)])
def test_general_only_exact(w, h, stream):
    """-force-no-fast-pipeline (minimal_mipmaps.cpp:266-270): general pipeline alone."""
    g = _check_exact(w, h, 12, stream, fast=False)
    f = H.gpu_run(w, h, H.level0_srgba8(w, h, 12), stream, fast=True)
    worst, where, _, _ = orc.compare(f, g, w, h)
    assert worst <= 1, f"fast vs general-only differ by {worst} at {where}"


SMALL = [1, 2, 3, 4, 5, 7, 8, 12, 16, 17, 31, 32, 33, 64, 65, 96, 127, 128, 129, 192, 256]


@pytest.mark.parametrize("w", SMALL)
def test_small_and_ragged_sizes_exact(w, stream):
    for h in SMALL:
        if w == 1 and h == 1:
            continue
        _check_exact(w, h, 100 + w * 7 + h, stream)


# Widths and heights around the two-level general tile (31 x 7 second-level
# texels: a first level of 63 x 15, one lane per column of a 64-lane wave) and
# around thin strips, where a lane lends its column to the lane before it.
TILE_W = [123, 124, 125, 126, 127, 249, 250, 251, 252, 253, 254, 255, 257, 499, 500, 501, 502]
TILE_H = [1, 2, 3, 5, 27, 28, 29, 30, 31, 57, 58, 59, 60, 61, 62]


@pytest.mark.parametrize("w", TILE_W)
def test_general_tile_edges_exact(w, stream):
    for h in TILE_H:
        _check_exact(w, h, 300 + w + 3 * h, stream)
        _check_exact(h, w, 301 + w + 3 * h, stream)


def test_one_by_one_is_a_no_op(stream):
    """A 1x1 image has a single level: nothing to fill."""
    l0 = H.level0_srgba8(1, 1, 3)
    gpu = H.gpu_run(1, 1, l0, stream)
    assert (gpu == l0.reshape(-1)).all()


@pytest.mark.parametrize("mip_levels", [2, 3, 4, 7])
def test_partial_chains_exact(mip_levels, stream):
    for (w, h) in [(256, 256), (300, 200), (1024, 768)]:
        _check_exact(w, h, 20 + mip_levels, stream, mip_levels=mip_levels)


@pytest.mark.parametrize("div,maxl", [(2, 6), (4, 5), (4, 3), (4, 1), (8, 6)])
def test_custom_fast_dispatchers_exact(div, maxl, stream):
    """The levels_1_6 / levels_1_5 / levels_1_3 variants of the dispatcher template
    (nvpro_pyramid_dispatch.hpp:195; strategy.md.html:425-539)."""
    fast = H.fast_dispatcher_ex(div, maxl)
    for (w, h) in [(4094, 4094), (2052, 2052), (1920, 1080), (512, 512), (1080, 4096)]:
        l0 = H.level0_srgba8(w, h, 30)
        gpu = H.gpu_run(w, h, l0, stream, fast_dispatcher=fast)
        steps = orc.plan(w, h, 0, True, div, maxl)
        emu = orc.pyramid_srgba8(H.host_chain(w, h, l0), w, h, steps)
        worst, where, ndiff, _ = orc.compare(gpu, emu, w, h)
        assert worst == 0, f"{w}x{h} <{div},{maxl}>: {worst} at {where} ({ndiff} texels)"


def test_null_dispatcher_leaves_poison(stream):
    """Negative control: a dispatcher that claims levels without dispatching
    (nvpro_pyramid_dispatch_alternative.cpp:66-77) leaves the magenta poison, which
    the comparison must flag."""
    null = H.py_dispatcher(lambda cmd, state: min(2, state.contents.remainingLevels))
    w, h = 512, 256
    l0 = H.level0_srgba8(w, h, 31)
    gpu = H.gpu_run(w, h, l0, stream, fast=False, general_dispatcher=null)
    assert (gpu[w * h * 4:].view(np.uint32) == P.POISON_SRGBA8).all()
    emu = H.emulate_srgba8(H.host_chain(w, h, l0), w, h, P.plan(w, h))
    assert orc.compare(gpu, emu, w, h)[0] > 100


def test_custom_layout_matches_packed(stream):
    """Levels at arbitrary offsets with padded rows give the packed layout's texels."""
    from vk_compute_mipmaps_amd import hip_rt

    w, h = 1000, 600
    l0 = H.level0_srgba8(w, h, 32)
    packed = H.gpu_run(w, h, l0, stream)
    levels = P.full_level_count(w, h)
    offsets, pitches, off = [], [], 4096
    for l in range(levels):
        lw, lh = P.level_dims(w, h, l)
        pitch = (lw * 4 + 255) // 256 * 256 + 4 * (l % 3)
        offsets.append(off)
        pitches.append(pitch)
        off += pitch * lh + 1024
    ptr = hip_rt.malloc(off)
    try:
        hip_rt.memset_d32_async(ptr, 0, off // 4, stream)
        raw = np.ascontiguousarray(l0)
        buf = np.zeros(pitches[0] * h, np.uint8)
        for y in range(h):
            buf[y * pitches[0]: y * pitches[0] + w * 4] = raw[y].reshape(-1)
        hip_rt.stream_sync(stream)
        hip_rt.memcpy_htod(ptr + offsets[0], buf.ctypes.data_as(P.ctypes.c_void_p), buf.nbytes)
        img = P.make_image(ptr, w, h, levels, offsets, pitches)
        P.nvpro_cmd_pyramid_dispatch(stream, P.pipelines(), img, 0)
        hip_rt.stream_sync(stream)
        out = np.empty(off, np.uint8)
        hip_rt.memcpy_dtoh(out.ctypes.data_as(P.ctypes.c_void_p), ptr, off)
    finally:
        hip_rt.free(ptr)
    for l, (o, lw, lh) in enumerate(H.level_slices(w, h)):
        got = np.stack([out[offsets[l] + y * pitches[l]: offsets[l] + y * pitches[l] + lw * 4] for y in range(lh)])
        np.testing.assert_array_equal(got.reshape(-1), packed[o:o + lw * lh * 4], err_msg=f"level {l}")
        for y in range(lh):  # row padding untouched
            pad = out[offsets[l] + y * pitches[l] + lw * 4: offsets[l] + (y + 1) * pitches[l]]
            assert not pad.any(), f"level {l} row {y} padding written"


def _device_chain(w, h, seed, stream):
    """Level 0 from the device generator (nvproPyramidSynthesize, bit-identical
    to synth.py: test_gpu_batch.py), a poisoned chain, one dispatch; returns the
    downloaded chain."""
    from vk_compute_mipmaps_amd import hip_rt

    chain = P.DeviceMipChain(w, h)
    try:
        chain.synthesize(seed, stream)
        chain.poison(stream=stream)
        chain.dispatch(stream=stream)
        hip_rt.stream_sync(stream)
        return chain.download(), chain.offsets
    finally:
        chain.free()


# Sizes whose first fast step runs on fewer persistent workgroups per compute
# unit than the 8 that fit: 5120^2 (6,400 groups: 4 per CU) and 8192x4096
# (8,192 groups: 3 per CU).
@pytest.mark.parametrize("w,h", [(5120, 5120), (8192, 4096)])
def test_large_grid_policies_exact(w, h, stream):
    gpu, _ = _device_chain(w, h, 40, stream)
    emu = H.emulate_srgba8(H.host_chain(w, h, gpu[: w * h * 4]), w, h, P.plan(w, h))
    worst, where, ndiff, total = orc.compare(gpu, emu, w, h)
    assert worst == 0, f"{w}x{h}: {worst} LSB from the FP32 oracle at {where} ({ndiff}/{total} texels)"


def test_16384_full_size(stream):
    """16384^2 (1 GiB base): plan F(0,6) F(6,6) F(12,2), the first step on 3
    persistent workgroups per compute unit.  No texel of levels >= 1 keeps the
    poison.  Each 64x64 tile of the first step depends only on its own input, so
    whole 64-row bands of level 0 (256 tiles each: the first, the last and a
    random one) are replayed by the oracle and compared on levels 1-6, and
    sampled single tiles elsewhere; the later steps are emulated in full from the
    GPU's level 6."""
    w = h = 16384
    steps = P.plan(w, h)
    assert [s.as_tuple()[:3] for s in steps] == [(1, 0, 6), (1, 6, 6), (1, 12, 2)]
    gpu, offs = _device_chain(w, h, 6, stream)
    assert not (gpu[offs[1]:].view(np.uint32) == P.POISON_SRGBA8).any(), "poison left in levels >= 1"
    rng = np.random.default_rng(6)
    band_plan = orc.plan(w, 64, 7, True, 4, 6)
    assert [(s.pipeline, s.inputLevel, s.levelCount) for s in band_plan] == [(1, 0, 6)]
    l0 = gpu[: w * h * 4].reshape(h, w, 4)
    for ty in sorted({0, 255, int(rng.integers(1, 255))}):
        band = H.host_chain(w, 64, l0[ty * 64: ty * 64 + 64])
        emu = orc.pyramid_srgba8(band, w, 64, band_plan, 7)
        for r in range(1, 7):
            s_ = 64 >> r
            got = gpu[offs[r]: offs[r] + (w >> r) * (h >> r) * 4].reshape(h >> r, w >> r, 4)[ty * s_:(ty + 1) * s_]
            np.testing.assert_array_equal(got, H.level_view(emu, w, 64, r), err_msg=f"band {ty} level {r}")
    tile_plan = orc.plan(64, 64, 7, True, 4, 6)
    for (tx, ty) in [tuple(t) for t in rng.integers(0, 256, (16, 2))]:
        tile = H.host_chain(64, 64, l0[ty * 64: ty * 64 + 64, tx * 64: tx * 64 + 64])
        emu = orc.pyramid_srgba8(tile, 64, 64, tile_plan, 7)
        for r in range(1, 7):
            s_ = 64 >> r
            got = gpu[offs[r]: offs[r] + (w >> r) * (h >> r) * 4].reshape(h >> r, w >> r, 4)[
                ty * s_:(ty + 1) * s_, tx * s_:(tx + 1) * s_]
            np.testing.assert_array_equal(got, H.level_view(emu, 64, 64, r), err_msg=f"tile {(tx, ty)} level {r}")
    # Steps 2 and 3 from the GPU's level 6 (256 x 256).
    sub = gpu[offs[6]:].copy()
    emu = H.emulate_srgba8(sub, 256, 256, P.plan(256, 256))
    np.testing.assert_array_equal(gpu[offs[6]:], emu)
