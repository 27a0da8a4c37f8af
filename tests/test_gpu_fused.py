"""Fused launches give the same bytes as one launch per step.

When the stream is not being captured, nvproPyramidDispatch runs a fast
step's successor inside the step's own launch, in the workgroup that finishes
last: a fast step of one group (the last step of 4096^2 and 2048^2) or a tail
whose first level has at most 128 texels (8192^2 after F(6,6), 128^2), when
the step's grid is small (at most 64 workgroups) or large (at least 1024).  Plans
of general steps launch as before.  NVPRO_PYRAMID_FUSE=0 turns fusion off.
Both must agree byte for byte, also when the chain is dispatched again on the
same stream (the arrival counters are left zeroed for the next launch), on a
second stream (each stream has its own counters) and on more streams than
there are counter blocks (64 per device; later streams do not fuse).
"""
import ctypes
import os

import numpy as np
import pytest

from tests import helpers as H
from vk_compute_mipmaps_amd import hip_rt
from vk_compute_mipmaps_amd import pyramid as P

pytestmark = pytest.mark.gpu

REPEATS = 3


def _run(chain, stream, fused, **kw):
    old = os.environ.get("NVPRO_PYRAMID_FUSE")
    if fused:
        os.environ.pop("NVPRO_PYRAMID_FUSE", None)
    else:
        os.environ["NVPRO_PYRAMID_FUSE"] = "0"
    try:
        chain.poison(stream=stream)
        chain.dispatch(stream=stream, **kw)
        launches = P.last_launch_count()
        hip_rt.stream_sync(stream)
        return chain.download(), launches
    finally:
        if old is None:
            os.environ.pop("NVPRO_PYRAMID_FUSE", None)
        else:
            os.environ["NVPRO_PYRAMID_FUSE"] = old


def _check(chain, stream, **kw):
    """Separate launches, then fused ones; returns (bytes, launches separate, launches fused)."""
    plain, n_plain = _run(chain, stream, False, **kw)
    other = hip_rt.stream_create()
    try:
        for k in range(REPEATS):
            got, n_fused = _run(chain, stream, True, **kw)
            assert np.array_equal(got, plain), f"fused dispatch {k} differs from separate launches"
        got, n_other = _run(chain, other, True, **kw)
        assert np.array_equal(got, plain), "fused dispatch on a second stream differs"
        assert n_other == n_fused
    finally:
        hip_rt.stream_sync(other)
        P.stream_release(other)
        hip_rt.stream_destroy(other)
    # The requested levels were all written.
    levels = kw.get("mip_levels", 0) or P.full_level_count(chain.width, chain.height)
    offs = chain.offsets
    end = offs[levels] if levels < len(offs) else plain.size
    if chain.fmt == P.FORMAT_SRGBA8:
        assert not (plain[offs[1]:end].view(np.uint32) == P.POISON_SRGBA8).any()
    return plain, n_plain, n_fused


# (size, dispatch arguments, launches separate, launches fused)
@pytest.mark.parametrize("w,h,kw,n_plain,n_fused", [
    (4096, 4096, {}, 2, 1),                    # F(0,6) + F(6,6)
    (1024, 1024, {}, 2, 2),                    # F(0,6) of 256 workgroups: not fused
    (2048, 2048, {}, 2, 1),                    # F(0,6) of 1024 workgroups + F(6,5)
    (128, 128, {}, 2, 1),                      # F(0,6) + tail G(6,1)
    (8192, 8192, {}, 3, 2),                    # F(0,6), F(6,6) + tail G(12,1)
    (4096, 4096, {"mip_levels": 9}, 2, 1),     # F(0,6) + F(6,2), levels 9.. untouched
    (4095, 2047, {}, 4, 4),                    # G(0,2), G(2,2), G(4,2), tail
    (1920, 1080, {}, 3, 3),                    # F(0,3), G(3,2), tail
    (2047, 4097, {}, 4, 4),
    (3095, 990, {}, None, None),
    (1024, 1024, {"fast": False}, None, None),  # general steps only
])
def test_fused_equals_separate_srgba8(w, h, kw, n_plain, n_fused, stream):
    chain = P.DeviceMipChain(w, h, P.FORMAT_SRGBA8)
    try:
        chain.synthesize(5, stream)
        _, got_plain, got_fused = _check(chain, stream, **kw)
    finally:
        chain.free()
    if n_plain is not None:
        assert (got_plain, got_fused) == (n_plain, n_fused)
    assert got_fused <= got_plain


def test_fused_custom_plan_srgba8(stream):
    """Three-level fast steps F(0,3) F(3,3) F(6,3) F(9,3).  Separate: F(0,3),
    F(3,3), then one tail launch for F(6,3) F(9,3).  Fused: F(0,3), F(3,3) with
    F(6,3) (a single group) in its last workgroup, then F(9,3): as many
    launches, different kernels, same bytes."""
    chain = P.DeviceMipChain(4096, 4096, P.FORMAT_SRGBA8)
    try:
        chain.synthesize(6, stream)
        _, n_plain, n_fused = _check(chain, stream, fast_dispatcher=H.fast_dispatcher_ex(8, 3))
    finally:
        chain.free()
    assert (n_plain, n_fused) == (3, 3)


@pytest.mark.parametrize("fmt,arith", [(P.FORMAT_RGBA16F, P.ARITH_FP32), (P.FORMAT_RGBA16F, P.ARITH_PACKED_FP16),
                                       (P.FORMAT_RGBA32F, P.ARITH_FP32)], ids=["rgba16f", "rgba16f_packed", "rgba32f"])
@pytest.mark.parametrize("w,h", [(2048, 2048), (1023, 767)])
def test_fused_equals_separate_float(fmt, arith, w, h, stream):
    chain = P.DeviceMipChain(w, h, fmt)
    try:
        if fmt == P.FORMAT_RGBA16F:
            chain.synthesize(7, stream)
        else:
            rng = np.random.default_rng(7)
            chain.upload_level0(rng.uniform(0.0, 4.0, (h, w, 4)).astype(np.float32))
        _check(chain, stream, arithmetic=arith)
    finally:
        chain.free()


def test_fused_concurrent_streams_srgba8(stream):
    """Fused dispatches on two streams in flight at once (each stream has its
    own arrival counters): every chain still matches its separate-launch bytes."""
    sizes = [(4096, 4096), (2048, 2048), (8192, 8192)]
    chains = [P.DeviceMipChain(w, h, P.FORMAT_SRGBA8) for w, h in sizes]
    other = hip_rt.stream_create()
    try:
        want = []
        for k, c in enumerate(chains):
            c.synthesize(20 + k, stream)
            want.append(_run(c, stream, False)[0])
        streams = [stream, other]
        for rep in range(6):
            for k, c in enumerate(chains):
                c.dispatch(stream=streams[(rep + k) % 2])
        hip_rt.stream_sync(stream)
        hip_rt.stream_sync(other)
        for k, c in enumerate(chains):
            assert np.array_equal(c.download(), want[k]), f"{sizes[k]}: concurrent fused dispatches differ"
    finally:
        hip_rt.stream_sync(other)
        P.stream_release(other)
        hip_rt.stream_destroy(other)
        for c in chains:
            c.free()


@pytest.mark.parametrize("w,h", [(2048, 2048), (128, 128), (4096, 64)])
def test_fused_custom_layout_srgba8(w, h, stream):
    """Fused launches with levels at arbitrary offsets and padded rows: the same
    texels as the packed layout, padding untouched."""
    l0 = H.level0_srgba8(w, h, 33)
    packed = H.gpu_run(w, h, l0, stream)
    levels = P.full_level_count(w, h)
    offsets, pitches, off = [], [], 4096
    for lvl in range(levels):
        lw, lh = P.level_dims(w, h, lvl)
        pitch = (lw * 4 + 255) // 256 * 256 + 4 * (lvl % 3)
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
        hip_rt.memcpy_htod(ptr + offsets[0], buf.ctypes.data_as(ctypes.c_void_p), buf.nbytes)
        img = P.make_image(ptr, w, h, levels, offsets, pitches)
        P.nvpro_cmd_pyramid_dispatch(stream, P.pipelines(), img, 0)
        launches = P.last_launch_count()
        hip_rt.stream_sync(stream)
        out = np.empty(off, np.uint8)
        hip_rt.memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), ptr, off)
    finally:
        hip_rt.free(ptr)
    assert launches < len(P.plan(w, h)), "expected a fused dispatch"
    for lvl, (o, lw, lh) in enumerate(H.level_slices(w, h)):
        got = np.stack([out[offsets[lvl] + y * pitches[lvl]: offsets[lvl] + y * pitches[lvl] + lw * 4]
                        for y in range(lh)])
        np.testing.assert_array_equal(got.reshape(-1), packed[o:o + lw * lh * 4], err_msg=f"level {lvl}")
        for y in range(lh):
            pad = out[offsets[lvl] + y * pitches[lvl] + lw * 4: offsets[lvl] + (y + 1) * pitches[lvl]]
            assert not pad.any(), f"level {lvl} row {y} padding written"


def test_fused_more_streams_than_counter_blocks_srgba8(stream):
    """80 streams, each dispatching 4096^2 twice: the streams that get a counter
    block fuse (1 launch), the rest do not (2 launches), and every result is the
    separate-launch bytes."""
    chain = P.DeviceMipChain(4096, 4096, P.FORMAT_SRGBA8)
    streams = []
    try:
        chain.synthesize(9, stream)
        want, _ = _run(chain, stream, False)
        counts = []
        for k in range(80):
            s = hip_rt.stream_create()
            streams.append(s)
            for rep in range(2):
                chain.poison(stream=s)
                chain.dispatch(stream=s)
                counts.append(P.last_launch_count())
                hip_rt.stream_sync(s)
                assert np.array_equal(chain.download(), want), f"stream {k} dispatch {rep} differs"
        assert set(counts) <= {1, 2}
        assert counts[-1] == 2, "a stream beyond the counter blocks fused"
        assert counts.count(1) <= 2 * 64
    finally:
        for s in streams:
            P.stream_release(s)
            hip_rt.stream_destroy(s)
        chain.free()


def test_released_counter_blocks_are_reused_srgba8(stream):
    """nvproPyramidStreamRelease: 100 streams one after another, each released
    before it is destroyed, all fuse (the released blocks go to the next
    streams), and every result is the separate-launch bytes."""
    chain = P.DeviceMipChain(4096, 4096, P.FORMAT_SRGBA8)
    try:
        chain.synthesize(10, stream)
        want, _ = _run(chain, stream, False)
        for k in range(100):
            s = hip_rt.stream_create()
            try:
                chain.poison(stream=s)
                chain.dispatch(stream=s)
                assert P.last_launch_count() == 1, f"stream {k} did not fuse"
                hip_rt.stream_sync(s)
                assert np.array_equal(chain.download(), want), f"stream {k} differs"
            finally:
                P.stream_release(s)
                hip_rt.stream_destroy(s)
        s = hip_rt.stream_create()  # a stream that never dispatched: nothing to release
        P.stream_release(s)
        hip_rt.stream_destroy(s)
    finally:
        chain.free()
