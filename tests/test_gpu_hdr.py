"""The HDR configuration: an 8192x8192 RGBA16F chain (log-uniform radiance in
[1e-3, 1e3), seed 7), reduced with FP32 arithmetic and with packed FP16,
against the FP64 oracle of the same filter (oracle/orc.c orc_pyramid_linear_f64:
the 2x2 / 3x3 kernels of every level in double, no intermediate rounding).

Per level: the maximum and mean relative error of RGB (radiance is positive,
so the relative error is defined everywhere) and the maximum absolute error of
alpha.  Stated tolerances:
  FP32 arithmetic        max relative RGB error <= 1.5e-3 (3 x 2^-11) per
                         level: each stored level is rounded to binary16 once
                         (at most 2^-11 relative), and a level computed from a
                         stored one inherits that rounding; the plan F(0,6)
                         F(6,6) G(12,1) has three generations of stored inputs.
  packed FP16 arithmetic max relative RGB error <= 4e-3 (8 x 2^-11) per
                         level: every add of the 2x2 box also rounds to
                         binary16, and levels inside a fast step carry binary16
                         values from level to level.
Both modes: alpha within the same bound, absolute (alpha is in [0, 1)).
The table is printed (pytest -s shows it).
"""
import numpy as np
import pytest

from oracle import orc
from vk_compute_mipmaps_amd import hip_rt
from vk_compute_mipmaps_amd import pyramid as P

pytestmark = pytest.mark.gpu

W = H = 8192
SEED = 7
TOL = {P.ARITH_FP32: 1.5e-3, P.ARITH_PACKED_FP16: 4e-3}


@pytest.fixture(scope="module")
def reference(hip):
    """Level 0 (device generator) and the FP64 chain of the same image."""
    s = hip.stream_create()
    c = P.DeviceMipChain(W, H, P.FORMAT_RGBA16F)
    try:
        c.synthesize(SEED, s)
        hip.stream_sync(s)
        halves = c.download()[: W * H * 8].view(np.float16)
    finally:
        c.free()
        hip.stream_destroy(s)
    offs, total = P.packed_layout(W, H, 0, 8)
    chain = np.zeros(total // 2, np.float64)  # 4 doubles per texel, same texel offsets
    chain[: halves.size] = halves.astype(np.float64)
    chain = orc.pyramid_linear_f64(chain, W, H)
    return halves, chain, offs


@pytest.mark.parametrize("arith", [P.ARITH_FP32, P.ARITH_PACKED_FP16], ids=["fp32", "packed_fp16"])
def test_hdr_per_level_error_vs_fp64(arith, reference, stream):
    halves, ref, offs = reference
    c = P.DeviceMipChain(W, H, P.FORMAT_RGBA16F)
    try:
        c.upload_level0(halves)
        c.dispatch(stream=stream, arithmetic=arith)
        hip_rt.stream_sync(stream)
        gpu = c.download().view(np.float16).astype(np.float64)
    finally:
        c.free()
    name = "fp32" if arith == P.ARITH_FP32 else "packed_fp16"
    print(f"\nRGBA16F {W}x{H} {name}: per-level error vs FP64 (plan "
          f"{' '.join(map(repr, P.plan(W, H, fmt=P.FORMAT_RGBA16F)))})")
    print("level  size        max_rel_rgb   mean_rel_rgb  max_abs_alpha")
    worst = 0.0
    for level in range(1, len(offs)):
        lw, lh = P.level_dims(W, H, level)
        a, b = offs[level] // 2, offs[level] // 2 + lw * lh * 4  # offsets in halves = 4 per texel
        g = gpu[a:b].reshape(-1, 4)
        r = ref[a:b].reshape(-1, 4)
        rel = np.abs(g[:, :3] - r[:, :3]) / r[:, :3]
        amax = float(np.abs(g[:, 3] - r[:, 3]).max())
        print(f"{level:5d}  {lw:5d}x{lh:<5d}  {rel.max():.3e}     {rel.mean():.3e}     {amax:.3e}")
        worst = max(worst, float(rel.max()))
        assert rel.max() <= TOL[arith], f"level {level}: relative RGB error {rel.max():.3e}"
        assert amax <= TOL[arith], f"level {level}: alpha error {amax:.3e}"
    assert worst > 0.0  # binary16 storage cannot be exact everywhere
