"""GPU parity of batched chains (nvproPyramidBatch*) and of the device input generator.

A batch runs each image's own plan, so every level it writes must equal, byte
for byte, what nvproPyramidDispatch writes for that image alone (itself pinned
to the FP32 oracle by test_gpu_srgba8.py); a sample of images is also checked
against the oracle directly.  Outputs are poisoned before every launch, so a
texel a launch skipped shows up.  Queue-mode counters reset themselves at the
end of a launch, which the repeated launches here exercise.
"""
import numpy as np
import pytest

from oracle import orc
from tests import helpers as H
from vk_compute_mipmaps_amd import hip_rt
from vk_compute_mipmaps_amd import pyramid as P
from vk_compute_mipmaps_amd import synth

pytestmark = pytest.mark.gpu

# Small mixed sizes: pow2 (fast steps of every count), even NP2, odd, thin and
# 1-texel edges, so every segment kind and the tail appear.
MIXED = [(64, 64), (1, 1), (2, 1), (1, 300), (257, 3), (128, 96), (512, 512), (255, 255), (640, 480),
         (333, 200), (96, 1024), (1024, 8), (1000, 1000), (32, 32), (4, 4), (127, 129), (384, 640),
         (2, 2), (60, 2), (700, 513), (256, 1024), (17, 1), (1024, 1024), (900, 16)]


def _chains(sizes, fmt, seed0, stream):
    chains = []
    for k, (w, h) in enumerate(sizes):
        c = P.DeviceMipChain(w, h, fmt)
        c.synthesize(seed0 + k, stream)
        chains.append(c)
    return chains


def _poison(chains, stream):
    for c in chains:
        c.poison(stream=stream)


def _plain_outputs(chains, stream, arithmetic=P.ARITH_FP32, mips=None):
    _poison(chains, stream)
    for k, c in enumerate(chains):
        c.dispatch(stream=stream, arithmetic=arithmetic, mip_levels=(mips[k] if mips else 0))
    hip_rt.stream_sync(stream)
    return [c.download() for c in chains]


def _batch_outputs(batch, chains, stream, launches=2):
    outs = None
    for _ in range(launches):
        _poison(chains, stream)
        batch.launch(stream)
        hip_rt.stream_sync(stream)
        now = [c.download() for c in chains]
        if outs is not None:
            for a, b in zip(outs, now):
                assert np.array_equal(a, b), "a repeated launch wrote different bytes"
        outs = now
    assert batch.read_errors() == 0
    return outs


def _free(chains):
    for c in chains:
        c.free()


CONFIGS = [
    dict(mode=P.BATCH_MODE_QUEUE, item_bytes=65536, tail_texels=128, use_graph=True),
    dict(mode=P.BATCH_MODE_QUEUE, item_bytes=4096, tail_texels=0, use_graph=True),
    dict(mode=P.BATCH_MODE_QUEUE, item_bytes=262144, tail_texels=64, use_graph=False, workgroups_per_cu=1),
    dict(mode=P.BATCH_MODE_ROUNDS, item_bytes=65536, tail_texels=128, use_graph=True),
    dict(mode=P.BATCH_MODE_ROUNDS, item_bytes=4096, tail_texels=0, use_graph=False),
]


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "{mode}-{item_bytes}-{tail_texels}-{use_graph}".format(**c))
def test_batch_srgba8_equals_plain(cfg, stream):
    print(f"batch sRGBA8 {cfg}", flush=True)
    chains = _chains(MIXED, P.FORMAT_SRGBA8, 100, stream)
    try:
        want = _plain_outputs(chains, stream)
        batch = P.PyramidBatch(P.pipelines(), [c.image for c in chains], config=P.batch_config(**cfg))
        info = batch.info()
        assert info.imageCount == len(MIXED)
        got = _batch_outputs(batch, chains, stream)
        for (w, h), a, b in zip(MIXED, got, want):
            assert np.array_equal(a, b), f"{w}x{h}: batch differs from nvproPyramidDispatch"
        batch.destroy()
    finally:
        _free(chains)


def test_batch_srgba8_against_oracle(stream):
    """Independent of the plain path: sampled images against the FP32 oracle."""
    chains = _chains(MIXED, P.FORMAT_SRGBA8, 200, stream)
    try:
        batch = P.PyramidBatch(P.pipelines(), [c.image for c in chains])
        got = _batch_outputs(batch, chains, stream, launches=1)
        for k in (0, 3, 6, 7, 9, 12, 15, 19):
            w, h = MIXED[k]
            host = H.host_chain(w, h, synth.srgba8(w, h, 200 + k))
            emu = H.emulate_srgba8(host, w, h, P.plan(w, h))
            worst, where, _, _ = orc.compare(got[k], emu, w, h)
            assert worst == 0, f"{w}x{h}: {worst} LSB from the FP32 oracle at {where}"
        batch.destroy()
    finally:
        _free(chains)


def test_batch_partial_levels(stream):
    sizes = MIXED[:12]
    mips = [0, 1, 1, 3, 2, 4, 5, 3, 0, 2, 6, 1]
    chains = _chains(sizes, P.FORMAT_SRGBA8, 300, stream)
    try:
        want = _plain_outputs(chains, stream, mips=mips)
        batch = P.PyramidBatch(P.pipelines(), [c.image for c in chains], mip_levels=mips)
        got = _batch_outputs(batch, chains, stream)
        for (w, h), a, b in zip(sizes, got, want):
            assert np.array_equal(a, b), f"{w}x{h}: partial batch differs"
        batch.destroy()
    finally:
        _free(chains)


@pytest.mark.parametrize("arith", [P.ARITH_FP32, P.ARITH_PACKED_FP16])
@pytest.mark.parametrize("mode", [P.BATCH_MODE_QUEUE, P.BATCH_MODE_ROUNDS])
def test_batch_rgba16f_equals_plain(arith, mode, stream):
    sizes = MIXED[:16]
    chains = _chains(sizes, P.FORMAT_RGBA16F, 400, stream)
    try:
        want = _plain_outputs(chains, stream, arithmetic=arith)
        pipes = P.pipelines(P.FORMAT_RGBA16F, True, arith)
        batch = P.PyramidBatch(pipes, [c.image for c in chains], config=P.batch_config(mode=mode))
        got = _batch_outputs(batch, chains, stream)
        for (w, h), a, b in zip(sizes, got, want):
            assert np.array_equal(a, b), f"{w}x{h}: RGBA16F batch differs"
        batch.destroy()
    finally:
        _free(chains)


def _digest(a: np.ndarray) -> bytes:
    import hashlib

    return hashlib.blake2b(memoryview(a), digest_size=20).digest()


def _mixed_sample(sizes, per_class=12):
    """Indices of a sample of the mixed workload: both sides odd, both sides
    multiples of 4 (plans that start with fast steps), and the other even
    sizes; the smallest and largest of each class plus evenly spaced ones."""
    classes = {
        "odd": [k for k, (w, h) in enumerate(sizes) if w % 2 and h % 2],
        "fast": [k for k, (w, h) in enumerate(sizes) if w % 4 == 0 and h % 4 == 0],
    }
    classes["even"] = [k for k in range(len(sizes)) if k not in classes["odd"] and k not in classes["fast"]]
    out = []
    for name, ks in classes.items():
        ks = sorted(ks, key=lambda k: sizes[k][0] * sizes[k][1])
        if len(ks) <= per_class:
            out += ks
        else:
            out += [ks[round(i * (len(ks) - 1) / (per_class - 1))] for i in range(per_class)]
    return sorted(set(out)), {name: len(ks) for name, ks in classes.items()}


def test_batch_mixed_workload_512(stream):
    """The whole 512-texture workload (sizes of synth.mixed_sizes(512, 8): each side
    log-uniform in [64, 4096], 40 % odd) in one batch, queue and rounds modes:
    every image byte-equal to its own nvproPyramidDispatch, and a sample of 36
    (odd, fast-step and other even sizes) equal to the FP32 oracle."""
    sizes = synth.mixed_sizes(512, 8)
    assert len(sizes) == 512
    chains = _chains(sizes, P.FORMAT_SRGBA8, 8000, stream)
    sample, counts = _mixed_sample(sizes)
    print(f"512 textures, classes {counts}, oracle sample {len(sample)}", flush=True)
    assert len(sample) >= 32 and min(counts.values()) > 0
    try:
        _poison(chains, stream)
        for c in chains:
            c.dispatch(stream=stream)
        hip_rt.stream_sync(stream)
        want, kept = [], {}
        for k, c in enumerate(chains):
            a = c.download()
            want.append(_digest(a))
            if k in sample:
                kept[k] = a
        for cfg in (P.batch_config(), P.batch_config(mode=P.BATCH_MODE_ROUNDS)):
            batch = P.PyramidBatch(P.pipelines(), [c.image for c in chains], config=cfg)
            try:
                info = batch.info()
                assert info.imageCount == 512
                _poison(chains, stream)
                batch.launch(stream)
                hip_rt.stream_sync(stream)
                assert batch.read_errors() == 0
            finally:
                batch.destroy()
            for k, c in enumerate(chains):
                a = c.download()
                assert _digest(a) == want[k], f"{sizes[k]}: batch differs from nvproPyramidDispatch ({cfg.as_dict()})"
                if k in kept:
                    assert np.array_equal(a, kept[k])
        for k in sample:
            w, h = sizes[k]
            gpu = kept[k]
            host = H.host_chain(w, h, gpu[: w * h * 4])  # level 0 as the device generated it
            emu = H.emulate_srgba8(host, w, h, P.plan(w, h))
            worst, where, ndiff, _ = orc.compare(gpu, emu, w, h)
            assert worst == 0, f"{w}x{h}: {worst} LSB from the FP32 oracle at {where} ({ndiff} texels)"
    finally:
        _free(chains)


def test_batch_autotune(stream):
    sizes = MIXED
    chains = _chains(sizes, P.FORMAT_SRGBA8, 500, stream)
    try:
        want = _plain_outputs(chains, stream)
        best, us = P.batch_autotune(P.pipelines(), [c.image for c in chains], stream=stream, repetitions=3)
        assert us > 0.0
        again, us2 = P.batch_autotune(P.pipelines(), [c.image for c in chains], stream=stream)
        assert again.as_dict() == best.as_dict() and us2 == us  # cached
        batch = P.PyramidBatch(P.pipelines(), [c.image for c in chains], config=best)
        got = _batch_outputs(batch, chains, stream)
        for (w, h), a, b in zip(sizes, got, want):
            assert np.array_equal(a, b), f"{w}x{h}: tuned batch differs"
        batch.destroy()
    finally:
        _free(chains)


# Widths below, at and above the generator's 256-thread stride, and 1-wide.
@pytest.mark.parametrize("w,h", [(1, 1), (67, 3), (256, 256), (1000, 333), (1, 700), (5, 1000), (4097, 3)])
def test_synthesize_srgba8_matches_spec(w, h, stream):
    c = P.DeviceMipChain(w, h, P.FORMAT_SRGBA8)
    try:
        c.synthesize(1234, stream)
        hip_rt.stream_sync(stream)
        got = c.download()[: w * h * 4].reshape(h, w, 4)
        assert np.array_equal(got, synth.srgba8(w, h, 1234))
    finally:
        c.free()


@pytest.mark.parametrize("w,h", [(1, 1), (67, 3), (300, 200), (1, 300), (257, 20)])
def test_synthesize_rgba16f_matches_spec(w, h, stream):
    c = P.DeviceMipChain(w, h, P.FORMAT_RGBA16F)
    try:
        c.synthesize(77, stream)
        hip_rt.stream_sync(stream)
        got = c.download()[: w * h * 8].view(np.uint16).reshape(h, w, 4)
        want = synth.rgba16f(w, h, 77).view(np.uint16)
        assert np.array_equal(got, want)
    finally:
        c.free()


def test_batch_spin_limit_reports_error(stream):
    """The error path of queue mode, forced from the host: with a spin limit of
    one poll, a step's items give up waiting for the previous step of their
    image (which is still running) and the launch reports
    NVPRO_PYRAMID_BATCH_ERROR_SPIN; every wait is bounded, so the launch still
    ends.  A batch of the same images with the default limit reports nothing
    and matches the plain dispatches."""
    sizes = [(2048, 2048), (2048, 1024), (1024, 2048), (1500, 1500)]
    chains = _chains(sizes, P.FORMAT_SRGBA8, 300, stream)
    try:
        want = _plain_outputs(chains, stream)
        images = [c.image for c in chains]
        bad = P.PyramidBatch(P.pipelines(), images,
                             config=P.batch_config(mode=P.BATCH_MODE_QUEUE, item_bytes=4096, use_graph=False,
                                                   spin_limit=1))
        try:
            bad.launch(stream)
            hip_rt.stream_sync(stream)
            assert bad.read_errors() & P.BATCH_ERROR_SPIN
        finally:
            bad.destroy()
        good = P.PyramidBatch(P.pipelines(), images, config=P.batch_config(mode=P.BATCH_MODE_QUEUE))
        try:
            got = _batch_outputs(good, chains, stream)
        finally:
            good.destroy()
        for (w, h), a, b in zip(sizes, got, want):
            assert np.array_equal(a, b), f"{w}x{h}: batch differs from nvproPyramidDispatch"
    finally:
        _free(chains)
