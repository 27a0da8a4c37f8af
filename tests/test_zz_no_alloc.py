"""nvproPyramidBatchDispatchNoAlloc: the batch entry point that allocates nothing.

On the GPU it must write exactly the bytes the allocating batch entry point
(nvproPyramidBatchCreate + nvproPyramidBatchLaunch, same rounds-mode
configuration) writes, on two calls in a row, the second with a smaller batch
(the scratch then still holds the first call's longer work list).  On the host,
without a GPU: a work list over the scratch size and a queue-mode config are
rejected before any HIP call.
"""
import numpy as np
import pytest

from vk_compute_mipmaps_amd import hip_rt
from vk_compute_mipmaps_amd import pyramid as P
from vk_compute_mipmaps_amd import synth

FIRST = [(1000, 1000), (4095, 2047), (64, 64), (1, 300), (257, 3), (2048, 2048), (640, 480), (333, 200),
         (96, 1024), (1920, 1080), (127, 129), (700, 513), (2, 2), (4096, 4096), (1023, 4), (900, 16)]
SECOND = [(300, 200), (128, 128), (17, 1), (511, 255)]


def _run(sizes, seed0, stream, no_alloc, cfg, fmt=P.FORMAT_SRGBA8, arith=P.ARITH_FP32):
    chains = []
    try:
        for k, (w, h) in enumerate(sizes):
            c = P.DeviceMipChain(w, h, fmt)
            c.synthesize(seed0 + k, stream)
            c.poison(stream=stream)
            chains.append(c)
        pipes = P.pipelines(fmt, True, arith)
        images = [c.image for c in chains]
        if no_alloc:
            P.batch_dispatch_no_alloc(stream, pipes, images, config=cfg)
        else:
            batch = P.PyramidBatch(pipes, images, config=cfg)
            try:
                batch.launch(stream)
                hip_rt.stream_sync(stream)
                assert batch.read_errors() == 0
            finally:
                batch.destroy()
        hip_rt.stream_sync(stream)
        return [c.download() for c in chains]
    finally:
        for c in chains:
            c.free()


@pytest.mark.gpu
@pytest.mark.parametrize("fmt,arith", [(P.FORMAT_SRGBA8, P.ARITH_FP32), (P.FORMAT_RGBA16F, P.ARITH_PACKED_FP16)],
                         ids=["srgba8", "rgba16f-packed"])
def test_no_alloc_equals_batch_twice(fmt, arith, stream):
    cfg = P.batch_config(mode=P.BATCH_MODE_ROUNDS, use_graph=False)
    for sizes, seed0 in ((FIRST, 6100), (SECOND, 6200)):
        print(f"no-alloc batch: {len(sizes)} images, format {fmt}", flush=True)
        want = _run(sizes, seed0, stream, False, cfg, fmt, arith)
        got = _run(sizes, seed0, stream, True, cfg, fmt, arith)
        for (w, h), a, b in zip(sizes, got, want):
            assert np.array_equal(a, b), f"{w}x{h}: no-alloc batch differs from the allocating batch"


@pytest.mark.gpu
def test_no_alloc_default_config_mixed_workload(stream):
    """The default configuration (rounds mode, 64 KiB items, 128-texel tails) on
    the first 64 textures of the mixed-size workload."""
    sizes = synth.mixed_sizes(512, 8)[:64]
    want = _run(sizes, 8000, stream, False, P.batch_config(mode=P.BATCH_MODE_ROUNDS, use_graph=False))
    got = _run(sizes, 8000, stream, True, None)
    for (w, h), a, b in zip(sizes, got, want):
        assert np.array_equal(a, b), f"{w}x{h}: no-alloc batch differs from the allocating batch"


def _err(fn, *args, **kw):
    with pytest.raises(P.NvproPyramidError) as e:
        fn(*args, **kw)
    return str(e.value)


def test_no_alloc_rejects_oversize_work_list_before_any_hip_call():
    """20,000 images (never dereferenced: the size check precedes every HIP call)."""
    images = [P.make_image(0x10000, 64, 64) for _ in range(20000)]
    msg = _err(P.batch_dispatch_no_alloc, 0, P.pipelines(), images)
    assert "exceeds" in msg and str(P.NO_ALLOC_SCRATCH_BYTES) in msg


def test_no_alloc_rejects_queue_mode():
    images = [P.make_image(0x10000, 64, 64)]
    msg = _err(P.batch_dispatch_no_alloc, 0, P.pipelines(), images, config=P.batch_config(mode=P.BATCH_MODE_QUEUE))
    assert "ROUNDS" in msg
