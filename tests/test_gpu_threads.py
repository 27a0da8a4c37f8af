"""Dispatches from several host threads at once.

Each thread owns a stream and its own images and dispatches chains on them over
and over while the other threads do the same; ctypes releases the GIL during
the foreign calls, so the library really runs concurrently.  The entry points
keep no per-call global state besides each stream's fused-launch counter block
(taken under a lock, handed back by nvproPyramidStreamRelease) and the calling
thread's error message and launch count, so every thread's bytes must equal a
dispatch made alone.  The threads come in waves; each wave's streams are
released and destroyed before the next wave creates its own, so later waves
run on counter blocks that earlier streams used.  Each thread also launches a
small batch of its own (rounds mode, no graph) between its dispatches.
"""
import threading

import numpy as np
import pytest

from vk_compute_mipmaps_amd import hip_rt
from vk_compute_mipmaps_amd import pyramid as P

pytestmark = pytest.mark.gpu

# Plans: fused F(0,6)+F(6,6); F(0,6) then fused F(6,4); general steps and a
# tail; fast steps of RGBA16F with a tail; one group with a fused tail.
SIZES = [(4096, 4096, P.FORMAT_SRGBA8), (1024, 1024, P.FORMAT_SRGBA8), (4095, 2047, P.FORMAT_SRGBA8),
         (1920, 1080, P.FORMAT_RGBA16F), (128, 128, P.FORMAT_SRGBA8)]
BATCH = [(64, 64), (257, 129), (1000, 750), (33, 1025)]
SEED = 41
THREADS = 6
WAVES = 3
REPEATS = 4


def _alone(stream):
    """Bytes of every size dispatched alone on the test's own stream."""
    out = {}
    for w, h, fmt in SIZES:
        c = P.DeviceMipChain(w, h, fmt)
        try:
            c.synthesize(SEED, stream)
            c.poison(stream=stream)
            c.dispatch(stream=stream)
            hip_rt.stream_sync(stream)
            out[(w, h, fmt)] = c.download()
        finally:
            c.free()
    for w, h in BATCH:
        c = P.DeviceMipChain(w, h)
        try:
            c.synthesize(SEED + w, stream)
            c.poison(stream=stream)
            c.dispatch(stream=stream)
            hip_rt.stream_sync(stream)
            out[("batch", w, h)] = c.download()
        finally:
            c.free()
    return out


def _worker(tid, start, want, errors):
    s = hip_rt.stream_create()
    chains, batch_chains, batch = [], [], None
    try:
        chains = [P.DeviceMipChain(w, h, fmt) for w, h, fmt in SIZES]
        for c in chains:
            c.synthesize(SEED, s)
        batch_chains = [P.DeviceMipChain(w, h) for w, h in BATCH]
        for c, (w, h) in zip(batch_chains, BATCH):
            c.synthesize(SEED + w, s)
        cfg = P.batch_config(mode=P.BATCH_MODE_ROUNDS, use_graph=False)
        batch = P.PyramidBatch(P.pipelines(), [c.image for c in batch_chains], None, cfg)
        hip_rt.stream_sync(s)
        start.wait()
        for k in range(REPEATS):
            for c in chains + batch_chains:
                c.poison(stream=s)
            for c in chains:
                c.dispatch(stream=s)
            batch.launch(s)
        hip_rt.stream_sync(s)
        for c, (w, h, fmt) in zip(chains, SIZES):
            if not np.array_equal(c.download(), want[(w, h, fmt)]):
                errors.append(f"thread {tid}: {w}x{h} format {fmt} differs from a dispatch alone")
        for c, (w, h) in zip(batch_chains, BATCH):
            if not np.array_equal(c.download(), want[("batch", w, h)]):
                errors.append(f"thread {tid}: batched {w}x{h} differs from a dispatch alone")
    except Exception as e:  # reported by the test
        errors.append(f"thread {tid}: {e!r}")
        start.abort()  # the others stop waiting for this thread
    finally:
        hip_rt.stream_sync(s)
        if batch is not None:
            batch.destroy()
        for c in chains + batch_chains:
            c.free()
        try:
            P.stream_release(s)
        except Exception as e:
            errors.append(f"thread {tid}: stream release: {e!r}")
        hip_rt.stream_destroy(s)


def test_threads_dispatch_concurrently_like_alone(stream):
    want = _alone(stream)
    for r in range(WAVES):
        errors = []
        start = threading.Barrier(THREADS, timeout=120)
        threads = [threading.Thread(target=_worker, args=(r * THREADS + t, start, want, errors))
                   for t in range(THREADS)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=300)
            assert not t.is_alive(), f"wave {r}: a thread did not finish"
        assert not errors, f"wave {r}: " + "; ".join(errors[:8])
