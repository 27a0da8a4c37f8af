#!/usr/bin/env python3
"""Benchmark: µs per 4096x4096 sRGBA8 full mip chain on one MI355X.

A step is one nvproPyramidDispatch of the full 13-level chain of a 4096x4096
sRGBA8 image (plan F(0,6) F(6,6)).  Chains rotate over >= 1 GiB of distinct
device buffers so the 256 MB Infinity Cache cannot hold the working set.
Timing uses HIP events on the dispatch stream around exactly --steps
dispatches, after --warmup untimed ones.

Printed (one JSON line): the metric, per-chain time, chains/s, the roofline of
the dominant kernel (the 6-level fast-pipeline launch over level 0), and the
reference's CPU generator (mipmap_storage.hpp:246-414, restated in C) timed on
the host as a baseline.  Beside it:
  batched   512 mixed-size sRGBA8 textures (seed 8): one batched launch against a
            loop of plain per-texture dispatches, timed by the C++ host
            (lib/nvpro_pyramid_demo batch), chains/s and p50/p99 per batch;
  serving   a Poisson stream of 20,000 requests at 20,000/s (sizes of the
            mixed-size workload, 70% sRGBA8 / 30% RGBA16F) through the serving
            loop with an 80 GB budget (lib/nvpro_pyramid_demo serve): chains/s,
            p50/p99 latency;
  configs   single chains of the other configurations (1024^2; the odd sizes
            4095x2047, 1920x1080, 2047x4097 and 4095^2; 16384^2 sRGBA8; 8192^2
            RGBA16F with FP32 and packed-FP16 arithmetic): us and GB/s.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "µs per 4096² sRGBA8 chain + %HBM roofline; batched mip-chains/s at p99 latency"
HBM_PEAK_GBS = 8000.0
W = H = 4096
SEED = 2


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--rotate-bytes", type=float, default=1.1e9)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--only-first-step", action="store_true",
                    help="dispatch only the 6-level fast step over level 0 (for profiling)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the 512-texture batched workload, in this process, --steps launches (for profiling)")
    ap.add_argument("--batch-mode", choices=["rounds", "queue"], default="rounds")
    ap.add_argument("--batch-no-graph", action="store_true", help="launch the batch's kernels directly")
    return ap.parse_args()


def batch_only(args):
    """The batched workload (512 textures, synth.mixed_sizes(512, 8), level 0
    seeded 8000 + k as in the C++ host) as one nvproPyramidBatch in this
    process: ms per launch over --steps launches, one JSON line."""
    from vk_compute_mipmaps_amd import hip_rt, pyramid as P, synth

    hip_rt.check(hip_rt.rt().hipSetDevice(0), "hipSetDevice")
    stream = hip_rt.stream_create()
    sizes = synth.mixed_sizes(512, 8)
    nbytes = [P.chain_bytes(w, h) for w, h in sizes]
    padded = [(b + 255) // 256 * 256 for b in nbytes]
    pool = hip_rt.malloc(sum(padded))
    images, off = [], 0
    for k, (w, h) in enumerate(sizes):
        images.append(P.make_image(pool + off, w, h))
        off += padded[k]
        P.synthesize(stream, P.FORMAT_SRGBA8, images[-1], 8 * 1000 + k)
    mode = P.BATCH_MODE_ROUNDS if args.batch_mode == "rounds" else P.BATCH_MODE_QUEUE
    batch = P.PyramidBatch(P.pipelines(), images,
                           config=P.batch_config(mode=mode, use_graph=not args.batch_no_graph))
    for _ in range(max(1, args.warmup)):
        batch.launch(stream)
    hip_rt.stream_sync(stream)
    e0, e1 = hip_rt.event_create(), hip_rt.event_create()
    hip_rt.event_record(e0, stream)
    for _ in range(args.steps):
        batch.launch(stream)
    hip_rt.event_record(e1, stream)
    hip_rt.event_sync(e1)
    ms = hip_rt.event_elapsed_ms(e0, e1) / args.steps
    info = batch.info()
    errors = batch.read_errors()
    batch.destroy()
    hip_rt.free(pool)
    print(json.dumps({"workload": "512 mixed sRGBA8 textures, synth.mixed_sizes(512, 8)", "mode": args.batch_mode,
                      "graph": not args.batch_no_graph, "launches": args.steps, "ms_per_batch": round(ms, 4),
                      "chains_per_s": round(512 / (ms * 1e-3), 1), "algorithmic_bytes": sum(nbytes),
                      "GBps": round(sum(nbytes) / (ms * 1e-3) / 1e9, 1), "kernel_launches_per_batch": info.launchCount,
                      "rounds": info.roundCount, "items": info.itemCount, "workgroups": info.workgroups,
                      "error_flags": errors}))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path has no multi-GPU leg",
              file=sys.stderr)
        sys.exit(2)
    if args.batch_only:
        batch_only(args)
        return
    import ctypes

    import numpy as np

    from vk_compute_mipmaps_amd import hip_rt, pyramid as P

    hip_rt.check(hip_rt.rt().hipSetDevice(0), "hipSetDevice")
    stream = hip_rt.stream_create()
    chain_bytes = P.chain_bytes(W, H)
    n_buf = max(2, int(np.ceil(args.rotate_bytes / chain_bytes)))
    chains = [P.DeviceMipChain(W, H) for _ in range(n_buf)]
    chains[0].synthesize(SEED, stream)  # synth.srgba8(W, H, SEED), generated on the device
    hip_rt.stream_sync(stream)
    level0 = chains[0].download()[: W * H * 4]
    for c in chains[1:]:
        hip_rt.memcpy_dtod_async(c.ptr, chains[0].ptr, W * H * 4, stream)
    for c in chains:
        c.poison(stream=stream)
    hip_rt.stream_sync(stream)

    pipes = P.pipelines()
    levels = P.full_level_count(W, H)
    first = P.plan(W, H)[0]
    step_levels = first.inputLevel + first.levelCount + 1  # mipLevels covering the first step only
    mip = step_levels if args.only_first_step else 0

    def dispatch(i, mip_levels=mip):
        P.nvpro_cmd_pyramid_dispatch(stream, pipes, chains[i % n_buf].image, mip_levels)

    for i in range(args.warmup):
        dispatch(i)
    hip_rt.stream_sync(stream)

    # Timed region: exactly --steps dispatches.
    e0, e1 = hip_rt.event_create(), hip_rt.event_create()
    hip_rt.device_sync()
    hip_rt.event_record(e0, stream)
    for i in range(args.steps):
        dispatch(i)
    hip_rt.event_record(e1, stream)
    hip_rt.event_sync(e1)
    hip_rt.device_sync()
    total_ms = hip_rt.event_elapsed_ms(e0, e1)
    us_per_chain = total_ms * 1000.0 / args.steps

    # Per-chain distribution (separate pass, an event pair per chain).
    reps = max(50, min(args.steps, 400))
    evs = [hip_rt.event_create() for _ in range(reps + 1)]
    hip_rt.event_record(evs[0], stream)
    for i in range(reps):
        dispatch(i)
        hip_rt.event_record(evs[i + 1], stream)
    hip_rt.event_sync(evs[-1])
    per = sorted(hip_rt.event_elapsed_ms(evs[i], evs[i + 1]) * 1000.0 for i in range(reps))
    p99 = per[min(len(per) - 1, int(round(0.99 * (len(per) - 1))))]

    # Dominant kernel: the first step alone (6-level fast pipeline over level 0).
    k_reps = max(50, min(args.steps, 400))
    hip_rt.event_record(evs[0], stream)
    for i in range(k_reps):
        dispatch(i, step_levels)
    hip_rt.event_record(evs[1], stream)
    hip_rt.event_sync(evs[1])
    k_us = hip_rt.event_elapsed_ms(evs[0], evs[1]) * 1000.0 / k_reps
    k_bytes = P.packed_layout(W, H, step_levels)[1]
    k_gbs = k_bytes / (k_us * 1e-6) / 1e9

    # Achievable copy bandwidth for reference (device-to-device, 1 GiB).
    cp_bytes = 1 << 30
    src, dst = hip_rt.malloc(cp_bytes), hip_rt.malloc(cp_bytes)
    hip_rt.memcpy_dtod_async(dst, src, cp_bytes, stream)
    hip_rt.stream_sync(stream)
    hip_rt.event_record(evs[0], stream)
    for _ in range(5):
        hip_rt.memcpy_dtod_async(dst, src, cp_bytes, stream)
    hip_rt.event_record(evs[1], stream)
    hip_rt.event_sync(evs[1])
    copy_gbs = 2 * cp_bytes * 5 / (hip_rt.event_elapsed_ms(evs[0], evs[1]) * 1e-3) / 1e9
    hip_rt.free(src)
    hip_rt.free(dst)

    for e in evs + [e0, e1]:
        hip_rt.event_destroy(e)
    for c in chains:
        c.free()
    configs = None if args.no_configs else large_configs(stream)
    hip_rt.stream_destroy(stream)
    batched = None if args.no_batched else batched_leg()
    serving = None if args.no_serving else serving_leg()

    traffic = None  # HBM bytes per F(0,6) launch, from the rocprofv3 --pmc summary
    tpath = os.path.join(REPO, "profiles", "fast_pipeline_srgba8_4096.md")
    if os.path.exists(tpath):
        with open(tpath) as f:
            for line in f:
                if line.startswith("- traffic_bytes_per_launch:"):
                    traffic = int(line.split(":")[1].split()[0])

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(level0, args.cpu_seconds)

    chain_total_bytes = chain_bytes
    out = {
        "metric": METRIC,
        "value": round(us_per_chain, 3),
        "unit": "us",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 6),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": "4096x4096 sRGBA8 full chain, 13 levels, plan F(0,6) F(6,6) (one fused launch)"
            if not args.only_first_step else "4096x4096 sRGBA8, first fast step only F(0,6)",
            "input": f"seeded value noise + uniform noise, seed {SEED}",
            "rotation_buffers": n_buf,
            "rotation_bytes": n_buf * chain_bytes,
        },
        "chains_per_s": round(1e6 / us_per_chain, 1),
        "latency_us": {"min": round(per[0], 3), "median": round(statistics.median(per), 3),
                       "p99": round(p99, 3)},
        "chain_algorithmic_bytes": chain_total_bytes,
        "chain_GBps": round(chain_total_bytes / (us_per_chain * 1e-6) / 1e9, 1),
        "chain_roofline_frac": round(chain_total_bytes / (us_per_chain * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {
            "bound": "hbm",
            "kernel": "fastPipelineKernel<Srgba8, 6, false> (F(0,6) over level 0, launched alone)",
            "achieved": round(k_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(k_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel_us": round(k_us, 3),
            "algorithmic_bytes_per_launch": k_bytes,
        },
        "copy_GBps": round(copy_gbs, 1),
        "cpu_baseline": cpu,
        "batched": batched,
        "serving": serving,
        "configs": configs,
    }
    print(json.dumps(out))


def batched_leg():
    """512 mixed textures: batched launch vs a loop of plain dispatches (C++ host)."""
    import subprocess

    from vk_compute_mipmaps_amd import build

    cmd = [build.DEMO_PATH, "batch", "--count", "512", "--seed", "8", "--reps", "30", "--autotune"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
    out = json.loads(r.stdout)
    out["command"] = "vk_compute_mipmaps_amd/lib/nvpro_pyramid_demo " + " ".join(cmd[1:])
    return out


def serving_leg():
    """Poisson request stream through the serving loop (C++ host)."""
    import subprocess

    from vk_compute_mipmaps_amd import build

    cmd = [build.DEMO_PATH, "serve", "--rate", "20000", "--count", "20000", "--seed", "9"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
    out = json.loads(r.stdout)
    out["command"] = "vk_compute_mipmaps_amd/lib/nvpro_pyramid_demo " + " ".join(cmd[1:])
    return out


def large_configs(stream, reps=50, warmup=10):
    """Single chains of the other configurations: median us and algorithmic GB/s
    over `reps` back-to-back chains after `warmup` untimed ones (a short burst
    from idle runs the large chains up to 10 % slower)."""
    from vk_compute_mipmaps_amd import hip_rt, pyramid as P

    res = {}
    for name, w, h, fmt, arith, seed in (("1024x1024 sRGBA8", 1024, 1024, P.FORMAT_SRGBA8, P.ARITH_FP32, 1),
                                         ("4095x2047 sRGBA8", 4095, 2047, P.FORMAT_SRGBA8, P.ARITH_FP32, 3),
                                         ("1920x1080 sRGBA8", 1920, 1080, P.FORMAT_SRGBA8, P.ARITH_FP32, 4),
                                         ("2047x4097 sRGBA8", 2047, 4097, P.FORMAT_SRGBA8, P.ARITH_FP32, 5),
                                         ("4095x4095 sRGBA8", 4095, 4095, P.FORMAT_SRGBA8, P.ARITH_FP32, 11),
                                         ("16384x16384 sRGBA8", 16384, 16384, P.FORMAT_SRGBA8, P.ARITH_FP32, 6),
                                         ("8192x8192 RGBA16F fp32", 8192, 8192, P.FORMAT_RGBA16F, P.ARITH_FP32, 7),
                                         ("8192x8192 RGBA16F packed_fp16", 8192, 8192, P.FORMAT_RGBA16F,
                                          P.ARITH_PACKED_FP16, 7)):
        c = P.DeviceMipChain(w, h, fmt)
        c.synthesize(seed, stream)
        for _ in range(warmup):
            c.dispatch(stream, arithmetic=arith)
        hip_rt.stream_sync(stream)
        evs = [hip_rt.event_create() for _ in range(reps + 1)]
        hip_rt.event_record(evs[0], stream)
        for i in range(reps):
            c.dispatch(stream, arithmetic=arith)
            hip_rt.event_record(evs[i + 1], stream)
        hip_rt.event_sync(evs[-1])
        per = sorted(hip_rt.event_elapsed_ms(evs[i], evs[i + 1]) * 1000.0 for i in range(reps))
        for e in evs:
            hip_rt.event_destroy(e)
        us = statistics.median(per)
        res[name] = {"us": round(us, 2), "GBps": round(c.nbytes / (us * 1e-6) / 1e9, 1),
                     "algorithmic_bytes": c.nbytes, "plan": " ".join(map(repr, P.plan(w, h, fmt=fmt)))}
        c.free()
    return res


def cpu_baseline(level0, seconds):
    """The reference CPU generator (restated in oracle/orc.c) on the same 4096^2
    input, one thread, repeated for about `seconds` of CPU work."""
    import numpy as np

    from oracle import orc
    from vk_compute_mipmaps_amd import pyramid as P

    offs, total = P.packed_layout(W, H)
    chain = np.zeros(total, np.uint8)
    chain[: level0.nbytes] = level0.reshape(-1)

    def timed(threads, budget, cap):
        times = []
        t_end = time.perf_counter() + budget
        while True:
            t0 = time.perf_counter()
            orc.reference_cpu(chain, W, H, threads=threads)
            times.append(time.perf_counter() - t0)
            if time.perf_counter() > t_end or len(times) >= cap:
                return times

    times = timed(1, seconds, 20)
    us = statistics.median(times) * 1e6
    # The same generator with its rows split over the host threads this process
    # may use (at most 16), reported beside the one-thread figure.
    try:
        nthreads = len(os.sched_getaffinity(0))
    except AttributeError:
        nthreads = os.cpu_count() or 1
    nthreads = max(1, min(nthreads, 16))
    mt = timed(nthreads, seconds / 4, 50)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round(us, 1), "unit": "us per 4096x4096 sRGBA8 chain", "cores": 1, "kind": "port",
            "sample": f"{len(times)} full 4096x4096 chains (mipmap_storage.hpp generateMipmaps restated in C, "
                      f"1 thread); median",
            "threads": {"value": round(statistics.median(mt) * 1e6, 1), "cores": nthreads,
                        "sample": f"{len(mt)} chains, rows of each level split over {nthreads} threads; median"},
            "host": {"cpu": model, "cpus_visible": os.cpu_count()}}


if __name__ == "__main__":
    main()
