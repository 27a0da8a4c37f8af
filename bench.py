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
  serving   Poisson request streams at 20k to 200k requests/s (sizes of the
            mixed-size workload, 70% sRGBA8 / 30% RGBA16F, level 0 generated on
            the device per request) through the serving loop with an 80 GB
            budget (lib/nvpro_pyramid_demo serve), each after a quarter second
            of uncounted warm-up arrivals: chains/s, p50/p99 latency, the
            highest rate within a 2 ms p99 and the saturation rate, beside the
            plain loop's chains/s;
  configs   chains of the other configurations (1024^2; the odd sizes
            4095x2047, 1920x1080, 2047x4097 and 4095^2; 16384^2 sRGBA8; 8192^2
            RGBA16F with FP32 and packed-FP16 arithmetic), each over rotating
            copies of its image: us and GB/s per chain back to back, and a
            lone chain's latency between events.
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
PREROLL_CHAINS = 2000


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

    # Clock settling: ~50 ms of chains before anything else, so that a short
    # run (the default --warmup is small) does not time a GPU still ramping
    # its clocks up from idle.
    for i in range(PREROLL_CHAINS):
        dispatch(i)
    hip_rt.stream_sync(stream)
    for i in range(args.warmup):
        dispatch(i)
    hip_rt.stream_sync(stream)

    # Timed region: exactly --steps dispatches.
    e0, e1 = hip_rt.event_create(), hip_rt.event_create()
    hip_rt.device_sync()
    hip_rt.event_record(e0, stream)
    t_issue = time.perf_counter()
    for i in range(args.steps):
        dispatch(i)
    host_us_per_dispatch = (time.perf_counter() - t_issue) * 1e6 / args.steps
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

    # Speed of light: the same rotation through a pass that reads level 0 once
    # and writes every level once without filtering (csrc/speed_of_light.hip,
    # the reference's `baseline` pipeline's traffic).
    sol = speed_of_light(stream, chains, n_buf, k_reps, evs[0], evs[1])
    for c in chains:  # the pass wrote non-mip values: restore the poison
        c.poison(stream=stream)

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
    configs = None if args.no_configs else large_configs(stream, args.rotate_bytes)
    # GPU outputs of the parity configurations (compared with the oracle in the
    # CPU-baseline leg, outside every timed region).
    parity_inputs = None if args.no_cpu_baseline else parity_outputs(stream)
    hip_rt.stream_destroy(stream)
    batched = None if args.no_batched else batched_leg()
    loop_cps = None
    if batched:
        loop_cps = max(batched["loop_eager"]["chains_per_s"], batched["loop_graph"]["chains_per_s"])
    serving = None if args.no_serving else serving_leg(loop_cps)

    # HBM bytes per launch from the committed rocprofv3 --pmc summary: the
    # chain's fused launch, and F(0,6) alone.
    traffic = {}
    tpath = os.path.join(REPO, "profiles", "fast_pipeline_srgba8_4096.md")
    if os.path.exists(tpath):
        with open(tpath) as f:
            for line in f:
                for key in ("chain_traffic_bytes_per_launch", "first_step_traffic_bytes_per_launch"):
                    if line.startswith(f"- {key}:"):
                        traffic[key] = int(line.split(":")[1].split()[0])

    cpu = parity = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(level0, args.cpu_seconds)
        parity = parity_stats(parity_inputs)

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
        "host_us_per_dispatch": round(host_us_per_dispatch, 3),
        "latency_us": {"min": round(per[0], 3), "median": round(statistics.median(per), 3),
                       "p99": round(p99, 3)},
        "chain_algorithmic_bytes": chain_total_bytes,
        "chain_GBps": round(chain_total_bytes / (us_per_chain * 1e-6) / 1e9, 1),
        "chain_roofline_frac": round(chain_total_bytes / (us_per_chain * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {
            "bound": "hbm",
            "kernel": "fastPipelineKernel<Srgba8, 6, true> (the chain's one launch: F(0,6), then F(6,6) "
                      "in its last workgroup)",
            "achieved": round(chain_total_bytes / (us_per_chain * 1e-6) / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(chain_total_bytes / (us_per_chain * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": traffic.get("chain_traffic_bytes_per_launch"),
            "kernel_us": round(us_per_chain, 3),
            "algorithmic_bytes_per_launch": chain_total_bytes,
            "timing": "HIP events on the dispatch stream over the timed region; one launch per chain",
        },
        "first_step_alone": {
            "kernel": "fastPipelineKernel<Srgba8, 6, false> (F(0,6) over level 0, mipLevels 7)",
            "us": round(k_us, 3),
            "GBps": round(k_gbs, 1),
            "frac": round(k_gbs / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_launch": k_bytes,
            "traffic": traffic.get("first_step_traffic_bytes_per_launch"),
        },
        "speed_of_light": sol,
        "chain_vs_speed_of_light": round(us_per_chain / sol["us"], 3),
        "copy_GBps": round(copy_gbs, 1),
        "cpu_baseline": cpu,
        "parity": parity,
        "batched": batched,
        "serving": serving,
        "configs": configs,
    }
    print(json.dumps(out))


def speed_of_light(stream, chains, n_buf, reps, ev0, ev1):
    """us per pass of csrc/speed_of_light.hip over the rotating 4096^2 chains:
    level 0 read once, every level written once, no filtering (incorrect values
    by design, like extras/general_pipelines/baseline/general_pipeline_alternative.glsl)."""
    import ctypes

    from vk_compute_mipmaps_amd import build, hip_rt, pyramid as P

    sol = ctypes.CDLL(build.SOL_PATH)
    sol.nvproPyramidSpeedOfLight.restype = ctypes.c_int
    sol.nvproPyramidSpeedOfLight.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                             ctypes.c_uint32]

    def run(i):
        rc = sol.nvproPyramidSpeedOfLight(ctypes.c_void_p(stream), ctypes.c_void_p(chains[i % n_buf].ptr), W, H, 0)
        if rc != 0:
            raise RuntimeError(f"nvproPyramidSpeedOfLight failed ({rc})")

    for i in range(20):
        run(i)
    hip_rt.stream_sync(stream)
    hip_rt.event_record(ev0, stream)
    for i in range(reps):
        run(i)
    hip_rt.event_record(ev1, stream)
    hip_rt.event_sync(ev1)
    us = hip_rt.event_elapsed_ms(ev0, ev1) * 1000.0 / reps
    nbytes = P.chain_bytes(W, H)
    return {"kernel": "speedOfLightKernel (csrc/speed_of_light.hip)", "us": round(us, 3),
            "GBps": round(nbytes / (us * 1e-6) / 1e9, 1), "algorithmic_bytes": nbytes, "passes": reps}


PARITY_CONFIGS = (("1024x1024", 1024, 1024, 1), ("4096x4096", 4096, 4096, 2), ("4095x2047", 4095, 2047, 3),
                  ("1920x1080", 1920, 1080, 4), ("2047x4097", 2047, 4097, 5))


def parity_outputs(stream):
    """One chain of each sRGBA8 parity configuration and the 8192^2 RGBA16F chain
    in both arithmetic modes, downloaded (level 0 from the device generator)."""
    from vk_compute_mipmaps_amd import hip_rt, pyramid as P

    out = {}
    for name, w, h, seed in PARITY_CONFIGS:
        c = P.DeviceMipChain(w, h)
        c.synthesize(seed, stream)
        c.dispatch(stream)
        hip_rt.stream_sync(stream)
        out[name] = (w, h, c.download())
        c.free()
    for arith, name in ((P.ARITH_FP32, "fp32"), (P.ARITH_PACKED_FP16, "packed_fp16")):
        c = P.DeviceMipChain(8192, 8192, P.FORMAT_RGBA16F)
        c.synthesize(7, stream)
        c.dispatch(stream, arithmetic=arith)
        hip_rt.stream_sync(stream)
        out["8192x8192 RGBA16F " + name] = (8192, 8192, c.download())
        c.free()
    return out


def parity_stats(outputs):
    """The GPU chains against the oracle (CPU-baseline leg, outside every timed
    region): for sRGBA8, the worst LSB and the bit-exact texel and RGB-channel
    rates against the FP64 filter oracle, the worst LSB against the FP32
    oracle, and the worst delta against the reference's CPU generator (the
    `delta` of demo_app/mipmaps_app.cpp:825-833); for 8192^2 RGBA16F, the
    per-level maximum relative RGB error against the FP64 chain."""
    import numpy as np

    from oracle import orc
    from vk_compute_mipmaps_amd import pyramid as P

    res = {}
    for name, w, h, _ in PARITY_CONFIGS:
        _, _, gpu = outputs[name]
        host = np.zeros_like(gpu)
        host[: w * h * 4] = gpu[: w * h * 4]
        steps = orc.plan(w, h)
        f32 = orc.pyramid_srgba8(host, w, h, steps, 0, "f32")
        f64 = orc.pyramid_srgba8(host, w, h, steps, 0, "f64")
        ref = orc.reference_cpu(host, w, h, threads=8)
        d = np.abs(gpu[w * h * 4:].astype(np.int16) - f64[w * h * 4:].astype(np.int16)).reshape(-1, 4)
        res[name + " sRGBA8"] = {
            "fp32_oracle_worst_lsb": int(orc.compare(gpu, f32, w, h)[0]),
            "fp64_worst_lsb": int(d.max()),
            "fp64_texels_exact_pct": round(100.0 * float((d == 0).all(axis=1).mean()), 4),
            "fp64_rgb_channels_exact_pct": round(100.0 * float((d[:, :3] == 0).mean()), 4),
            "reference_cpu_delta": int(orc.compare(gpu, ref, w, h)[0]),
            "texels": int(d.shape[0]),
        }
    w = h = 8192
    offs, total = P.packed_layout(w, h, 0, 8)
    halves = outputs["8192x8192 RGBA16F fp32"][2][: w * h * 8].view(np.float16)
    chain = np.zeros(total // 2, np.float64)
    chain[: halves.size] = halves.astype(np.float64)
    f64 = orc.pyramid_linear_f64(chain, w, h)
    for name in ("fp32", "packed_fp16"):
        gpu = outputs["8192x8192 RGBA16F " + name][2].view(np.float16)
        per = []
        for level in range(1, len(offs)):
            lw, lh = P.level_dims(w, h, level)
            a, b = offs[level] // 2, offs[level] // 2 + lw * lh * 4
            g = gpu[a:b].astype(np.float64).reshape(-1, 4)
            r = f64[a:b].reshape(-1, 4)
            per.append(float(f"{(np.abs(g[:, :3] - r[:, :3]) / r[:, :3]).max():.3e}"))
        res["8192x8192 RGBA16F " + name] = {"max_rel_rgb_error_per_level": per}
    return res


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


SERVE_RATES = (20000, 40000, 80000, 120000, 140000, 160000, 200000)
SERVE_P99_BOUND_US = 2000.0
SERVE_WARMUP_S = 0.25  # arrivals served before the measured ones at each rate, not counted


def serving_leg(loop_chains_per_s=None):
    """Poisson request streams through the serving loop (C++ host), one run per
    arrival rate: SERVE_WARMUP_S seconds of arrivals that bring the GPU's clocks
    and the server to a steady state and are not counted, then half a second of
    measured arrivals: chains/s, p50/p99 latency and batch sizes; `capacity` is
    the highest rate whose p99 stays within SERVE_P99_BOUND_US, `saturation` the
    most chains/s any rate reached."""
    import subprocess

    from vk_compute_mipmaps_amd import build

    runs = []
    for rate in SERVE_RATES:
        count = max(10000, rate // 2)
        warm = int(rate * SERVE_WARMUP_S)
        cmd = [build.DEMO_PATH, "serve", "--rate", str(rate), "--count", str(count), "--warmup", str(warm),
               "--seed", "9"]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
        if r.returncode != 0:
            raise RuntimeError(f"{' '.join(cmd)} failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
        out = json.loads(r.stdout)
        out["command"] = "vk_compute_mipmaps_amd/lib/nvpro_pyramid_demo " + " ".join(cmd[1:])
        runs.append(out)
    ok = [o for o in runs if o["latency_us"]["p99"] <= SERVE_P99_BOUND_US and o["failed"] == 0]
    cap = max(ok, key=lambda o: o["rate_per_s"]) if ok else None
    best = max(runs, key=lambda o: o["chains_per_s"])
    summary = {
        "p99_bound_us": SERVE_P99_BOUND_US,
        "capacity": None if cap is None else {
            "rate_per_s": cap["rate_per_s"], "chains_per_s": cap["chains_per_s"], "latency_us": cap["latency_us"],
            "mean_batch_images": cap["mean_batch_images"]},
        "saturation": {"rate_per_s": best["rate_per_s"], "chains_per_s": best["chains_per_s"],
                       "latency_us": best["latency_us"], "mean_batch_images": best["mean_batch_images"]},
        "plain_loop_chains_per_s": loop_chains_per_s,
        "runs": [{k: o[k] for k in ("rate_per_s", "requests", "warmup_requests", "chains_per_s", "latency_us",
                                    "latency_us_second_half", "mean_batch_images",
                                    "batches", "peak_resident_bytes", "failed", "host_loop_us", "command")}
                 for o in runs],
        "mix": runs[0]["mix"], "budget_bytes": runs[0]["budget_bytes"], "in_flight": runs[0]["in_flight"],
        "hold_ms": runs[0]["hold_ms"], "batch_mode": runs[0]["batch_mode"],
    }
    if loop_chains_per_s:
        if cap is not None:
            summary["capacity"]["vs_plain_loop"] = round(cap["chains_per_s"] / loop_chains_per_s, 2)
        summary["saturation"]["vs_plain_loop"] = round(best["chains_per_s"] / loop_chains_per_s, 2)
    return summary


def large_configs(stream, rotate_bytes, bursts=10, burst=10, singles=20, warmup=10):
    """Single chains of the other configurations, each over enough copies of its
    image to exceed `rotate_bytes` (at most 64), so that level 0 is not in the
    caches when a chain starts.  `us`: the median over `bursts` runs of `burst`
    back-to-back chains between two events, per chain (the headline's method);
    `single_us`: the median of `singles` chains with an event between every
    two, which adds the events' own cost (~2-3 us per chain) and shows a lone
    chain's latency.  Both after `warmup` untimed chains (a short burst from
    idle runs the large chains up to 10 % slower)."""
    import numpy as np

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
        first = P.DeviceMipChain(w, h, fmt)
        n_buf = int(max(1, min(64, np.ceil(rotate_bytes / first.nbytes))))
        chains = [first] + [P.DeviceMipChain(w, h, fmt) for _ in range(n_buf - 1)]
        for c in chains:
            c.synthesize(seed, stream)
        k = 0

        def dispatch():
            nonlocal k
            chains[k % n_buf].dispatch(stream, arithmetic=arith)
            k += 1

        for _ in range(warmup):
            dispatch()
        hip_rt.stream_sync(stream)
        evs = [hip_rt.event_create() for _ in range(max(bursts, singles) + 1)]
        hip_rt.event_record(evs[0], stream)
        for i in range(bursts):
            for _ in range(burst):
                dispatch()
            hip_rt.event_record(evs[i + 1], stream)
        hip_rt.event_sync(evs[bursts])
        per = sorted(hip_rt.event_elapsed_ms(evs[i], evs[i + 1]) * 1000.0 / burst for i in range(bursts))
        hip_rt.event_record(evs[0], stream)
        for i in range(singles):
            dispatch()
            hip_rt.event_record(evs[i + 1], stream)
        hip_rt.event_sync(evs[singles])
        single = sorted(hip_rt.event_elapsed_ms(evs[i], evs[i + 1]) * 1000.0 for i in range(singles))
        for e in evs:
            hip_rt.event_destroy(e)
        us = statistics.median(per)
        res[name] = {"us": round(us, 2), "GBps": round(first.nbytes / (us * 1e-6) / 1e9, 1),
                     "single_us": round(statistics.median(single), 2), "buffers": n_buf,
                     "algorithmic_bytes": first.nbytes, "plan": " ".join(map(repr, P.plan(w, h, fmt=fmt)))}
        for c in chains:
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
