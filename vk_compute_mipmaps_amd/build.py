"""Build the HIP library (gfx950) in-tree.

    python -m vk_compute_mipmaps_amd.build

produces vk_compute_mipmaps_amd/lib/libnvpro_pyramid_hip.so from csrc/.  Every
HIP translation unit is compiled with -ffp-contract=off (no implicit FMA: each
FP32 operation rounds where the source writes it) and with packed-FP32
instruction selection disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libnvpro_pyramid_hip.so")
INCLUDE = os.path.join(REPO_DIR, "include")

DEMO_PATH = os.path.join(LIB_DIR, "nvpro_pyramid_demo")
# Measurement only: the speed-of-light pass the benchmark times beside the chain.
SOL_PATH = os.path.join(LIB_DIR, "libnvpro_pyramid_sol.so")

SOURCES = [
    "fast_pipeline.hip",
    "general_pipeline.hip",
    "tail_pipeline.hip",
    "batch_pipeline.hip",
    "synth.hip",
    "convert.hip",
    "nvpro_pyramid_dispatch.cpp",
    "nvpro_pyramid_batch.cpp",
    "nvpro_pyramid_noalloc.cpp",
    "nvpro_pyramid_server.cpp",
    "nvpro_pyramid_io.cpp",
]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    path = os.path.join(rocm, "bin", "hipcc")
    return path if os.path.exists(path) else "hipcc"


def build_library(verbose: bool = True) -> str:
    """Compile every source for gfx950 (in parallel) and link the shared library."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(LIB_DIR, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(LIB_DIR, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([hipcc(), *HIPCC_FLAGS, "-I", INCLUDE, "-I", CSRC, *lang, "-c",
                     os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        list(pool.map(run, cmds))
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    for obj in objs:
        os.remove(obj)
    build_demo(verbose)
    build_sol(verbose)
    return LIB_PATH


def build_sol(verbose: bool = True) -> str:
    """The benchmark's speed-of-light kernel (csrc/speed_of_light.hip), its own library."""
    tmp = SOL_PATH + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-shared", os.path.join(CSRC, "speed_of_light.hip"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, SOL_PATH)
    return SOL_PATH


def build_demo(verbose: bool = True) -> str:
    """The C++ host (csrc/pyramid_demo.cpp) linked against the library."""
    tmp = DEMO_PATH + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-O2", "-std=c++17", "-I", INCLUDE,
           os.path.join(CSRC, "pyramid_demo.cpp"), "-o", tmp, "-L", LIB_DIR, "-lnvpro_pyramid_hip",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, DEMO_PATH)
    return DEMO_PATH


if __name__ == "__main__":
    build_library(verbose="-q" not in sys.argv)
