"""Writes tests/fixtures/oracle_fixtures.json: the frozen bytes of the oracle.

For every seeded benchmark configuration it records the plan the dispatcher
makes, a hash of level 0 (synth.py) and hashes of the oracle's FP32 schedule
emulator and FP64 filter oracle outputs (sRGBA8), or of both RGBA16F emulators
(FP32 and packed binary16).  tests/test_fixtures.py recomputes them, so a change
to the oracle, the planner or the input generator shows up as a changed hash
instead of moving together with the kernels it checks.

    python tests/fixtures/make_fixtures.py      # rewrite after an intended change
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

from oracle import orc  # noqa: E402
from vk_compute_mipmaps_amd import pyramid as P  # noqa: E402
from vk_compute_mipmaps_amd import synth  # noqa: E402

SRGBA8 = [("1024x1024", 1024, 1024, 1), ("4096x4096", 4096, 4096, 2), ("4095x2047", 4095, 2047, 3),
          ("1920x1080", 1920, 1080, 4), ("2047x4097", 2047, 4097, 5)]
RGBA16F = [("1024x1024 RGBA16F", 1024, 1024, 7), ("1000x777 RGBA16F", 1000, 777, 7)]
MIXED = (512, 8)


def digest(a) -> str:
    return hashlib.blake2b(memoryview(np.ascontiguousarray(a)).cast("B"), digest_size=16).hexdigest()


def plan_text(w, h, fmt=P.FORMAT_SRGBA8):
    return " ".join(map(repr, P.plan(w, h, fmt=fmt)))


def srgba8_entry(w, h, seed):
    l0 = synth.srgba8(w, h, seed)
    offs, total = P.packed_layout(w, h)
    chain = np.zeros(total, np.uint8)
    chain[: l0.nbytes] = l0.reshape(-1)
    steps = orc.plan(w, h)
    return {"width": w, "height": h, "seed": seed, "plan": plan_text(w, h), "level0": digest(l0),
            "fp32": digest(orc.pyramid_srgba8(chain, w, h, steps, 0, "f32")),
            "fp64": digest(orc.pyramid_srgba8(chain, w, h, steps, 0, "f64"))}


def rgba16f_entry(w, h, seed):
    l0 = synth.rgba16f(w, h, seed)
    offs, total = P.packed_layout(w, h, 0, 8)
    chain = np.zeros(total // 2, np.uint16)
    chain[: l0.size] = l0.view(np.uint16).reshape(-1)
    steps = orc.plan(w, h)
    return {"width": w, "height": h, "seed": seed, "plan": plan_text(w, h, P.FORMAT_RGBA16F), "level0": digest(l0),
            "fp32": digest(orc.pyramid_rgba16f(chain, w, h, steps)),
            "packed_fp16": digest(orc.pyramid_rgba16f(chain, w, h, steps, packed=True))}


def mixed_entry(count, seed):
    sizes = synth.mixed_sizes(count, seed)
    return {"count": count, "seed": seed, "sizes": digest(np.array(sizes, np.uint32)),
            "plans": digest(np.frombuffer("|".join(plan_text(w, h) for w, h in sizes).encode(), np.uint8)),
            "texels": int(sum(w * h for w, h in sizes)),
            "odd": sum(1 for w, h in sizes if w % 2 and h % 2)}


def build():
    out = {name: srgba8_entry(w, h, s) for name, w, h, s in SRGBA8}
    out.update({name: rgba16f_entry(w, h, s) for name, w, h, s in RGBA16F})
    out["512 mixed"] = mixed_entry(*MIXED)
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "oracle_fixtures.json"), "w") as f:
        json.dump(build(), f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote", os.path.join(HERE, "oracle_fixtures.json"))
