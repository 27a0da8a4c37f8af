"""Host side of the C-ABI, without a GPU: exports, planner, argument checks.

Every check here returns before any HIP call is made (validation precedes
launch), so these run in a container without a device.
"""
import ctypes
import re

import pytest

from tests import helpers as H
from vk_compute_mipmaps_amd import pyramid as P

HEADER = P.PKG_DIR + "/../include/nvpro_pyramid_hip.h"


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:int|uint32_t|const char\*)\s+(nvpro\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = P.lib()
    names = declared_functions()
    assert len(names) >= 11
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/nvpro_pyramid_hip.h but not exported"
    assert set(names) == set(P.EXPORTED_SYMBOLS)


def test_full_level_count_and_bytes_per_texel():
    L = P.lib()
    assert [P.full_level_count(w, h) for w, h in [(1, 1), (2, 1), (4096, 4096), (4095, 1), (1, 65535)]] == [1, 2, 13, 12, 16]
    assert [L.nvproPyramidBytesPerTexel(f) for f in (0, 1, 2, 3)] == [4, 8, 16, 0]


def _err(fn, *args):
    with pytest.raises(P.NvproPyramidError) as e:
        fn(*args)
    return str(e.value)


def test_plan_argument_errors():
    assert "zero image size" in _err(P.plan, 0, 5)
    assert "zero image size" in _err(P.plan, 5, 0)
    assert "exceeds" in _err(P.plan, 16, 16, 6)
    assert "more than 16" in _err(P.plan, 65536, 4)
    assert "unknown format" in _err(P.plan, 8, 8, 0, True, 7)
    assert P.plan(1, 1) == []


def _dispatch(img, pipes=None, mip=0):
    P.nvpro_cmd_pyramid_dispatch(0, pipes or P.pipelines(), img, mip)


def test_dispatch_argument_errors_before_any_launch():
    good = 0x10000  # never dereferenced: every case fails validation first
    assert "null image data" in _err(_dispatch, P.make_image(0, 64, 64))
    assert "4-byte aligned" in _err(_dispatch, P.make_image(good + 2, 64, 64))
    assert "zero image size" in _err(_dispatch, P.make_image(good, 0, 64))
    assert "levelCount" in _err(_dispatch, P.make_image(good, 64, 64, 9))
    assert "exceeds the image's" in _err(_dispatch, P.make_image(good, 64, 64, 3), None, 5)
    assert "exceeds the 7" in _err(_dispatch, P.make_image(good, 64, 64), None, 9)
    pitches = [64 * 4 - 4] + [0] * 6
    assert "row pitch" in _err(_dispatch, P.make_image(good, 64, 64, 7, None, pitches))
    assert "offset" in _err(_dispatch, P.make_image(good, 64, 64, 7, [0, 6, 100, 200, 300, 400, 500]))
    bad = P.pipelines()
    bad.reserved = 1
    assert "reserved" in _err(_dispatch, P.make_image(good, 64, 64), bad)
    arith = P.pipelines(P.FORMAT_SRGBA8, True, P.ARITH_PACKED_FP16)
    assert "arithmetic" in _err(_dispatch, P.make_image(good, 64, 64), arith)
    img = P.make_image(good, 64, 64)
    img.reserved = 3
    assert "reserved" in _err(_dispatch, img)


def test_one_level_dispatch_is_a_no_op():
    """mipLevels 1 fills nothing (the reference asserts, nvpro_pyramid_dispatch.hpp:169)."""
    _dispatch(P.make_image(0x10000, 64, 64), None, 1)
    _dispatch(P.make_image(0x10000, 1, 1))


def test_custom_dispatchers_shape_the_plan():
    # levels_1_6: divisibility 2 (strategy.md.html:432).
    f26 = H.fast_dispatcher_ex(2, 6)
    assert [s.as_tuple()[:3] for s in P.plan(4094, 4094, 0, True, P.FORMAT_SRGBA8, None, f26)][:2] == [
        (1, 0, 1), (0, 1, 2)]
    # levels_1_3: at most 3 levels per fast dispatch.
    f43 = H.fast_dispatcher_ex(4, 3)
    assert [s.as_tuple()[:3] for s in P.plan(1024, 1024, 0, True, P.FORMAT_SRGBA8, None, f43)] == [
        (1, 0, 3), (1, 3, 3), (1, 6, 3), (0, 9, 1)]
    # invalid template arguments decline (odd divisibility, > 6 levels).
    assert [s.pipeline for s in P.plan(64, 64, 0, True, P.FORMAT_SRGBA8, None, H.fast_dispatcher_ex(3, 6))] == [0, 0, 0]
    assert [s.pipeline for s in P.plan(64, 64, 0, True, P.FORMAT_SRGBA8, None, H.fast_dispatcher_ex(4, 7))] == [0, 0, 0]


def test_dispatcher_contract_violations():
    zero = H.py_dispatcher(lambda cmd, st: 0)
    assert "general dispatcher filled no level" in _err(P.plan, 100, 60, 0, False, P.FORMAT_SRGBA8, zero, None)
    greedy = H.py_dispatcher(lambda cmd, st: st.contents.remainingLevels + 1)
    assert "claimed" in _err(P.plan, 100, 60, 0, False, P.FORMAT_SRGBA8, greedy, None)

    L = P.lib()

    def fast_too_deep(cmd, st):  # 3 levels on a level only divisible by 4
        s = st.contents
        L.nvproPyramidCmdDispatch(ctypes.c_void_p(cmd), P.PIPELINE_FAST, s.currentLevel << 5 | 3, 1)
        return 3

    assert "divisible by 8" in _err(P.plan, 12, 12, 0, True, P.FORMAT_SRGBA8, None, H.py_dispatcher(fast_too_deep))

    def general_three(cmd, st):
        s = st.contents
        L.nvproPyramidCmdDispatch(ctypes.c_void_p(cmd), P.PIPELINE_GENERAL, s.currentLevel << 5 | 3, 1)
        return 3

    assert "1 or 2 levels" in _err(P.plan, 64, 64, 0, False, P.FORMAT_SRGBA8, H.py_dispatcher(general_three), None)

    def fast_when_unavailable(cmd, st):
        s = st.contents
        L.nvproPyramidCmdDispatch(ctypes.c_void_p(cmd), P.PIPELINE_FAST, s.currentLevel << 5 | 1, 1)
        return 1

    assert "not available" in _err(P.plan, 64, 64, 0, False, P.FORMAT_SRGBA8,
                                   H.py_dispatcher(fast_when_unavailable), None)

    def past_the_end(cmd, st):
        s = st.contents
        L.nvproPyramidCmdDispatch(ctypes.c_void_p(cmd), P.PIPELINE_GENERAL, (s.currentLevel + 5) << 5 | 2, 1)
        return 2

    assert "outside" in _err(P.plan, 64, 64, 0, False, P.FORMAT_SRGBA8, H.py_dispatcher(past_the_end), None)


def test_null_dispatcher_records_nothing():
    """The demo's null dispatchers claim levels without dispatching
    (nvpro_pyramid_dispatch_alternative.cpp:66-77): a legal, empty plan."""
    null = H.py_dispatcher(lambda cmd, st: min(2, st.contents.remainingLevels))
    assert P.plan(256, 256, 0, False, P.FORMAT_SRGBA8, null, None) == []


def test_plan_reports_reference_workgroups_and_push_constants():
    steps = P.plan(4096, 4096)
    assert [(s.pushConstant, s.groupCountX, s.inputWidth) for s in steps] == [(6, 4096, 4096), ((6 << 5) | 6, 1, 64)]


def test_batch_argument_errors_before_any_device_call():
    """Batch creation validates config, pipelines and every image before touching the device."""
    fake = [P.make_image(0x10000, 64, 64), P.make_image(0x20000, 33, 7)]
    pipes = P.pipelines()
    assert "tailTexels" in _err(P.PyramidBatch, pipes, fake, None, P.batch_config(tail_texels=129))
    cfg = P.batch_config()
    cfg.reserved[1] = 5
    assert "reserved" in _err(P.PyramidBatch, pipes, fake, None, cfg)
    cfg = P.batch_config()
    cfg.mode = 7
    assert "unknown mode" in _err(P.PyramidBatch, pipes, fake, None, cfg)
    assert "no images" in _err(P.PyramidBatch, pipes, [], None, None)
    assert "unknown format" in _err(P.PyramidBatch, P.NvproPyramidPipelines(9, 1, 0, 0), fake, None, None)
    bad = [fake[0], P.make_image(0x30000, 0, 5)]
    assert "zero image size" in _err(P.PyramidBatch, pipes, bad, None, None)
    assert "exceeds" in _err(P.PyramidBatch, pipes, fake, [0, 7], None)
    assert "no images" in _err(P.batch_autotune, pipes, [])


def test_batch_default_config():
    c = P.batch_config()
    d = P.NvproPyramidBatchConfig()
    assert P.lib().nvproPyramidBatchDefaultConfig(ctypes.byref(d)) == 0
    assert d.as_dict() == {"mode": "queue", "workgroupsPerCU": 0, "itemBytes": 65536, "tailTexels": 128,
                           "useGraph": True}
    assert c.as_dict() == d.as_dict()


def test_synthesize_argument_errors():
    img = P.make_image(0x10000, 8, 8)
    assert "no generator" in _err(P.synthesize, 0, P.FORMAT_RGBA32F, img, 1)
    assert "zero image size" in _err(P.synthesize, 0, P.FORMAT_SRGBA8, P.make_image(0x10000, 0, 8), 1)
    assert "aligned" in _err(P.synthesize, 0, P.FORMAT_RGBA16F, P.make_image(0x10004, 8, 8), 1)


def test_demo_workload_sizes_match_python():
    """The C++ host's mixed-size workload equals synth.mixed_sizes (no device use)."""
    import subprocess

    from vk_compute_mipmaps_amd import build, synth

    out = subprocess.run([build.DEMO_PATH, "sizes", "512", "8"], check=True, capture_output=True, text=True).stdout
    got = [tuple(int(v) for v in line.split()) for line in out.splitlines()]
    assert got == synth.mixed_sizes(512, 8)
