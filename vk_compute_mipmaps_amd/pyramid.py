"""Host-side mirror of nvpro_pyramid_dispatch.hpp over the C-ABI library.

Names follow the reference interface (nvpro_pyramid/nvpro_pyramid_dispatch.hpp):

    NvproPyramidPipelines, NvproPyramidState       structs of :29-35 and :63-75
    nvpro_cmd_pyramid_dispatch(...)                nvproCmdPyramidDispatch, :54-59 / :294-304
    nvpro_cmd_pyramid_dispatch_custom(...)         the dispatcher overload, :109-188
    DEFAULT_FAST_DISPATCHER / DEFAULT_GENERAL_DISPATCHER   :195-242 / :247-292

The library is loaded from vk_compute_mipmaps_amd/lib/; importing this module
never falls back to anything else: a missing library raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libnvpro_pyramid_hip.so")

INPUT_LEVEL_SHIFT = 5
MAX_LEVELS = 16
PIPELINE_GENERAL = 0
PIPELINE_FAST = 1
FORMAT_SRGBA8 = 0
FORMAT_RGBA16F = 1
FORMAT_RGBA32F = 2
ARITH_FP32 = 0
ARITH_PACKED_FP16 = 1
BYTES_PER_TEXEL = {FORMAT_SRGBA8: 4, FORMAT_RGBA16F: 8, FORMAT_RGBA32F: 16}

# Every symbol include/nvpro_pyramid_hip.h declares.
EXPORTED_SYMBOLS = (
    "nvproPyramidCmdDispatch",
    "nvproPyramidDefaultFastDispatcher",
    "nvproPyramidFastDispatcherEx",
    "nvproPyramidDefaultGeneralDispatcher",
    "nvproPyramidDispatch",
    "nvproPyramidDispatchCustom",
    "nvproPyramidPlan",
    "nvproPyramidFullLevelCount",
    "nvproPyramidPackedLayout",
    "nvproPyramidBytesPerTexel",
    "nvproPyramidGetLastError",
    "nvproPyramidLastLaunchCount",
    "nvproPyramidStreamRelease",
    "nvproPyramidSynthesize",
    "nvproPyramidPremultiplyAlpha",
    "nvproPyramidWriteTga",
    "nvproPyramidReadTga",
    "nvproPyramidBatchDefaultConfig",
    "nvproPyramidBatchCreate",
    "nvproPyramidBatchLaunch",
    "nvproPyramidBatchGetInfo",
    "nvproPyramidBatchDispatchNoAlloc",
    "nvproPyramidBatchReadErrors",
    "nvproPyramidBatchDestroy",
    "nvproPyramidBatchAutotune",
    "nvproPyramidServerDefaultConfig",
    "nvproPyramidServerCreate",
    "nvproPyramidServerSubmit",
    "nvproPyramidServerPoll",
    "nvproPyramidServerRelease",
    "nvproPyramidServerGetStats",
    "nvproPyramidServerDestroy",
)

BATCH_MODE_ROUNDS = 0
BATCH_MODE_QUEUE = 1
BATCH_ERROR_SPIN = 1
BATCH_MAX_TAIL_TEXELS = 128


class NvproPyramidError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"nvpro_pyramid error {code}: {message}")
        self.code = code


class NvproPyramidPipelines(ctypes.Structure):
    _fields_ = [("format", ctypes.c_uint32), ("fastPipeline", ctypes.c_uint32),
                ("arithmetic", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class NvproPyramidImage(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("levelCount", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("levelOffsets", ctypes.POINTER(ctypes.c_uint64)),
                ("rowPitches", ctypes.POINTER(ctypes.c_uint32))]


class NvproPyramidState(ctypes.Structure):
    _fields_ = [("currentLevel", ctypes.c_uint32), ("remainingLevels", ctypes.c_uint32),
                ("currentX", ctypes.c_uint32), ("currentY", ctypes.c_uint32)]


class NvproPyramidStep(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "pipeline", "inputLevel", "levelCount", "pushConstant", "groupCountX",
        "inputWidth", "inputHeight", "reserved")]

    def as_tuple(self):
        return (self.pipeline, self.inputLevel, self.levelCount, self.groupCountX)

    def __repr__(self):
        kind = "F" if self.pipeline == PIPELINE_FAST else "G"
        return f"{kind}({self.inputLevel},{self.levelCount},{self.groupCountX})"


class NvproPyramidBatchConfig(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_uint32), ("workgroupsPerCU", ctypes.c_uint32),
                ("itemBytes", ctypes.c_uint32), ("tailTexels", ctypes.c_uint32),
                ("useGraph", ctypes.c_uint32), ("spinLimit", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 2)]

    def as_dict(self):
        return {"mode": "queue" if self.mode == BATCH_MODE_QUEUE else "rounds",
                "workgroupsPerCU": self.workgroupsPerCU, "itemBytes": self.itemBytes,
                "tailTexels": self.tailTexels, "useGraph": bool(self.useGraph)}


class NvproPyramidBatchInfo(ctypes.Structure):
    _fields_ = [("imageCount", ctypes.c_uint32), ("segmentCount", ctypes.c_uint32),
                ("itemCount", ctypes.c_uint32), ("roundCount", ctypes.c_uint32),
                ("launchCount", ctypes.c_uint32), ("workgroups", ctypes.c_uint32),
                ("algorithmicBytes", ctypes.c_uint64), ("deviceBytes", ctypes.c_uint64),
                ("config", NvproPyramidBatchConfig)]


class NvproPyramidServerConfig(ctypes.Structure):
    _fields_ = [("budgetBytes", ctypes.c_uint64), ("maxInFlight", ctypes.c_uint32),
                ("maxBatchImages", ctypes.c_uint32), ("arithmetic16f", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 2), ("batch", NvproPyramidBatchConfig)]


class NvproPyramidRequest(ctypes.Structure):
    _fields_ = [("format", ctypes.c_uint32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("mipLevels", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("level0", ctypes.c_void_p)]


class NvproPyramidCompletion(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("status", ctypes.c_int), ("format", ctypes.c_uint32),
                ("image", NvproPyramidImage), ("bytes", ctypes.c_uint64), ("latencyUs", ctypes.c_double),
                ("waitUs", ctypes.c_double)]


class NvproPyramidServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "submitted", "admitted", "completed", "released", "batches", "batchedImages", "failedBatches",
        "budgetBytes", "residentBytes", "peakResidentBytes")] + [
        ("waiting", ctypes.c_uint32), ("inFlightBatches", ctypes.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


DISPATCHER = ctypes.CFUNCTYPE(ctypes.c_uint32, ctypes.c_void_p, ctypes.POINTER(NvproPyramidState))

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"HIP library {LIB_PATH} is not built; run `python -m vk_compute_mipmaps_amd.build`")
        L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        u32, vp = ctypes.c_uint32, ctypes.c_void_p
        L.nvproPyramidFullLevelCount.restype = u32
        L.nvproPyramidFullLevelCount.argtypes = [u32, u32]
        L.nvproPyramidBytesPerTexel.restype = u32
        L.nvproPyramidBytesPerTexel.argtypes = [u32]
        L.nvproPyramidGetLastError.restype = ctypes.c_char_p
        L.nvproPyramidGetLastError.argtypes = []
        L.nvproPyramidLastLaunchCount.restype = ctypes.c_uint32
        L.nvproPyramidLastLaunchCount.argtypes = []
        L.nvproPyramidStreamRelease.restype = ctypes.c_int
        L.nvproPyramidStreamRelease.argtypes = [vp]
        L.nvproPyramidPackedLayout.restype = ctypes.c_int
        L.nvproPyramidPackedLayout.argtypes = [u32, u32, u32, u32, ctypes.POINTER(ctypes.c_uint64),
                                               ctypes.POINTER(ctypes.c_uint64)]
        L.nvproPyramidPlan.restype = ctypes.c_int
        L.nvproPyramidPlan.argtypes = [ctypes.POINTER(NvproPyramidPipelines), u32, u32, u32, vp, vp,
                                       ctypes.POINTER(NvproPyramidStep), u32, ctypes.POINTER(u32)]
        L.nvproPyramidDispatch.restype = ctypes.c_int
        L.nvproPyramidDispatch.argtypes = [vp, ctypes.POINTER(NvproPyramidPipelines),
                                           ctypes.POINTER(NvproPyramidImage), u32]
        L.nvproPyramidDispatchCustom.restype = ctypes.c_int
        L.nvproPyramidDispatchCustom.argtypes = [vp, ctypes.POINTER(NvproPyramidPipelines),
                                                 ctypes.POINTER(NvproPyramidImage), u32, vp, vp]
        L.nvproPyramidCmdDispatch.restype = ctypes.c_int
        L.nvproPyramidCmdDispatch.argtypes = [vp, u32, u32, u32]
        for name in ("nvproPyramidDefaultFastDispatcher", "nvproPyramidDefaultGeneralDispatcher"):
            getattr(L, name).restype = u32
            getattr(L, name).argtypes = [vp, ctypes.POINTER(NvproPyramidState)]
        L.nvproPyramidFastDispatcherEx.restype = u32
        L.nvproPyramidFastDispatcherEx.argtypes = [vp, ctypes.POINTER(NvproPyramidState), u32, u32]
        L.nvproPyramidSynthesize.restype = ctypes.c_int
        L.nvproPyramidSynthesize.argtypes = [vp, u32, ctypes.POINTER(NvproPyramidImage), ctypes.c_uint64]
        L.nvproPyramidPremultiplyAlpha.restype = ctypes.c_int
        L.nvproPyramidPremultiplyAlpha.argtypes = [vp, u32, ctypes.POINTER(NvproPyramidImage)]
        L.nvproPyramidWriteTga.restype = ctypes.c_int
        L.nvproPyramidWriteTga.argtypes = [vp, u32, u32, u32, ctypes.c_char_p]
        L.nvproPyramidReadTga.restype = ctypes.c_int
        L.nvproPyramidReadTga.argtypes = [ctypes.c_char_p, vp, ctypes.c_uint64, ctypes.POINTER(u32),
                                          ctypes.POINTER(u32)]
        cfgp = ctypes.POINTER(NvproPyramidBatchConfig)
        for name, args in (
                ("nvproPyramidBatchDefaultConfig", [cfgp]),
                ("nvproPyramidBatchCreate", [ctypes.POINTER(NvproPyramidPipelines), ctypes.POINTER(NvproPyramidImage),
                                             ctypes.POINTER(u32), u32, cfgp, ctypes.POINTER(vp)]),
                ("nvproPyramidBatchLaunch", [vp, vp]),
                ("nvproPyramidBatchGetInfo", [vp, ctypes.POINTER(NvproPyramidBatchInfo)]),
                ("nvproPyramidBatchReadErrors", [vp, ctypes.POINTER(u32)]),
                ("nvproPyramidBatchDestroy", [vp]),
                ("nvproPyramidBatchDispatchNoAlloc", [vp, ctypes.POINTER(NvproPyramidPipelines),
                                                      ctypes.POINTER(NvproPyramidImage), ctypes.POINTER(u32), u32,
                                                      cfgp]),
                ("nvproPyramidBatchAutotune", [ctypes.POINTER(NvproPyramidPipelines),
                                               ctypes.POINTER(NvproPyramidImage), ctypes.POINTER(u32), u32, vp, u32,
                                               cfgp, ctypes.POINTER(ctypes.c_float)])):
            getattr(L, name).restype = ctypes.c_int
            getattr(L, name).argtypes = args
        for name, args in (
                ("nvproPyramidServerDefaultConfig", [ctypes.POINTER(NvproPyramidServerConfig)]),
                ("nvproPyramidServerCreate", [ctypes.POINTER(NvproPyramidServerConfig), ctypes.POINTER(vp)]),
                ("nvproPyramidServerSubmit", [vp, ctypes.POINTER(NvproPyramidRequest),
                                              ctypes.POINTER(ctypes.c_uint64)]),
                ("nvproPyramidServerPoll", [vp, ctypes.POINTER(NvproPyramidCompletion), u32, ctypes.POINTER(u32)]),
                ("nvproPyramidServerRelease", [vp, ctypes.c_uint64]),
                ("nvproPyramidServerGetStats", [vp, ctypes.POINTER(NvproPyramidServerStats)]),
                ("nvproPyramidServerDestroy", [vp])):
            getattr(L, name).restype = ctypes.c_int
            getattr(L, name).argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().nvproPyramidGetLastError().decode()


def stream_release(stream: int):
    """nvproPyramidStreamRelease: hand the stream's fused-launch counter block back."""
    _check(lib().nvproPyramidStreamRelease(ctypes.c_void_p(stream)))


def last_launch_count() -> int:
    """Kernel launches of this thread's last successful dispatch (fused or not)."""
    return int(lib().nvproPyramidLastLaunchCount())


def _check(code: int):
    if code != 0:
        raise NvproPyramidError(code, last_error())


def dispatcher_address(fn) -> int:
    """Address of a library dispatcher or of a DISPATCHER-wrapped Python callable."""
    return ctypes.cast(fn, ctypes.c_void_p).value


def default_fast_dispatcher() -> int:
    return dispatcher_address(lib().nvproPyramidDefaultFastDispatcher)


def default_general_dispatcher() -> int:
    return dispatcher_address(lib().nvproPyramidDefaultGeneralDispatcher)


def full_level_count(width: int, height: int) -> int:
    return lib().nvproPyramidFullLevelCount(width, height)


def level_dims(width: int, height: int, level: int):
    for _ in range(level):
        width = (width >> 1) | (width == 1)
        height = (height >> 1) | (height == 1)
    return width, height


def packed_layout(width: int, height: int, levels: int = 0, bytes_per_texel: int = 4):
    """MipmapStorage layout: (list of per-level byte offsets, total bytes)."""
    levels = levels or full_level_count(width, height)
    offs = (ctypes.c_uint64 * levels)()
    total = ctypes.c_uint64()
    _check(lib().nvproPyramidPackedLayout(width, height, levels, bytes_per_texel, offs, ctypes.byref(total)))
    return [int(o) for o in offs], int(total.value)


def chain_bytes(width: int, height: int, levels: int = 0, bytes_per_texel: int = 4) -> int:
    """Algorithmic bytes of a chain: base read once plus every output level written once."""
    return packed_layout(width, height, levels, bytes_per_texel)[1]


def pipelines(fmt: int = FORMAT_SRGBA8, fast: bool = True, arithmetic: int = ARITH_FP32):
    return NvproPyramidPipelines(fmt, 1 if fast else 0, arithmetic, 0)


def plan(width: int, height: int, mip_levels: int = 0, fast: bool = True, fmt: int = FORMAT_SRGBA8,
         general_dispatcher=None, fast_dispatcher=None):
    pipes = pipelines(fmt, fast)
    steps = (NvproPyramidStep * 64)()
    n = ctypes.c_uint32()
    _check(lib().nvproPyramidPlan(ctypes.byref(pipes), width, height, mip_levels,
                                  _fn_ptr(general_dispatcher), _fn_ptr(fast_dispatcher), steps, 64,
                                  ctypes.byref(n)))
    return [steps[i] for i in range(min(n.value, 64))]


def _fn_ptr(fn):
    if fn is None:
        return None
    if isinstance(fn, int):
        return ctypes.c_void_p(fn)
    return ctypes.cast(fn, ctypes.c_void_p)


def make_image(data_ptr: int, width: int, height: int, level_count: int = 0, level_offsets=None,
               row_pitches=None):
    img = NvproPyramidImage()
    img.data = data_ptr
    img.width, img.height = width, height
    img.levelCount = level_count
    img.reserved = 0
    keep = []
    if level_offsets is not None:
        arr = (ctypes.c_uint64 * len(level_offsets))(*level_offsets)
        img.levelOffsets = arr
        keep.append(arr)
    if row_pitches is not None:
        arr = (ctypes.c_uint32 * len(row_pitches))(*row_pitches)
        img.rowPitches = arr
        keep.append(arr)
    img._keep = keep
    return img


def nvpro_cmd_pyramid_dispatch(stream: int, pipes: NvproPyramidPipelines, image: NvproPyramidImage,
                               mip_levels: int = 0):
    """nvproCmdPyramidDispatch over a HIP stream; raises NvproPyramidError."""
    _check(lib().nvproPyramidDispatch(ctypes.c_void_p(stream), ctypes.byref(pipes), ctypes.byref(image),
                                      mip_levels))


def nvpro_cmd_pyramid_dispatch_custom(stream: int, pipes: NvproPyramidPipelines, image: NvproPyramidImage,
                                      mip_levels: int, general_dispatcher, fast_dispatcher):
    _check(lib().nvproPyramidDispatchCustom(ctypes.c_void_p(stream), ctypes.byref(pipes), ctypes.byref(image),
                                            mip_levels, _fn_ptr(general_dispatcher), _fn_ptr(fast_dispatcher)))


def synthesize(stream: int, fmt: int, image: NvproPyramidImage, seed: int):
    """nvproPyramidSynthesize: seeded synthetic level 0 of `image` (synth.py's spec)."""
    _check(lib().nvproPyramidSynthesize(ctypes.c_void_p(stream), fmt, ctypes.byref(image), seed))


def premultiply_alpha(stream: int, fmt: int, image: NvproPyramidImage):
    """nvproPyramidPremultiplyAlpha: premultiply level 0 of `image` in place."""
    _check(lib().nvproPyramidPremultiplyAlpha(ctypes.c_void_p(stream), fmt, ctypes.byref(image)))


def write_tga(chain: np.ndarray, width: int, height: int, base_filename: str, levels: int = 0):
    """writeMipmapsTga: one TGA file per level of a packed sRGBA8 host chain."""
    a = np.ascontiguousarray(chain, dtype=np.uint8)
    _check(lib().nvproPyramidWriteTga(a.ctypes.data, width, height, levels, base_filename.encode()))


def read_tga(filename: str) -> np.ndarray:
    """A 24/32-bit true-colour TGA as a (height, width, 4) uint8 RGBA array."""
    w, h = ctypes.c_uint32(), ctypes.c_uint32()
    _check(lib().nvproPyramidReadTga(filename.encode(), None, 0, ctypes.byref(w), ctypes.byref(h)))
    out = np.empty((h.value, w.value, 4), np.uint8)
    _check(lib().nvproPyramidReadTga(filename.encode(), out.ctypes.data, out.nbytes, ctypes.byref(w),
                                     ctypes.byref(h)))
    return out


def tga_level_filename(base: str, level: int) -> str:
    """The file writeMipmapsTga names level `level` of `base`."""
    if level == 0:
        return base
    dot = base.rfind(".")
    return base[: dot + 1] + str(level) + base[dot:] if dot >= 0 else base + str(level)


POISON_SRGBA8 = 0xFFFF00FF  # magenta, opaque (bytes FF 00 FF FF), as the reference's test clears


class DeviceMipChain:
    """A packed MipmapStorage chain in device memory (one hipMalloc)."""

    def __init__(self, width: int, height: int, fmt: int = FORMAT_SRGBA8, levels: int = 0):
        from . import hip_rt

        self.rt = hip_rt
        self.width, self.height, self.fmt = width, height, fmt
        self.levels = levels or full_level_count(width, height)
        self.bpt = BYTES_PER_TEXEL[fmt]
        self.offsets, self.nbytes = packed_layout(width, height, self.levels, self.bpt)
        self.ptr = hip_rt.malloc(self.nbytes)
        self.image = make_image(self.ptr, width, height, self.levels)

    def upload_level0(self, level0: np.ndarray):
        a = np.ascontiguousarray(level0)
        assert a.nbytes == self.width * self.height * self.bpt
        self.rt.memcpy_htod(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes)

    def upload(self, chain: np.ndarray):
        a = np.ascontiguousarray(chain)
        assert a.nbytes == self.nbytes
        self.rt.memcpy_htod(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes)

    def premultiply(self, stream: int = 0):
        """Premultiply level 0's alpha in place (device)."""
        premultiply_alpha(stream, self.fmt, self.image)

    def synthesize(self, seed: int, stream: int = 0):
        """Write the seeded synthetic level 0 (synth.py's images) on the device."""
        synthesize(stream, self.fmt, self.image, seed)

    def poison(self, word: int = POISON_SRGBA8, stream: int = 0):
        """Fill levels >= 1 with a poison pattern (unwritten texels stay visible)."""
        start = self.offsets[1] if self.levels > 1 else self.nbytes
        if self.nbytes > start:
            self.rt.memset_d32_async(self.ptr + start, word, (self.nbytes - start) // 4, stream)

    def download(self) -> np.ndarray:
        out = np.empty(self.nbytes, dtype=np.uint8)
        self.rt.memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes)
        return out

    def dispatch(self, stream: int = 0, fast: bool = True, mip_levels: int = 0, arithmetic: int = ARITH_FP32,
                 general_dispatcher=None, fast_dispatcher=None):
        pipes = pipelines(self.fmt, fast, arithmetic)
        if general_dispatcher is None and fast_dispatcher is None:
            nvpro_cmd_pyramid_dispatch(stream, pipes, self.image, mip_levels)
        else:
            nvpro_cmd_pyramid_dispatch_custom(stream, pipes, self.image, mip_levels,
                                              general_dispatcher or default_general_dispatcher(),
                                              fast_dispatcher or default_fast_dispatcher())

    def free(self):
        if self.ptr:
            self.rt.free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def batch_config(mode: int = BATCH_MODE_QUEUE, workgroups_per_cu: int = 0, item_bytes: int = 65536,
                 tail_texels: int = BATCH_MAX_TAIL_TEXELS, use_graph: bool = True,
                 spin_limit: int = 0) -> NvproPyramidBatchConfig:
    c = NvproPyramidBatchConfig()
    _check(lib().nvproPyramidBatchDefaultConfig(ctypes.byref(c)))
    c.mode, c.workgroupsPerCU, c.itemBytes = mode, workgroups_per_cu, item_bytes
    c.tailTexels, c.useGraph, c.spinLimit = tail_texels, 1 if use_graph else 0, spin_limit
    return c


def _image_array(images):
    arr = (NvproPyramidImage * len(images))()
    for i, im in enumerate(images):
        arr[i] = im
    return arr


def _mip_array(mip_levels, n):
    if mip_levels is None:
        return None
    return (ctypes.c_uint32 * n)(*mip_levels)


class PyramidBatch:
    """nvproPyramidBatch*: many images' chains in one persistent launch (or one per round)."""

    def __init__(self, pipes: NvproPyramidPipelines, images, mip_levels=None,
                 config: NvproPyramidBatchConfig | None = None):
        self._images = _image_array(images)  # keeps offset/pitch arrays alive with the structs
        self._keep = [getattr(im, "_keep", None) for im in images]
        mips = _mip_array(mip_levels, len(images))
        h = ctypes.c_void_p()
        _check(lib().nvproPyramidBatchCreate(ctypes.byref(pipes), self._images, mips, len(images),
                                             ctypes.byref(config) if config is not None else None,
                                             ctypes.byref(h)))
        self.handle = h.value

    def launch(self, stream: int = 0):
        _check(lib().nvproPyramidBatchLaunch(ctypes.c_void_p(self.handle), ctypes.c_void_p(stream)))

    def info(self) -> NvproPyramidBatchInfo:
        out = NvproPyramidBatchInfo()
        _check(lib().nvproPyramidBatchGetInfo(ctypes.c_void_p(self.handle), ctypes.byref(out)))
        return out

    def read_errors(self) -> int:
        flags = ctypes.c_uint32()
        _check(lib().nvproPyramidBatchReadErrors(ctypes.c_void_p(self.handle), ctypes.byref(flags)))
        return flags.value

    def destroy(self):
        if self.handle:
            lib().nvproPyramidBatchDestroy(ctypes.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


NO_ALLOC_SCRATCH_BYTES = 4 << 20  # NVPRO_PYRAMID_NO_ALLOC_SCRATCH_BYTES


def batch_dispatch_no_alloc(stream: int, pipes: NvproPyramidPipelines, images, mip_levels=None,
                            config: NvproPyramidBatchConfig | None = None):
    """nvproPyramidBatchDispatchNoAlloc: one rounds-mode batch through the static scratch."""
    arr = _image_array(images)
    _check(lib().nvproPyramidBatchDispatchNoAlloc(ctypes.c_void_p(stream), ctypes.byref(pipes), arr,
                                                  _mip_array(mip_levels, len(images)), len(images),
                                                  ctypes.byref(config) if config is not None else None))


def batch_autotune(pipes: NvproPyramidPipelines, images, mip_levels=None, stream: int = 0, repetitions: int = 5):
    """The project autotuner: (best NvproPyramidBatchConfig, its microseconds per launch)."""
    arr = _image_array(images)
    best = NvproPyramidBatchConfig()
    us = ctypes.c_float()
    _check(lib().nvproPyramidBatchAutotune(ctypes.byref(pipes), arr, _mip_array(mip_levels, len(images)),
                                           len(images), ctypes.c_void_p(stream), repetitions, ctypes.byref(best),
                                           ctypes.byref(us)))
    return best, us.value


def server_config(budget_bytes: int = 0, max_in_flight: int = 0, max_batch_images: int = 0,
                  arithmetic16f: int = ARITH_FP32, batch: NvproPyramidBatchConfig | None = None):
    c = NvproPyramidServerConfig()
    _check(lib().nvproPyramidServerDefaultConfig(ctypes.byref(c)))
    if budget_bytes:
        c.budgetBytes = budget_bytes
    if max_in_flight:
        c.maxInFlight = max_in_flight
    if max_batch_images:
        c.maxBatchImages = max_batch_images
    c.arithmetic16f = arithmetic16f
    if batch is not None:
        c.batch = batch
    return c


class PyramidServer:
    """nvproPyramidServer*: admission by HBM budget, resident chains, continuous batching."""

    def __init__(self, config: NvproPyramidServerConfig | None = None):
        h = ctypes.c_void_p()
        _check(lib().nvproPyramidServerCreate(ctypes.byref(config) if config is not None else None,
                                              ctypes.byref(h)))
        self.handle = h.value
        self._host = {}  # ticket -> host level 0 kept alive until completion

    def submit(self, fmt: int, width: int, height: int, mip_levels: int = 0, seed: int = 0, level0=None) -> int:
        r = NvproPyramidRequest(fmt, width, height, mip_levels, seed, None)
        if level0 is not None:
            import numpy as np

            level0 = np.ascontiguousarray(level0)
            r.level0 = level0.ctypes.data
        t = ctypes.c_uint64()
        _check(lib().nvproPyramidServerSubmit(ctypes.c_void_p(self.handle), ctypes.byref(r), ctypes.byref(t)))
        if level0 is not None:
            self._host[t.value] = level0
        return t.value

    def poll(self, capacity: int = 256):
        arr = (NvproPyramidCompletion * capacity)()
        n = ctypes.c_uint32()
        _check(lib().nvproPyramidServerPoll(ctypes.c_void_p(self.handle), arr, capacity, ctypes.byref(n)))
        out = [arr[i] for i in range(n.value)]
        for c in out:
            self._host.pop(c.ticket, None)
        return out

    def release(self, ticket: int):
        _check(lib().nvproPyramidServerRelease(ctypes.c_void_p(self.handle), ticket))

    def stats(self) -> NvproPyramidServerStats:
        st = NvproPyramidServerStats()
        _check(lib().nvproPyramidServerGetStats(ctypes.c_void_p(self.handle), ctypes.byref(st)))
        return st

    def destroy(self):
        if self.handle:
            lib().nvproPyramidServerDestroy(ctypes.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
