"""MI355X (gfx950) HIP implementation of the nvpro_pyramid mip-chain schedule.

The compute path is the C-ABI library lib/libnvpro_pyramid_hip.so (built from
csrc/ by `python -m vk_compute_mipmaps_amd.build`); `pyramid` is the Python
mirror of the reference's host interface over that library.
"""
from . import pyramid  # noqa: F401
