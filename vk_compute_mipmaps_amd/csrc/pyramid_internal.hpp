// Internal declarations shared by the C-ABI host code and the HIP kernels.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

#include "nvpro_pyramid_hip.h"

namespace nvpro_pyramid {

// One mip level in linear device memory.
struct LevelDesc
{
  uint64_t offset;  // bytes from ImageDesc::base
  uint32_t width, height;
  uint32_t pitch;   // bytes per row
  uint32_t pad;
};

// A bound image: the kernels' view of NvproPyramidImage.
struct ImageDesc
{
  uint8_t* base;
  uint32_t levelCount;
  uint32_t format;
  uint32_t arithmetic;
  uint32_t pad;
  LevelDesc lv[NVPRO_PYRAMID_MAX_LEVELS];
};

// Launch one dispatch of the fast or general pipeline over levels
// inputLevel+1 .. inputLevel+levelCount.  The caller has validated the step.
hipError_t launchStep(hipStream_t stream, const ImageDesc& img, uint32_t pipeline,
                      uint32_t inputLevel, uint32_t levelCount);

// Number of compute units of the current device (cached per device).
uint32_t deviceComputeUnits();

}  // namespace nvpro_pyramid
