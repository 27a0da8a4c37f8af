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

// The steps a single tail launch runs (see tail_pipeline.hip).
struct TailStep
{
  uint32_t pipeline, inputLevel, levelCount, pad;
};
struct TailPlan
{
  uint32_t count, pad[3];
  TailStep steps[NVPRO_PYRAMID_MAX_LEVELS];
};

// A step whose first output level has at most kTailMaxTexels texels, and every
// step after it, runs in the one-workgroup tail kernel of kTailBS threads.
constexpr uint32_t kTailMaxTexels = 1024;
constexpr int kTailBS             = 256;

// Per-pipeline launchers.  The caller has validated every step.
hipError_t launchFast(hipStream_t stream, const ImageDesc& img, uint32_t inputLevel, uint32_t levelCount,
                      uint32_t wavesPerCU);
hipError_t launchGeneral(hipStream_t stream, const ImageDesc& img, uint32_t inputLevel, uint32_t levelCount);
hipError_t launchTail(hipStream_t stream, const ImageDesc& img, const TailPlan& plan);

// Launch a validated plan: per-step kernels for the large steps, then one tail
// launch for the small remainder.
hipError_t executePlan(hipStream_t stream, const ImageDesc& img, const NvproPyramidStep* steps, uint32_t count);

// Number of compute units of the current device (cached per device).
uint32_t deviceComputeUnits();

}  // namespace nvpro_pyramid
