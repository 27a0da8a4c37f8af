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
  uint32_t count, image, pad[2];  // image: index into the batch's ImageDesc array
  TailStep steps[NVPRO_PYRAMID_MAX_LEVELS];
};

// One image's share of a batched launch: tile groups
// [groupStart, next segment's groupStart) belong to step (inputLevel) of image.
struct BatchSeg
{
  uint32_t image, inputLevel, tilesX, numTiles, groupStart, pad[3];
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

// Batched launchers: one launch covers the segments of many images (device
// arrays).  Fast: all segments fill `count` levels.  General: `count` 1 or 2.
hipError_t launchFastBatch(hipStream_t stream, uint32_t format, uint32_t arithmetic, uint32_t count,
                           const ImageDesc* imgs, const BatchSeg* segs, uint32_t numSegs, uint32_t totalGroups,
                           uint32_t wgPerCU);
hipError_t launchGeneralBatch(hipStream_t stream, uint32_t format, uint32_t arithmetic, uint32_t count,
                              const ImageDesc* imgs, const BatchSeg* segs, uint32_t numSegs, uint32_t totalGroups);
hipError_t launchTailBatch(hipStream_t stream, uint32_t format, uint32_t arithmetic, const ImageDesc* imgs,
                           const TailPlan* plans, uint32_t numPlans);
// Tile groups one launch of the fast / general kernel uses for a step.
uint32_t fastGroupsPerStep(uint32_t width, uint32_t height, uint32_t count);
uint32_t generalGroupsPerStep(uint32_t dstWidth, uint32_t dstHeight, uint32_t count, uint32_t* tilesX);

// Launch a validated plan: per-step kernels for the large steps, then one tail
// launch for the small remainder.
hipError_t executePlan(hipStream_t stream, const ImageDesc& img, const NvproPyramidStep* steps, uint32_t count);

// Number of compute units of the current device (cached per device).
uint32_t deviceComputeUnits();

}  // namespace nvpro_pyramid
