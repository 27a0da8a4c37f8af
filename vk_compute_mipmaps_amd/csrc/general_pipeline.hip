// General pipeline kernels (nvpro_pyramid.glsl:536-884) for gfx950: one launch
// per general step of 1 or 2 levels.  The group bodies (generalOneGroup,
// generalTwoGroup) are in groups.hpp.  Also the per-step executor.
#include "groups.hpp"

namespace nvpro_pyramid {

template <class F>
struct GeneralShared
{
  typename F::Tables tab;
  typename F::Val mid[gen2MidTexels<F>()];
};

template <class F>
__global__ void __launch_bounds__(kGroupBS) generalOneLevelKernel(ImageDesc img, uint32_t inLevel, uint32_t groups)
{
  __shared__ typename F::Tables tab;
  F::template fill<kGroupBS>(tab, 1.0f);
  __syncthreads();
  for(uint32_t g = blockIdx.x; g < groups; g += gridDim.x)
    generalOneGroup<F>(tab, img, inLevel, g);
}

template <class F>
__global__ void __launch_bounds__(kGroupBS) generalTwoLevelKernel(ImageDesc img, uint32_t inLevel, uint32_t tilesX,
                                                                uint32_t numTiles)
{
  __shared__ GeneralShared<F> sh;
  F::template fill<kGroupBS>(sh.tab, 1.0f);
  __syncthreads();
  for(uint32_t tile = blockIdx.x; tile < numTiles; tile += gridDim.x)
    generalTwoGroup<F>(sh.tab, sh.mid, img, inLevel, tilesX, tile);
}

// Workgroups of `kernel` that are resident on one compute unit at once.  A
// persistent launch is capped there: more workgroups would only start after
// the first ones had finished all their tiles.
template <class K>
static uint32_t residentPerCU(K kernel)
{
  int n = 0;
  if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, kGroupBS, 0) != hipSuccess || n <= 0)
    n = 1;
  return (uint32_t)n;
}

template <class F>
static hipError_t launchGeneralTyped(hipStream_t stream, const ImageDesc& img, uint32_t inLevel, uint32_t count)
{
  static const uint32_t oneFit = residentPerCU(generalOneLevelKernel<F>);
  static const uint32_t twoFit = residentPerCU(generalTwoLevelKernel<F>);
  if(count == 1u)
  {
    const LevelDesc& d    = img.lv[inLevel + 1u];
    const uint64_t want   = ((uint64_t)d.width * d.height + kGroupBS - 1) / kGroupBS;
    const uint64_t cap    = (uint64_t)deviceComputeUnits() * oneFit;
    const uint32_t blocks = (uint32_t)(want < cap ? want : cap);
    hipLaunchKernelGGL((generalOneLevelKernel<F>), dim3(blocks ? blocks : 1), dim3(kGroupBS), 0, stream, img,
                       inLevel, (uint32_t)want);
  }
  else
  {
    const LevelDesc& d    = img.lv[inLevel + 2u];
    uint32_t tilesX       = 0;
    const uint64_t want   = generalGroupCount(d.width, d.height, 2u, F::kBytes, &tilesX);
    const uint64_t cap    = (uint64_t)deviceComputeUnits() * twoFit;
    const uint32_t blocks = (uint32_t)(want < cap ? want : cap);
    hipLaunchKernelGGL((generalTwoLevelKernel<F>), dim3(blocks ? blocks : 1), dim3(kGroupBS), 0, stream, img,
                       inLevel, tilesX, (uint32_t)want);
  }
  return hipGetLastError();
}

hipError_t launchGeneral(hipStream_t stream, const ImageDesc& img, uint32_t inLevel, uint32_t count)
{
  switch(img.format)
  {
    case NVPRO_PYRAMID_FORMAT_SRGBA8:
      return launchGeneralTyped<Srgba8>(stream, img, inLevel, count);
    case NVPRO_PYRAMID_FORMAT_RGBA16F:
      return img.arithmetic == NVPRO_PYRAMID_ARITH_PACKED_FP16
                 ? launchGeneralTyped<Rgba16fPacked>(stream, img, inLevel, count)
                 : launchGeneralTyped<Rgba16f>(stream, img, inLevel, count);
    case NVPRO_PYRAMID_FORMAT_RGBA32F:
      return launchGeneralTyped<Rgba32f>(stream, img, inLevel, count);
    default:
      return hipErrorInvalidValue;
  }
}

uint32_t deviceComputeUnits()
{
  static uint32_t cached[64] = {};
  int dev = 0;
  if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
    return 256u;
  if(cached[dev] == 0u)
  {
    int cus = 0;
    if(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    cached[dev] = (uint32_t)cus;
  }
  return cached[dev];
}

hipError_t executePlan(hipStream_t stream, const ImageDesc& img, const NvproPyramidStep* steps, uint32_t count)
{
  for(uint32_t i = 0; i < count; ++i)
  {
    const NvproPyramidStep& s = steps[i];
    // The one-workgroup tail pays off when it saves launches: with a single step
    // left, that step's own kernel is as quick.
    const LevelDesc& first = img.lv[s.inputLevel + 1u];
    if((uint64_t)first.width * first.height <= kTailMaxTexels && count - i >= 2u)
    {
      TailPlan plan{};
      for(uint32_t k = i; k < count && plan.count < NVPRO_PYRAMID_MAX_LEVELS; ++k)
        plan.steps[plan.count++] = TailStep{steps[k].pipeline, steps[k].inputLevel, steps[k].levelCount, 0u};
      return launchTail(stream, img, plan);
    }
    hipError_t e = s.pipeline == NVPRO_PYRAMID_PIPELINE_FAST ? launchFast(stream, img, s.inputLevel, s.levelCount, 8u)
                                                             : launchGeneral(stream, img, s.inputLevel, s.levelCount);
    if(e != hipSuccess)
      return e;
  }
  return hipSuccess;
}

}  // namespace nvpro_pyramid
