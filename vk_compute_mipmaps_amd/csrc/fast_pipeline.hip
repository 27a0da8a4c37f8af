// Fast pipeline kernels (nvpro_pyramid.glsl:209-534) for gfx950, wave64: one
// launch per fast step.  The group body (fastGroup) is in groups.hpp.
//
// Workgroups are persistent (the LDS tables are filled once, then the
// workgroup loops over tile groups) and small enough in registers for eight
// waves per SIMD, so the hardware overlaps one wave's loads with another's
// arithmetic.
#include "groups.hpp"

namespace nvpro_pyramid {

template <class F, int COUNT>
struct FastShared
{
  typename F::Tables tab;
  typename F::Val level2[FastGeom<COUNT>::UPPER ? kGroupBS : 1];
};

template <class F, int COUNT>
__global__ void __launch_bounds__(kGroupBS) __attribute__((amdgpu_waves_per_eu(8, 8)))
fastPipelineKernel(ImageDesc img, uint32_t inLevel, uint32_t tilesX, uint32_t numTiles)
{
  __shared__ FastShared<F, COUNT> sh;
  F::template fill<kGroupBS>(sh.tab, F::kPrescale ? 0.25f : 1.0f);
  __syncthreads();
  const uint32_t numGroups = (numTiles + FastGeom<COUNT>::TILES_PER_WG - 1u) / FastGeom<COUNT>::TILES_PER_WG;
  uint32_t iter            = 0;
  for(uint32_t g = blockIdx.x; g < numGroups; g += gridDim.x, ++iter)
    fastGroup<F, COUNT>(sh.tab, sh.level2, img, inLevel, tilesX, numTiles, g, iter);
}

template <class F, int COUNT>
static void launchFastCount(hipStream_t stream, const ImageDesc& img, uint32_t inLevel, uint32_t wgPerCU)
{
  const LevelDesc& in     = img.lv[inLevel];
  const uint32_t tilesX   = in.width >> COUNT, tilesY = in.height >> COUNT;
  const uint64_t groups   = fastGroupCount(in.width, in.height, COUNT);
  const uint64_t cap      = (uint64_t)deviceComputeUnits() * wgPerCU;
  const uint32_t blocks   = (uint32_t)(groups < cap ? groups : cap);
  hipLaunchKernelGGL((fastPipelineKernel<F, COUNT>), dim3(blocks ? blocks : 1), dim3(kGroupBS), 0, stream, img,
                     inLevel, tilesX, tilesX * tilesY);
}

template <class F>
static hipError_t launchFastTyped(hipStream_t stream, const ImageDesc& img, uint32_t inLevel, uint32_t count,
                                  uint32_t wgPerCU)
{
  switch(count)
  {
    case 1: launchFastCount<F, 1>(stream, img, inLevel, wgPerCU); break;
    case 2: launchFastCount<F, 2>(stream, img, inLevel, wgPerCU); break;
    case 3: launchFastCount<F, 3>(stream, img, inLevel, wgPerCU); break;
    case 4: launchFastCount<F, 4>(stream, img, inLevel, wgPerCU); break;
    case 5: launchFastCount<F, 5>(stream, img, inLevel, wgPerCU); break;
    case 6: launchFastCount<F, 6>(stream, img, inLevel, wgPerCU); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launchFast(hipStream_t stream, const ImageDesc& img, uint32_t inLevel, uint32_t count, uint32_t wgPerCU)
{
  switch(img.format)
  {
    case NVPRO_PYRAMID_FORMAT_SRGBA8:
      return launchFastTyped<Srgba8>(stream, img, inLevel, count, wgPerCU);
    case NVPRO_PYRAMID_FORMAT_RGBA16F:
      return img.arithmetic == NVPRO_PYRAMID_ARITH_PACKED_FP16
                 ? launchFastTyped<Rgba16fPacked>(stream, img, inLevel, count, wgPerCU)
                 : launchFastTyped<Rgba16f>(stream, img, inLevel, count, wgPerCU);
    case NVPRO_PYRAMID_FORMAT_RGBA32F:
      return launchFastTyped<Rgba32f>(stream, img, inLevel, count, wgPerCU);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace nvpro_pyramid
