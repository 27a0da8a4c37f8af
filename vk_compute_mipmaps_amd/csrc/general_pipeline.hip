// General pipeline (nvpro_pyramid.glsl:536-884) for gfx950: 1 or 2 levels of
// any size with the 2x2 / 2x3 / 3x2 / 3x3 energy-conserving kernel.
//
// One level: a thread per output texel (glsl:828-842).
// Two levels: a workgroup owns an 8x8 tile of the second level; it computes the
// (up to) 17x17 first-level tile it needs into LDS, writing it out as well
// (fillIntermediateTile_, glsl:732-781), then the 8x8 tile from LDS
// (fillLastTile_, glsl:791-820).  Neighbouring tiles recompute their shared
// first-level border texels and store identical values.
#include "kernel_common.hpp"

namespace nvpro_pyramid {

constexpr int kGeneralBS = 256;

// One level: outputs [256 g, 256 g + 256) of level inLevel+1, a thread each.
template <class F>
__device__ __forceinline__ void generalOneGroup(const typename F::Tables& tab, const ImageDesc& img, uint32_t inLevel,
                                                uint32_t g)
{
  const LevelDesc& src = img.lv[inLevel];
  const LevelDesc& dst = img.lv[inLevel + 1u];
  const int kx = kernelSize(src.width), ky = kernelSize(src.height);
  const uint32_t i = g * kGeneralBS + threadIdx.x;
  if(i >= dst.width * dst.height)
    return;
  const uint32_t dy = i / dst.width, dx = i - dy * dst.width;
  const uint8_t* p  = img.base + src.offset + (uint64_t)(2u * dy) * src.pitch + (uint64_t)(2u * dx) * F::kBytes;
  auto at           = [&](int j, int r) {
    return F::template load<false>(tab, loadTexel<F>(p + (uint64_t)r * src.pitch + (uint64_t)j * F::kBytes));
  };
  const typename F::Val v = generalSample<F>(at, kx, ky, dst.width, dst.height, dx, dy);
  storeTexel<F>(texelPtr(img, inLevel + 1u, dx, dy, F::kBytes), F::store(tab, v));
}

template <class F>
struct GeneralShared
{
  typename F::Tables tab;
  typename F::Val mid[17][17];  // sharedLevel_ (nvpro_pyramid.glsl:555), [y][x]
};

// Two levels: the 8x8 tile `tile` of level inLevel+2 and the first-level texels
// it needs.  Every thread of the workgroup calls this (it contains barriers).
template <class F>
__device__ __forceinline__ void generalTwoGroup(GeneralShared<F>& sh, const ImageDesc& img, uint32_t inLevel,
                                                uint32_t tilesX, uint32_t tile)
{
  const auto& tab       = sh.tab;
  const LevelDesc& src  = img.lv[inLevel];
  const LevelDesc& l1   = img.lv[inLevel + 1u];
  const LevelDesc& l2   = img.lv[inLevel + 2u];
  const int kx1 = kernelSize(src.width), ky1 = kernelSize(src.height);
  const int kx2 = kernelSize(l1.width), ky2 = kernelSize(l1.height);
  const uint32_t nx = kx2 == 3 ? 17u : 16u, ny = ky2 == 3 ? 17u : 16u;
  const uint32_t ty = tile / tilesX, tx = tile - ty * tilesX;
  // First level: the tile of level inLevel+1 the 8x8 output tile needs
  // (fillIntermediateTile_, nvpro_pyramid.glsl:732-781), with bounds checks.
  for(uint32_t e = threadIdx.x; e < nx * ny; e += kGeneralBS)
  {
    const uint32_t ly = e / nx, lx = e - ly * nx;
    const uint32_t gx = 16u * tx + lx, gy = 16u * ty + ly;
    if(gx < l1.width && gy < l1.height)
    {
      const uint8_t* p = img.base + src.offset + (uint64_t)(2u * gy) * src.pitch + (uint64_t)(2u * gx) * F::kBytes;
      auto at          = [&](int j, int r) {
        return F::template load<false>(tab, loadTexel<F>(p + (uint64_t)r * src.pitch + (uint64_t)j * F::kBytes));
      };
      const typename F::Val v = generalSample<F>(at, kx1, ky1, l1.width, l1.height, gx, gy);
      storeTexel<F>(texelPtr(img, inLevel + 1u, gx, gy, F::kBytes), F::store(tab, v));
      sh.mid[ly][lx] = v;
    }
  }
  __syncthreads();
  // Second level from LDS (fillLastTile_, nvpro_pyramid.glsl:791-820).
  if(threadIdx.x < 64u)
  {
    const uint32_t ox = threadIdx.x & 7u, oy = threadIdx.x >> 3;
    const uint32_t gx = 8u * tx + ox, gy = 8u * ty + oy;
    if(gx < l2.width && gy < l2.height)
    {
      auto at                 = [&](int j, int r) { return sh.mid[2u * oy + r][2u * ox + j]; };
      const typename F::Val v = generalSample<F>(at, kx2, ky2, l2.width, l2.height, gx, gy);
      storeTexel<F>(texelPtr(img, inLevel + 2u, gx, gy, F::kBytes), F::store(tab, v));
    }
  }
  __syncthreads();
}

template <class F>
__global__ void __launch_bounds__(kGeneralBS) generalOneLevelKernel(ImageDesc img, uint32_t inLevel, uint32_t groups)
{
  __shared__ typename F::Tables tab;
  F::template fill<kGeneralBS>(tab, 1.0f);
  __syncthreads();
  for(uint32_t g = blockIdx.x; g < groups; g += gridDim.x)
    generalOneGroup<F>(tab, img, inLevel, g);
}

template <class F>
__global__ void __launch_bounds__(kGeneralBS) generalTwoLevelKernel(ImageDesc img, uint32_t inLevel, uint32_t tilesX,
                                                                  uint32_t numTiles)
{
  __shared__ GeneralShared<F> sh;
  F::template fill<kGeneralBS>(sh.tab, 1.0f);
  __syncthreads();
  for(uint32_t tile = blockIdx.x; tile < numTiles; tile += gridDim.x)
    generalTwoGroup<F>(sh, img, inLevel, tilesX, tile);
}

template <class F>
__global__ void __launch_bounds__(kGeneralBS)
    generalOneLevelBatchKernel(const ImageDesc* __restrict__ imgs, const BatchSeg* __restrict__ segs, uint32_t numSegs,
                               uint32_t totalGroups)
{
  __shared__ typename F::Tables tab;
  F::template fill<kGeneralBS>(tab, 1.0f);
  __syncthreads();
  for(uint32_t g = blockIdx.x; g < totalGroups; g += gridDim.x)
  {
    const BatchSeg s = segs[findSegment(segs, numSegs, g)];
    generalOneGroup<F>(tab, imgs[s.image], s.inputLevel, g - s.groupStart);
  }
}

template <class F>
__global__ void __launch_bounds__(kGeneralBS)
    generalTwoLevelBatchKernel(const ImageDesc* __restrict__ imgs, const BatchSeg* __restrict__ segs,
                               uint32_t numSegs, uint32_t totalGroups)
{
  __shared__ GeneralShared<F> sh;
  F::template fill<kGeneralBS>(sh.tab, 1.0f);
  __syncthreads();
  for(uint32_t g = blockIdx.x; g < totalGroups; g += gridDim.x)
  {
    const BatchSeg s = segs[findSegment(segs, numSegs, g)];
    generalTwoGroup<F>(sh, imgs[s.image], s.inputLevel, s.tilesX, g - s.groupStart);
  }
}

uint32_t generalGroupsPerStep(uint32_t dstWidth, uint32_t dstHeight, uint32_t count, uint32_t* tilesX)
{
  if(count == 1u)
  {
    *tilesX = 0;
    return (uint32_t)(((uint64_t)dstWidth * dstHeight + kGeneralBS - 1) / kGeneralBS);
  }
  *tilesX = (dstWidth + 7u) / 8u;
  return *tilesX * ((dstHeight + 7u) / 8u);
}

template <class F>
static hipError_t launchGeneralBatchTyped(hipStream_t stream, uint32_t count, const ImageDesc* imgs,
                                          const BatchSeg* segs, uint32_t numSegs, uint32_t totalGroups)
{
  const uint64_t cap    = (uint64_t)deviceComputeUnits() * 8u;
  const uint32_t blocks = (uint32_t)(totalGroups < cap ? totalGroups : cap);
  if(count == 1u)
    hipLaunchKernelGGL((generalOneLevelBatchKernel<F>), dim3(blocks ? blocks : 1), dim3(kGeneralBS), 0, stream,
                       imgs, segs, numSegs, totalGroups);
  else
    hipLaunchKernelGGL((generalTwoLevelBatchKernel<F>), dim3(blocks ? blocks : 1), dim3(kGeneralBS), 0, stream,
                       imgs, segs, numSegs, totalGroups);
  return hipGetLastError();
}

hipError_t launchGeneralBatch(hipStream_t stream, uint32_t format, uint32_t arithmetic, uint32_t count,
                              const ImageDesc* imgs, const BatchSeg* segs, uint32_t numSegs, uint32_t totalGroups)
{
  switch(format)
  {
    case NVPRO_PYRAMID_FORMAT_SRGBA8:
      return launchGeneralBatchTyped<Srgba8>(stream, count, imgs, segs, numSegs, totalGroups);
    case NVPRO_PYRAMID_FORMAT_RGBA16F:
      return arithmetic == NVPRO_PYRAMID_ARITH_PACKED_FP16
                 ? launchGeneralBatchTyped<Rgba16fPacked>(stream, count, imgs, segs, numSegs, totalGroups)
                 : launchGeneralBatchTyped<Rgba16f>(stream, count, imgs, segs, numSegs, totalGroups);
    case NVPRO_PYRAMID_FORMAT_RGBA32F:
      return launchGeneralBatchTyped<Rgba32f>(stream, count, imgs, segs, numSegs, totalGroups);
    default:
      return hipErrorInvalidValue;
  }
}

template <class F>
static hipError_t launchGeneralTyped(hipStream_t stream, const ImageDesc& img, uint32_t inLevel, uint32_t count)
{
  const uint64_t cap = (uint64_t)deviceComputeUnits() * 8u;
  if(count == 1u)
  {
    const LevelDesc& d    = img.lv[inLevel + 1u];
    const uint64_t want   = ((uint64_t)d.width * d.height + kGeneralBS - 1) / kGeneralBS;
    const uint32_t blocks = (uint32_t)(want < cap ? want : cap);
    hipLaunchKernelGGL((generalOneLevelKernel<F>), dim3(blocks ? blocks : 1), dim3(kGeneralBS), 0, stream, img,
                       inLevel, (uint32_t)want);
  }
  else
  {
    const LevelDesc& d    = img.lv[inLevel + 2u];
    const uint32_t tilesX = (d.width + 7u) / 8u, tilesY = (d.height + 7u) / 8u;
    const uint64_t want   = (uint64_t)tilesX * tilesY;
    const uint32_t blocks = (uint32_t)(want < cap ? want : cap);
    hipLaunchKernelGGL((generalTwoLevelKernel<F>), dim3(blocks ? blocks : 1), dim3(kGeneralBS), 0, stream, img,
                       inLevel, tilesX, tilesX * tilesY);
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
    const LevelDesc& first    = img.lv[s.inputLevel + 1u];
    if((uint64_t)first.width * first.height <= kTailMaxTexels)
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
