// Tail kernel: the small remainder of a plan in one workgroup of kTailBS
// threads (tailRun in groups.hpp): with kTailMaxTexels outputs at most in its
// first level, every level is one pass of the workgroup.
#include "groups.hpp"

namespace nvpro_pyramid {

template <class F, int MAXT>
struct TailShared
{
  typename F::Tables tab;
  typename F::Val buf[2][MAXT];
};

constexpr int kTailBS = 1024;

template <class F, int MAXT>
__global__ void __launch_bounds__(kTailBS) tailKernel(ImageDesc img, TailPlan plan)
{
  __shared__ TailShared<F, MAXT> sh;
  uint32_t lines = prefetchLevels(img) ^ plan.count;  // the plan's lines too
#pragma unroll
  for(uint32_t k = 0; k < NVPRO_PYRAMID_MAX_LEVELS; ++k)
    lines ^= plan.steps[k].inputLevel;
  F::template fill<kTailBS>(sh.tab, 1.0f);
  __syncthreads();
  levelsReady(lines);
  tailRun<F, kTailBS>(sh.tab, sh.buf[0], sh.buf[1], img, plan);
}

template <class F>
static hipError_t launchTailTyped(hipStream_t stream, const ImageDesc& img, const TailPlan& plan)
{
  hipLaunchKernelGGL((tailKernel<F, kTailMaxTexels>), dim3(1), dim3(kTailBS), 0, stream, img, plan);
  return hipGetLastError();
}

hipError_t launchTail(hipStream_t stream, const ImageDesc& img, const TailPlan& plan)
{
  switch(img.format)
  {
    case NVPRO_PYRAMID_FORMAT_SRGBA8:
      return launchTailTyped<Srgba8>(stream, img, plan);
    case NVPRO_PYRAMID_FORMAT_RGBA16F:
      return img.arithmetic == NVPRO_PYRAMID_ARITH_PACKED_FP16 ? launchTailTyped<Rgba16fPacked>(stream, img, plan)
                                                               : launchTailTyped<Rgba16f>(stream, img, plan);
    case NVPRO_PYRAMID_FORMAT_RGBA32F:
      return launchTailTyped<Rgba32f>(stream, img, plan);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace nvpro_pyramid
