// Tail of a plan in one workgroup.
//
// Once a plan reaches levels of at most kTailMaxTexels texels, launching one
// kernel per step only buys launch latency.  This kernel runs all remaining
// steps of the plan, fast and general alike, in a single 1024-thread workgroup.
// Every level is computed by the whole workgroup, stored to memory, and kept in
// LDS as float4 for the next level; at a step boundary the value kept is the
// decode of the stored texel, exactly what the next dispatch would load.  The
// arithmetic per level is the same as the per-step kernels (fast: box with the
// reference's order; general: reduceStoreSample_), so results do not depend on
// which kernel ran a step.
#include "kernel_common.hpp"

namespace nvpro_pyramid {

template <class F, int MAXT>
struct TailShared
{
  typename F::Tables tab;
  typename F::Val buf[2][MAXT];
};

// Runs every step of `plan` on `img` with the BS threads of one workgroup.
template <class F, int BS, int MAXT>
__device__ __forceinline__ void tailRun(TailShared<F, MAXT>& sh, const ImageDesc& img, const TailPlan& plan)
{
  using V         = typename F::Val;
  const auto& tab = sh.tab;
  int cur         = -1;  // buffer holding the current source level, -1: read it from memory
  for(uint32_t s = 0; s < plan.count; ++s)
  {
    const uint32_t L     = plan.steps[s].inputLevel;
    const uint32_t count = plan.steps[s].levelCount;
    const bool fast      = plan.steps[s].pipeline == NVPRO_PYRAMID_PIPELINE_FAST;
    for(uint32_t r = 1; r <= count; ++r)
    {
      const LevelDesc& src = img.lv[L + r - 1u];
      const LevelDesc& dst = img.lv[L + r];
      const int dbuf       = cur == 0 ? 1 : 0;
      const int kx = kernelSize(src.width), ky = kernelSize(src.height);
      const int order      = fastOrder(count, r);
      const uint32_t n     = dst.width * dst.height;
      for(uint32_t i = threadIdx.x; i < n; i += BS)
      {
        const uint32_t dy = i / dst.width, dx = i - dy * dst.width;
        auto at           = [&](int j, int k) -> V {
          const uint32_t x = 2u * dx + (uint32_t)j, y = 2u * dy + (uint32_t)k;
          if(cur < 0)
          {
            const uint8_t* p = img.base + src.offset + (uint64_t)y * src.pitch + (uint64_t)x * F::kBytes;
            return F::template load<false>(tab, loadTexel<F>(p));
          }
          return sh.buf[cur][y * src.width + x];
        };
        V v;
        if(fast)
          v = box<F, false>(at(0, 0), at(1, 0), at(0, 1), at(1, 1), order);
        else
          v = generalSample<F>(at, kx, ky, dst.width, dst.height, dx, dy);
        const typename F::Raw q = F::store(tab, v);
        storeTexel<F>(img.base + dst.offset + (uint64_t)dy * dst.pitch + (uint64_t)dx * F::kBytes, q);
        if(r == count)
          v = F::template load<false>(tab, q);  // next step reads the stored texel
        sh.buf[dbuf][i] = v;
      }
      __syncthreads();
      cur = dbuf;
    }
  }
}

template <class F, int BS, int MAXT>
__global__ void __launch_bounds__(BS) tailKernel(ImageDesc img, TailPlan plan)
{
  __shared__ TailShared<F, MAXT> sh;
  F::template fill<BS>(sh.tab, 1.0f);
  __syncthreads();
  tailRun<F, BS, MAXT>(sh, img, plan);
}

// Batched form: workgroup i runs plans[i] on imgs[plans[i].image].
template <class F, int BS, int MAXT>
__global__ void __launch_bounds__(BS) tailBatchKernel(const ImageDesc* __restrict__ imgs,
                                                      const TailPlan* __restrict__ plans)
{
  __shared__ TailShared<F, MAXT> sh;
  F::template fill<BS>(sh.tab, 1.0f);
  __syncthreads();
  const TailPlan& plan = plans[blockIdx.x];
  tailRun<F, BS, MAXT>(sh, imgs[plan.image], plan);
}

template <class F>
static hipError_t launchTailBatchTyped(hipStream_t stream, const ImageDesc* imgs, const TailPlan* plans, uint32_t n)
{
  hipLaunchKernelGGL((tailBatchKernel<F, kTailBS, kTailMaxTexels>), dim3(n), dim3(kTailBS), 0, stream, imgs, plans);
  return hipGetLastError();
}

hipError_t launchTailBatch(hipStream_t stream, uint32_t format, uint32_t arithmetic, const ImageDesc* imgs,
                           const TailPlan* plans, uint32_t numPlans)
{
  if(numPlans == 0)
    return hipSuccess;
  switch(format)
  {
    case NVPRO_PYRAMID_FORMAT_SRGBA8:
      return launchTailBatchTyped<Srgba8>(stream, imgs, plans, numPlans);
    case NVPRO_PYRAMID_FORMAT_RGBA16F:
      return arithmetic == NVPRO_PYRAMID_ARITH_PACKED_FP16 ? launchTailBatchTyped<Rgba16fPacked>(stream, imgs, plans, numPlans)
                                                           : launchTailBatchTyped<Rgba16f>(stream, imgs, plans, numPlans);
    case NVPRO_PYRAMID_FORMAT_RGBA32F:
      return launchTailBatchTyped<Rgba32f>(stream, imgs, plans, numPlans);
    default:
      return hipErrorInvalidValue;
  }
}

template <class F>
static hipError_t launchTailTyped(hipStream_t stream, const ImageDesc& img, const TailPlan& plan)
{
  hipLaunchKernelGGL((tailKernel<F, kTailBS, kTailMaxTexels>), dim3(1), dim3(kTailBS), 0, stream, img, plan);
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
