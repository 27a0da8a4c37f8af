// Fast pipeline kernels (nvpro_pyramid.glsl:209-534) for gfx950, wave64: one
// launch per fast step, or per two when the second is a single group or a
// small tail.  The group body (fastGroup) is in groups.hpp.
//
// Workgroups are persistent (the LDS tables are filled once, then the
// workgroup loops over tile groups) and small enough in registers for eight
// waves per SIMD, so the hardware overlaps one wave's loads with another's
// arithmetic (four for 16-byte texels, whose 4x4 block alone takes 64
// registers).
#include <map>
#include <mutex>

#include "groups.hpp"

namespace nvpro_pyramid {

// The arrival counters of fused launches: zero when the code object is loaded,
// and every fused launch leaves its block zeroed, so no memset is ever needed.
// Per device, this array is the only device memory the dispatch path uses
// besides the caller's image.
static __device__ __align__(256) uint32_t gFusedCounters[kFusedSlots * kArriveWords];

uint32_t* fusedCounterBlock(uint32_t slot)
{
  static std::mutex lock;
  static std::map<int, uint32_t*> base;  // by device
  int device = 0;
  if(slot >= kFusedSlots || hipGetDevice(&device) != hipSuccess)
  {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> guard(lock);
  auto it = base.find(device);
  if(it == base.end())
  {
    uint32_t* p = nullptr;
    if(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(gFusedCounters)) != hipSuccess)
    {
      (void)hipGetLastError();
      p = nullptr;
    }
    it = base.emplace(device, p).first;
  }
  return it->second ? it->second + (size_t)slot * kArriveWords : nullptr;
}

template <class F, int COUNT, bool CONT>
struct FastShared
{
  typename F::Tables tab;
  typename F::Val level2[(FastGeom<COUNT>::UPPER || CONT) ? kGroupBS : 1];
};

// A fused launch (CONT): the step's last level is stored coherently, and the
// last workgroup to finish (arriveLast) then runs `next`, a fast step of one
// group with the tables it already holds, or a tail plan with its ping-pong in
// the level-2 scratch and the tables refilled at the unscaled decode the tail's
// bodies load with.
static_assert(2u * kFastTailTexels <= (uint32_t)kGroupBS, "tail ping-pong fits the level-2 scratch");

// NT: the step's input loads carry the non-temporal hint.
template <class F, int COUNT, bool CONT, bool NT>
__global__ void __launch_bounds__(kGroupBS) __attribute__((amdgpu_waves_per_eu(F::kBytes > 8u ? 4 : 8, 8)))
fastPipelineKernel(ImageDesc img, uint32_t inLevel, uint32_t tilesX, uint32_t numTiles, NextStep next, TailPlan tail)
{
  __shared__ FastShared<F, COUNT, CONT> sh;
  const uint32_t lines = prefetchLevels(img);
  F::template fill<kGroupBS>(sh.tab, F::kPrescale ? 0.25f : 1.0f);
  __syncthreads();
  levelsReady(lines);
  const uint32_t numGroups = (numTiles + FastGeom<COUNT>::TILES_PER_WG - 1u) / FastGeom<COUNT>::TILES_PER_WG;
  uint32_t iter            = 0;
  for(uint32_t g = blockIdx.x; g < numGroups; g += gridDim.x, ++iter)
    fastGroup<F, COUNT, CONT, COUNT == 6, NT>(sh.tab, sh.level2, img, inLevel, tilesX, numTiles, g, iter);
  if constexpr(CONT)
  {
    if(!arriveLast(next.arrive, gridDim.x, reinterpret_cast<uint32_t*>(sh.level2)))
      return;
    if(next.kind == kNextTail)
    {
      if(F::kPrescale)
      {
        F::template fill<kGroupBS>(sh.tab, 1.0f);
        __syncthreads();
      }
      tailRun<F, kGroupBS, 1, false>(sh.tab, sh.level2, sh.level2 + kFastTailTexels, img, tail);
      return;
    }
    switch(next.count)
    {
      case 1: fastGroup<F, 1>(sh.tab, sh.level2, img, next.inLevel, next.tilesX, next.numTiles, 0u, iter); break;
      case 2: fastGroup<F, 2>(sh.tab, sh.level2, img, next.inLevel, next.tilesX, next.numTiles, 0u, iter); break;
      case 3: fastGroup<F, 3>(sh.tab, sh.level2, img, next.inLevel, next.tilesX, next.numTiles, 0u, iter); break;
      case 4: fastGroup<F, 4>(sh.tab, sh.level2, img, next.inLevel, next.tilesX, next.numTiles, 0u, iter); break;
      case 5: fastGroup<F, 5>(sh.tab, sh.level2, img, next.inLevel, next.tilesX, next.numTiles, 0u, iter); break;
      case 6: fastGroup<F, 6>(sh.tab, sh.level2, img, next.inLevel, next.tilesX, next.numTiles, 0u, iter); break;
      default: break;
    }
  }
}

// Groups per compute unit for which 6-level steps load their input with the
// non-temporal hint (measured on MI355X, sRGBA8 F(0,6), 4096^2 at 16 groups per
// unit: 22.4 -> 20.3 us; 16384^2 at 256 per unit, on 3 workgroups per unit:
// 291 -> 302 us).
constexpr uint64_t kNtMinGroupsPerCU = 8, kNtMaxGroupsPerCU = 16;

template <class F, int COUNT>
static void launchFastCount(hipStream_t stream, const ImageDesc& img, uint32_t inLevel, uint32_t wgPerCU,
                            const NextStep* next, const TailPlan* tail)
{
  const LevelDesc& in     = img.lv[inLevel];
  const uint32_t tilesX   = in.width >> COUNT, tilesY = in.height >> COUNT;
  const uint64_t groups   = fastGroupCount(in.width, in.height, COUNT);
  const uint64_t cap      = (uint64_t)deviceComputeUnits() * wgPerCU;
  const uint32_t blocks   = (uint32_t)(groups < cap ? groups : cap);
  const TailPlan none{};
  const dim3 grid(blocks ? blocks : 1), wg(kGroupBS);
  // Non-temporal input loads for 6-level steps in the kNtMin..kNtMaxGroupsPerCU range.
  const uint64_t cus = deviceComputeUnits();
  const bool nt      = COUNT == 6 && groups >= kNtMinGroupsPerCU * cus && groups <= kNtMaxGroupsPerCU * cus;
  if(next && nt)
    hipLaunchKernelGGL((fastPipelineKernel<F, COUNT, true, COUNT == 6>), grid, wg, 0, stream, img, inLevel, tilesX,
                       tilesX * tilesY, *next, tail ? *tail : none);
  else if(next)
    hipLaunchKernelGGL((fastPipelineKernel<F, COUNT, true, false>), grid, wg, 0, stream, img, inLevel, tilesX,
                       tilesX * tilesY, *next, tail ? *tail : none);
  else if(nt)
    hipLaunchKernelGGL((fastPipelineKernel<F, COUNT, false, COUNT == 6>), grid, wg, 0, stream, img, inLevel, tilesX,
                       tilesX * tilesY, NextStep{}, none);
  else
    hipLaunchKernelGGL((fastPipelineKernel<F, COUNT, false, false>), grid, wg, 0, stream, img, inLevel, tilesX,
                       tilesX * tilesY, NextStep{}, none);
}

template <class F>
static hipError_t launchFastTyped(hipStream_t stream, const ImageDesc& img, uint32_t inLevel, uint32_t count,
                                  uint32_t wgPerCU, const NextStep* next, const TailPlan* tail)
{
  switch(count)
  {
    case 1: launchFastCount<F, 1>(stream, img, inLevel, wgPerCU, next, tail); break;
    case 2: launchFastCount<F, 2>(stream, img, inLevel, wgPerCU, next, tail); break;
    case 3: launchFastCount<F, 3>(stream, img, inLevel, wgPerCU, next, tail); break;
    case 4: launchFastCount<F, 4>(stream, img, inLevel, wgPerCU, next, tail); break;
    case 5: launchFastCount<F, 5>(stream, img, inLevel, wgPerCU, next, tail); break;
    case 6: launchFastCount<F, 6>(stream, img, inLevel, wgPerCU, next, tail); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launchFast(hipStream_t stream, const ImageDesc& img, uint32_t inLevel, uint32_t count, uint32_t wgPerCU,
                      const NextStep* next, const TailPlan* tail)
{
  if(next && (!next->arrive || (next->kind == kNextFast && (next->count < 1u || next->count > 6u)) ||
              (next->kind == kNextTail && (!tail || tail->count == 0u || tail->count > NVPRO_PYRAMID_MAX_LEVELS)) ||
              (next->kind != kNextFast && next->kind != kNextTail)))
    return hipErrorInvalidValue;
  switch(img.format)
  {
    case NVPRO_PYRAMID_FORMAT_SRGBA8:
      return launchFastTyped<Srgba8>(stream, img, inLevel, count, wgPerCU, next, tail);
    case NVPRO_PYRAMID_FORMAT_RGBA16F:
      return img.arithmetic == NVPRO_PYRAMID_ARITH_PACKED_FP16
                 ? launchFastTyped<Rgba16fPacked>(stream, img, inLevel, count, wgPerCU, next, tail)
                 : launchFastTyped<Rgba16f>(stream, img, inLevel, count, wgPerCU, next, tail);
    case NVPRO_PYRAMID_FORMAT_RGBA32F:
      return launchFastTyped<Rgba32f>(stream, img, inLevel, count, wgPerCU, next, tail);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace nvpro_pyramid
