// General pipeline kernels (nvpro_pyramid.glsl:536-884) for gfx950: one launch
// per general step of 1 or 2 levels.  The group bodies (generalOneGroup,
// generalTwoGroup) are in groups.hpp.  Also the per-step executor.
#include <stdlib.h>

#include "groups.hpp"

namespace nvpro_pyramid {

template <class F, bool THIN>
struct GeneralShared
{
  typename F::Tables tab;
  typename F::Val mid[gen2MidTexels<F, THIN>()];
};

template <class F>
__global__ void __launch_bounds__(kGroupBS) generalOneLevelKernel(ImageDesc img, uint32_t inLevel, uint32_t groups)
{
  __shared__ typename F::Tables tab;
  const uint32_t lines = prefetchLevels(img);
  F::template fill<kGroupBS>(tab, 1.0f);
  __syncthreads();
  levelsReady(lines);
  for(uint32_t g = blockIdx.x; g < groups; g += gridDim.x)
    generalOneGroup<F>(tab, img, inLevel, g);
}

// KX, KY: the first level's kernel sizes, or 0 (read from the image).  THIN:
// the tile geometry (Gen2Geom) and one tile per workgroup (the grid is
// numTiles); otherwise persistent workgroups loop over the tiles.
template <class F, int KX, int KY, bool THIN>
__global__ void __launch_bounds__(kGroupBS) generalTwoLevelKernel(ImageDesc img, uint32_t inLevel, uint32_t tilesX,
                                                                uint32_t numTiles)
{
  __shared__ GeneralShared<F, THIN> sh;
  if(THIN && blockIdx.x >= numTiles)  // the same for the whole workgroup
    return;
  const uint32_t lines = prefetchLevels(img);
  F::template fill<kGroupBS>(sh.tab, 1.0f);
  __syncthreads();
  levelsReady(lines);
  if constexpr(THIN)  // one tile per workgroup, straight-line: ~20 fewer VGPRs than the loop
    generalTwoGroup<F, false, KX, KY, true>(sh.tab, sh.mid, img, inLevel, tilesX, blockIdx.x);
  else
    for(uint32_t tile = blockIdx.x; tile < numTiles; tile += gridDim.x)
      generalTwoGroup<F, false, KX, KY, false>(sh.tab, sh.mid, img, inLevel, tilesX, tile);
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

// A step whose thin tiles all fit on the GPU at once runs them (its time is
// then one tile's latency, which thin tiles halve); a larger one runs the full
// tiles, which recompute fewer border texels, on persistent workgroups.  The
// environment variable NVPRO_PYRAMID_THIN=1 / =0 forces thin / full tiles for
// every step (the bytes are the same either way).
template <class F, int KX, int KY>
static void launchTwoLevel(hipStream_t stream, const ImageDesc& img, uint32_t inLevel)
{
  static const uint32_t fit     = residentPerCU(generalTwoLevelKernel<F, KX, KY, false>);
  static const uint32_t fitThin = residentPerCU(generalTwoLevelKernel<F, KX, KY, true>);
  const LevelDesc& d            = img.lv[inLevel + 2u];
  uint32_t tilesX = 0, thinX = 0;
  const uint32_t numTiles = generalGroupCount(d.width, d.height, 2u, F::kBytes, &tilesX);
  const uint32_t numThin  = generalGroupCount(d.width, d.height, 2u, F::kBytes, &thinX, true);
  const char* force       = getenv("NVPRO_PYRAMID_THIN");
  const bool thin         = force && (force[0] == '0' || force[0] == '1')
                                ? force[0] == '1'
                                : (uint64_t)numThin <= (uint64_t)deviceComputeUnits() * fitThin;
  if(thin)
  {
    hipLaunchKernelGGL((generalTwoLevelKernel<F, KX, KY, true>), dim3(numThin), dim3(kGroupBS), 0, stream, img,
                       inLevel, thinX, numThin);
    return;
  }
  const uint64_t cap    = (uint64_t)deviceComputeUnits() * fit;
  const uint32_t blocks = (uint32_t)(numTiles < cap ? numTiles : cap);
  hipLaunchKernelGGL((generalTwoLevelKernel<F, KX, KY, false>), dim3(blocks ? blocks : 1), dim3(kGroupBS), 0, stream,
                     img, inLevel, tilesX, numTiles);
}

template <class F>
static hipError_t launchGeneralTyped(hipStream_t stream, const ImageDesc& img, uint32_t inLevel, uint32_t count)
{
  if(count == 1u)
  {
    static const uint32_t oneFit = residentPerCU(generalOneLevelKernel<F>);
    const LevelDesc& d           = img.lv[inLevel + 1u];
    uint32_t unused              = 0;
    const uint64_t want          = generalGroupCount(d.width, d.height, 1u, F::kBytes, &unused);
    const uint64_t cap    = (uint64_t)deviceComputeUnits() * oneFit;
    const uint32_t blocks = (uint32_t)(want < cap ? want : cap);
    hipLaunchKernelGGL((generalOneLevelKernel<F>), dim3(blocks ? blocks : 1), dim3(kGroupBS), 0, stream, img,
                       inLevel, (uint32_t)want);
    return hipGetLastError();
  }
  // Two levels: the kernel specialised for the first level's kernel sizes.
  const int kx = kernelSize(img.lv[inLevel].width), ky = kernelSize(img.lv[inLevel].height);
  if(kx == 3 && ky == 3)
    launchTwoLevel<F, 3, 3>(stream, img, inLevel);
  else if(kx == 2 && ky == 2)
    launchTwoLevel<F, 2, 2>(stream, img, inLevel);
  else if(kx == 3 && ky == 2)
    launchTwoLevel<F, 3, 2>(stream, img, inLevel);
  else if(kx == 2 && ky == 3)
    launchTwoLevel<F, 2, 3>(stream, img, inLevel);
  else
    launchTwoLevel<F, 0, 0>(stream, img, inLevel);
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

static TailPlan tailFrom(const NvproPyramidStep* steps, uint32_t first, uint32_t count)
{
  TailPlan plan{};
  for(uint32_t k = first; k < count && plan.count < NVPRO_PYRAMID_MAX_LEVELS; ++k)
    plan.steps[plan.count++] = TailStep{steps[k].pipeline, steps[k].inputLevel, steps[k].levelCount, 0u};
  return plan;
}

static uint64_t firstOutput(const ImageDesc& img, const NvproPyramidStep& s)
{
  const LevelDesc& first = img.lv[s.inputLevel + 1u];
  return (uint64_t)first.width * first.height;
}

// Persistent workgroups per compute unit of a fast step of `groups` groups
// (measured on MI355X, sRGBA8 F(0,6), per-step time): up to 16 groups per
// compute unit, 8 (the most that fit: 4096^2 22.3 us; with 4 or 3, 22.6-22.8);
// fewer workgroups on larger steps, although each then has fewer waves to
// hide latency behind: 5120^2 (25 groups per unit) 33.2 -> 32.6 us with 4,
// 6144^2 (36) 47.1 -> 45.2 with 3 (44.8 with 4), 8192x4096 (32) 43.0 -> 39.8 with 3, 8192^2
// (64) 86.8 -> 76.7 with 3, 16384^2 (256) 350.1 -> 296.0 with 3 (with 2:
// 344.3).
static uint32_t fastWorkgroupsPerCU(uint64_t groups)
{
  const uint64_t cus = deviceComputeUnits();
  return groups >= 32u * cus ? 3u : groups > 16u * cus ? 4u : 8u;
}

static uint32_t fastWorkgroupsPerCU(const ImageDesc& img, const NvproPyramidStep& s)
{
  const LevelDesc& in = img.lv[s.inputLevel];
  return fastWorkgroupsPerCU(fastGroupCount(in.width, in.height, s.levelCount));
}

// Whether fusing a successor into fast step s pays, by the step's workgroups
// (measured on MI355X, sRGBA8): a small grid counts at one arrival counter and
// gains the successor's launch and start-up (4 workgroups with a one-texel
// tail: 8.7 -> 7.2 us); a large grid also skips waiting for its own end of
// kernel, the write-back of everything it stored (2048 workgroups of 4096^2
// with F(6,6): 28.1 -> 26.5 us per chain); in between, two arrival counts and
// the continuation cost what they save (256 workgroups of 1024^2 with F(6,4):
// 8.7 -> 9.2 us).
static bool fusionPays(const ImageDesc& img, const NvproPyramidStep& s)
{
  const LevelDesc& in   = img.lv[s.inputLevel];
  const uint64_t groups = fastGroupCount(in.width, in.height, s.levelCount);
  const uint64_t cap    = (uint64_t)fastWorkgroupsPerCU(groups) * deviceComputeUnits();
  const uint64_t blocks = groups < cap ? groups : cap;
  return blocks <= kArriveLeaves || blocks >= 4ull * deviceComputeUnits();
}

hipError_t executePlan(hipStream_t stream, const ImageDesc& img, const NvproPyramidStep* steps, uint32_t count,
                       uint32_t* arrive, uint32_t* launches)
{
  *launches = 0;
  for(uint32_t i = 0; i < count; ++i, ++*launches)
  {
    const NvproPyramidStep& s = steps[i];
    // The one-workgroup tail pays off when it saves launches: with a single step
    // left, that step's own kernel is as quick.
    if(firstOutput(img, s) <= kTailMaxTexels && count - i >= 2u)
    {
      ++*launches;
      return launchTail(stream, img, tailFrom(steps, i, count));
    }
    const bool fast = s.pipeline == NVPRO_PYRAMID_PIPELINE_FAST;
    if(arrive && i + 1u < count && fast && fusionPays(img, s))
    {
      // Fusion: the successor runs in this launch's last workgroup, saving a
      // launch and the second kernel's start-up (tables, first loads).
      const NvproPyramidStep& n = steps[i + 1u];
      const LevelDesc& nin      = img.lv[n.inputLevel];
      if(n.pipeline == NVPRO_PYRAMID_PIPELINE_FAST && fastGroupCount(nin.width, nin.height, n.levelCount) == 1u)
      {
        const uint32_t tilesX = nin.width >> n.levelCount, tilesY = nin.height >> n.levelCount;
        const NextStep next{arrive, kNextFast, n.inputLevel, n.levelCount, tilesX, tilesX * tilesY};
        if(hipError_t e = launchFast(stream, img, s.inputLevel, s.levelCount, fastWorkgroupsPerCU(img, s), &next))
          return e;
        ++i;
        continue;
      }
      // A tail small enough for the fast kernel's scratch.  (A general step
      // followed by a tail measured no faster fused: the tail's first level
      // then takes several passes of a 256-thread workgroup, where the tail
      // kernel takes one pass of 1024 threads.)
      if(firstOutput(img, n) <= kFastTailTexels)
      {
        const TailPlan tail = tailFrom(steps, i + 1u, count);
        const NextStep next{arrive, kNextTail, 0u, 0u, 0u, 0u};
        ++*launches;
        return launchFast(stream, img, s.inputLevel, s.levelCount, fastWorkgroupsPerCU(img, s), &next, &tail);
      }
    }
    hipError_t e = fast ? launchFast(stream, img, s.inputLevel, s.levelCount, fastWorkgroupsPerCU(img, s))
                        : launchGeneral(stream, img, s.inputLevel, s.levelCount);
    if(e != hipSuccess)
      return e;
  }
  return hipSuccess;
}

}  // namespace nvpro_pyramid
