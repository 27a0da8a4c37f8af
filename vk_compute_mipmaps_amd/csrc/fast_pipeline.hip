// Fast pipeline (nvpro_pyramid.glsl:209-534) for gfx950, wave64.
//
// One dispatch fills COUNT (1..6) levels from an input level whose edges are
// divisible by 2^COUNT.  The input is cut into 2^COUNT x 2^COUNT tiles.
//
// Each lane owns a 4x4 input block (2x2 when COUNT is 1), loads it with one
// 16-byte load per row and reduces it in registers to levels 1 and 2.  A
// 256-thread workgroup thus covers 4096 input texels: one 64x64 tile for COUNT
// 6, four 32x32 tiles for COUNT 5, and so on.  Lanes are numbered in Morton
// order within a tile, so the level-2 values of a 2x2 are the values of four
// consecutive workgroup threads.  Levels 3..COUNT are reduced by one "duty" wave
// per iteration (rotating over the workgroup's waves): the level-2 values go
// through a 4 KiB LDS tile, duty lane k reduces threads 4k..4k+3 to level 3
// (all 64 lanes busy), and the levels above are cross-lane reductions inside
// the duty wave (DPP row shifts, then ds_bpermute).  The other waves meanwhile
// load and reduce the next blocks.
//
// Workgroups are persistent (the LDS tables are filled once) and small enough
// in registers for eight waves per SIMD, so the hardware overlaps one wave's
// loads with another's arithmetic.  Per level the 2x2 summation order is the
// reference's (fastOrder), so results are bit-identical to the reference
// schedule's arithmetic.
#include "kernel_common.hpp"

namespace nvpro_pyramid {

constexpr int kFastBS = 256;

template <int COUNT>
struct FastGeom
{
  static constexpr int LB           = COUNT >= 2 ? 2 : 1;                 // lane block edge, log2
  static constexpr int TEAM         = COUNT >= 2 ? 2 * (COUNT - 2) : 0;   // threads per tile, log2 (<= 8)
  static constexpr int TILES_PER_WG = kFastBS >> TEAM;
  static constexpr bool UPPER       = COUNT >= 3;                         // levels 3.. exist
};

template <class F, int COUNT>
struct FastShared
{
  typename F::Tables tab;
  typename F::Val level2[FastGeom<COUNT>::UPPER ? kFastBS : 1];
};

// One workgroup iteration: tile group g of a fast step over `img`, levels
// inLevel+1 .. inLevel+COUNT.  `iter` picks the duty wave.  Every thread of the
// workgroup calls this (it contains barriers when COUNT >= 3).
template <class F, int COUNT>
__device__ __forceinline__ void fastGroup(FastShared<F, COUNT>& sh, const ImageDesc& img, uint32_t inLevel,
                                          uint32_t tilesX, uint32_t numTiles, uint32_t g, uint32_t iter)
{
  using G            = FastGeom<COUNT>;
  using Raw          = typename F::Raw;
  using V            = typename F::Val;
  constexpr int LB   = G::LB;
  constexpr int BLK  = 1 << LB;
  constexpr bool PRE = F::kPrescale;
  const auto& tab    = sh.tab;

  const uint32_t lane   = threadIdx.x & 63u;
  const uint32_t wave   = threadIdx.x >> 6;
  const uint32_t inTile = threadIdx.x & ((1u << G::TEAM) - 1u);
  const uint32_t px     = compactEven(inTile);
  const uint32_t py     = compactEven(inTile >> 1);
  const LevelDesc& in   = img.lv[inLevel];

  // Levels 1 and 2 of this thread's block.
  const uint32_t tile = g * G::TILES_PER_WG + (threadIdx.x >> G::TEAM);
  V v{};
  if(tile < numTiles)  // whole tiles
  {
    const uint32_t ty = tile / tilesX, tx = tile - ty * tilesX;
    const uint32_t x0 = (tx << COUNT) + (px << LB), y0 = (ty << COUNT) + (py << LB);
    Raw raw[BLK][BLK];
    const uint8_t* p = img.base + in.offset + (uint64_t)y0 * in.pitch + (uint64_t)x0 * F::kBytes;
#pragma unroll
    for(int r = 0; r < BLK; ++r)
      loadRow<F, BLK>(p + (uint64_t)r * in.pitch, raw[r]);

    // Level 1: LOAD_REDUCE4 as four loads (premultiplied by 0.25 when exact).
    V l1[BLK / 2][BLK / 2];
#pragma unroll
    for(int i = 0; i < BLK / 2; ++i)
    {
      Raw row[BLK / 2];
#pragma unroll
      for(int j = 0; j < BLK / 2; ++j)
      {
        l1[i][j] = box<F, PRE>(F::template load<PRE>(tab, raw[2 * i][2 * j]),
                               F::template load<PRE>(tab, raw[2 * i][2 * j + 1]),
                               F::template load<PRE>(tab, raw[2 * i + 1][2 * j]),
                               F::template load<PRE>(tab, raw[2 * i + 1][2 * j + 1]), fastOrder(COUNT, 1u));
        row[j] = F::store(tab, l1[i][j]);
      }
      storeRow<F, BLK / 2>(texelPtr(img, inLevel + 1u, x0 >> 1, (y0 >> 1) + i, F::kBytes), row);
    }
    v = l1[0][0];
    if constexpr(LB == 2)
    {
      v = box<F, false>(l1[0][0], l1[0][1], l1[1][0], l1[1][1], fastOrder(COUNT, 2u));
      storeTexel<F>(texelPtr(img, inLevel + 2u, x0 >> 2, y0 >> 2, F::kBytes), F::store(tab, v));
    }
  }
  if constexpr(G::UPPER)
  {
    __syncthreads();  // the previous duty wave has read level2[]
    sh.level2[threadIdx.x] = v;
    __syncthreads();
    if(wave == (iter & 3u))
    {
      // Duty lane k: level 3 from threads 4k..4k+3 (a Morton 2x2 of level 2).
      constexpr int TEAM3 = 2 * (COUNT - 3);  // duty lanes per tile, log2 (<= 6)
      const uint32_t t3   = g * G::TILES_PER_WG + (lane >> TEAM3);
      const uint32_t k3   = lane & ((1u << TEAM3) - 1u);
      const uint32_t qx   = compactEven(k3), qy = compactEven(k3 >> 1);
      const V* c          = &sh.level2[4u * lane];
      V u                 = box<F, false>(c[0], c[1], c[2], c[3], fastOrder(COUNT, 3u));
      if(t3 < numTiles)
      {
        const uint32_t ty = t3 / tilesX, tx = t3 - ty * tilesX;
        storeTexel<F>(texelPtr(img, inLevel + 3u, (tx << (COUNT - 3)) + qx, (ty << (COUNT - 3)) + qy, F::kBytes),
                      F::store(tab, u));
#pragma unroll
        for(int k = 0; k < 3; ++k)
        {
          const int r = 4 + k;
          if(r > COUNT)
            break;
          if(k == 0)
            u = crossLaneBox<F, 1>(u, fastOrder(COUNT, (uint32_t)r));
          else if(k == 1)
            u = crossLaneBox<F, 4>(u, fastOrder(COUNT, (uint32_t)r));
          else
            u = crossLaneBox<F, 16>(u, fastOrder(COUNT, (uint32_t)r));
          if((k3 & ((4u << (2 * k)) - 1u)) == 0u)
          {
            const uint32_t sft = (uint32_t)(r - 3);
            storeTexel<F>(texelPtr(img, inLevel + (uint32_t)r, (tx << (COUNT - r)) + (qx >> sft),
                                   (ty << (COUNT - r)) + (qy >> sft), F::kBytes),
                          F::store(tab, u));
          }
        }
      }
    }
  }
}

template <class F, int COUNT>
__global__ void __launch_bounds__(kFastBS) __attribute__((amdgpu_waves_per_eu(8, 8)))
fastPipelineKernel(ImageDesc img, uint32_t inLevel, uint32_t tilesX, uint32_t numTiles)
{
  __shared__ FastShared<F, COUNT> sh;
  F::template fill<kFastBS>(sh.tab, F::kPrescale ? 0.25f : 1.0f);
  __syncthreads();
  const uint32_t numGroups = (numTiles + FastGeom<COUNT>::TILES_PER_WG - 1u) / FastGeom<COUNT>::TILES_PER_WG;
  uint32_t iter            = 0;
  for(uint32_t g = blockIdx.x; g < numGroups; g += gridDim.x, ++iter)
    fastGroup<F, COUNT>(sh, img, inLevel, tilesX, numTiles, g, iter);
}

// Batched form: the tile groups of many images' fast steps with the same COUNT,
// concatenated.  segs[i] covers groups [segs[i].groupStart, segs[i+1].groupStart).
template <class F, int COUNT>
__global__ void __launch_bounds__(kFastBS) __attribute__((amdgpu_waves_per_eu(8, 8)))
fastPipelineBatchKernel(const ImageDesc* __restrict__ imgs, const BatchSeg* __restrict__ segs, uint32_t numSegs,
                        uint32_t totalGroups)
{
  __shared__ FastShared<F, COUNT> sh;
  F::template fill<kFastBS>(sh.tab, F::kPrescale ? 0.25f : 1.0f);
  __syncthreads();
  uint32_t iter = 0;
  for(uint32_t g = blockIdx.x; g < totalGroups; g += gridDim.x, ++iter)
  {
    const BatchSeg s = segs[findSegment(segs, numSegs, g)];
    fastGroup<F, COUNT>(sh, imgs[s.image], s.inputLevel, s.tilesX, s.numTiles, g - s.groupStart, iter);
  }
}

template <class F, int COUNT>
static void launchFastCount(hipStream_t stream, const ImageDesc& img, uint32_t inLevel, uint32_t wgPerCU)
{
  using G                 = FastGeom<COUNT>;
  const LevelDesc& in     = img.lv[inLevel];
  const uint32_t tilesX   = in.width >> COUNT, tilesY = in.height >> COUNT;
  const uint32_t numTiles = tilesX * tilesY;
  const uint64_t groups   = ((uint64_t)numTiles + G::TILES_PER_WG - 1) / G::TILES_PER_WG;
  const uint64_t cap      = (uint64_t)deviceComputeUnits() * wgPerCU;
  const uint32_t blocks   = (uint32_t)(groups < cap ? groups : cap);
  hipLaunchKernelGGL((fastPipelineKernel<F, COUNT>), dim3(blocks ? blocks : 1), dim3(kFastBS), 0, stream, img,
                     inLevel, tilesX, numTiles);
}

template <class F>
hipError_t launchFastTyped(hipStream_t stream, const ImageDesc& img, uint32_t inLevel, uint32_t count,
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

template <class F>
static hipError_t launchFastBatchTyped(hipStream_t stream, uint32_t count, const ImageDesc* imgs, const BatchSeg* segs,
                                       uint32_t numSegs, uint32_t totalGroups, uint32_t wgPerCU)
{
  const uint64_t cap    = (uint64_t)deviceComputeUnits() * wgPerCU;
  const uint32_t blocks = (uint32_t)(totalGroups < cap ? totalGroups : cap);
  switch(count)
  {
#define NVPRO_FAST_BATCH_CASE(C)                                                                                     \
  case C:                                                                                                            \
    hipLaunchKernelGGL((fastPipelineBatchKernel<F, C>), dim3(blocks ? blocks : 1), dim3(kFastBS), 0, stream, imgs,   \
                       segs, numSegs, totalGroups);                                                                  \
    break;
    NVPRO_FAST_BATCH_CASE(1)
    NVPRO_FAST_BATCH_CASE(2)
    NVPRO_FAST_BATCH_CASE(3)
    NVPRO_FAST_BATCH_CASE(4)
    NVPRO_FAST_BATCH_CASE(5)
    NVPRO_FAST_BATCH_CASE(6)
#undef NVPRO_FAST_BATCH_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launchFastBatch(hipStream_t stream, uint32_t format, uint32_t arithmetic, uint32_t count,
                           const ImageDesc* imgs, const BatchSeg* segs, uint32_t numSegs, uint32_t totalGroups,
                           uint32_t wgPerCU)
{
  switch(format)
  {
    case NVPRO_PYRAMID_FORMAT_SRGBA8:
      return launchFastBatchTyped<Srgba8>(stream, count, imgs, segs, numSegs, totalGroups, wgPerCU);
    case NVPRO_PYRAMID_FORMAT_RGBA16F:
      return arithmetic == NVPRO_PYRAMID_ARITH_PACKED_FP16
                 ? launchFastBatchTyped<Rgba16fPacked>(stream, count, imgs, segs, numSegs, totalGroups, wgPerCU)
                 : launchFastBatchTyped<Rgba16f>(stream, count, imgs, segs, numSegs, totalGroups, wgPerCU);
    case NVPRO_PYRAMID_FORMAT_RGBA32F:
      return launchFastBatchTyped<Rgba32f>(stream, count, imgs, segs, numSegs, totalGroups, wgPerCU);
    default:
      return hipErrorInvalidValue;
  }
}

uint32_t fastGroupsPerStep(uint32_t width, uint32_t height, uint32_t count)
{
  const uint32_t tiles = (width >> count) * (height >> count);
  uint32_t perWG       = 0;
  switch(count)
  {
    case 1: perWG = FastGeom<1>::TILES_PER_WG; break;
    case 2: perWG = FastGeom<2>::TILES_PER_WG; break;
    case 3: perWG = FastGeom<3>::TILES_PER_WG; break;
    case 4: perWG = FastGeom<4>::TILES_PER_WG; break;
    case 5: perWG = FastGeom<5>::TILES_PER_WG; break;
    default: perWG = FastGeom<6>::TILES_PER_WG; break;
  }
  return (tiles + perWG - 1u) / perWG;
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
