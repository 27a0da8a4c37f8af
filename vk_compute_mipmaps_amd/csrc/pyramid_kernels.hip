// HIP kernels of the nvpro_pyramid schedule for gfx950 (wave64).
//
// Fast pipeline (nvpro_pyramid.glsl:209-534): fills `count` (1..6) levels from an
// input level whose edges are divisible by 2^count.  The input is cut into
// 2^count x 2^count tiles.  Each lane loads a BLK x BLK block (BLK = 2^LB) with
// 16-byte row loads and reduces it in registers through LB levels; a team of
// 4^(count-LB) lanes (at most one wave) then finishes the tile with cross-lane
// exchanges, lanes ordered in Morton order so the four quadrants of every 2x2
// are lanes l, l^m, l^2m, l^3m.  No LDS traffic or barrier is needed between
// levels.  Summation order per level follows the reference (fastOrder).
//
// General pipeline (nvpro_pyramid.glsl:536-884): 1 or 2 levels of any size with
// the 2x2 / 2x3 / 3x2 / 3x3 energy-conserving kernel.  Two levels: a workgroup
// owns an 8x8 tile of the second level, computes the (up to) 17x17 first-level
// tile it needs into LDS (writing it out as well), then the 8x8 tile.
#include "functors.hpp"
#include "pyramid_internal.hpp"

namespace nvpro_pyramid {

template <int NW>
struct __attribute__((aligned(4))) Words
{
  uint32_t w[NW];
};

// Load / store NBYTES (a multiple of 4) at a 4-byte aligned address using the
// widest dword-aligned vector accesses (global_load_dwordx4 for 16 bytes).
template <int NBYTES>
__device__ __forceinline__ void loadWords(const uint8_t* p, uint32_t* out)
{
  constexpr int NW = NBYTES / 4;
  if constexpr(NW >= 4)
  {
#pragma unroll
    for(int i = 0; i < NW / 4; ++i)
    {
      Words<4> v = reinterpret_cast<const Words<4>*>(p)[i];
#pragma unroll
      for(int k = 0; k < 4; ++k)
        out[4 * i + k] = v.w[k];
    }
  }
  else
  {
    Words<NW> v = *reinterpret_cast<const Words<NW>*>(p);
#pragma unroll
    for(int k = 0; k < NW; ++k)
      out[k] = v.w[k];
  }
}

template <int NBYTES>
__device__ __forceinline__ void storeWords(uint8_t* p, const uint32_t* in)
{
  constexpr int NW = NBYTES / 4;
  if constexpr(NW >= 4)
  {
#pragma unroll
    for(int i = 0; i < NW / 4; ++i)
    {
      Words<4> v;
#pragma unroll
      for(int k = 0; k < 4; ++k)
        v.w[k] = in[4 * i + k];
      reinterpret_cast<Words<4>*>(p)[i] = v;
    }
  }
  else
  {
    Words<NW> v;
#pragma unroll
    for(int k = 0; k < NW; ++k)
      v.w[k] = in[k];
    *reinterpret_cast<Words<NW>*>(p) = v;
  }
}

template <class F, int N>
__device__ __forceinline__ void loadRow(const uint8_t* p, typename F::Raw (&out)[N])
{
  static_assert(sizeof(typename F::Raw) == F::kBytes, "raw texel size");
  uint32_t w[N * F::kBytes / 4];
  loadWords<N * F::kBytes>(p, w);
  __builtin_memcpy(out, w, sizeof w);
}

template <class F, int N>
__device__ __forceinline__ void storeRow(uint8_t* p, const typename F::Raw (&in)[N])
{
  uint32_t w[N * F::kBytes / 4];
  __builtin_memcpy(w, in, sizeof w);
  storeWords<N * F::kBytes>(p, w);
}

__device__ __forceinline__ uint8_t* texelPtr(const ImageDesc& img, uint32_t level, uint32_t x, uint32_t y,
                                             uint32_t bytes)
{
  const LevelDesc& d = img.lv[level];
  return img.base + d.offset + (uint64_t)y * d.pitch + (uint64_t)x * bytes;
}

// Even bits of v packed together (Morton decode of one coordinate).
__device__ __forceinline__ uint32_t compactEven(uint32_t v)
{
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0f0f0f0fu;
  v = (v | (v >> 4)) & 0x00ff00ffu;
  v = (v | (v >> 8)) & 0x0000ffffu;
  return v;
}

__device__ __forceinline__ float4 shflXor(float4 v, int mask)
{
  return make_float4(__shfl_xor(v.x, mask, 64), __shfl_xor(v.y, mask, 64), __shfl_xor(v.z, mask, 64),
                     __shfl_xor(v.w, mask, 64));
}

// In-register reduction of an N x N grid of level-(r-1) values into
// (N/2) x (N/2) level-r values, stored as rows of level `dstLevel` at (dx, dy).
template <class F, bool PRE, int N>
__device__ __forceinline__ void reduceGrid(const typename F::Smem& smem, const float4 (&src)[N][N],
                                           float4 (&dst)[N / 2][N / 2], int order, const ImageDesc& img,
                                           uint32_t dstLevel, uint32_t dx, uint32_t dy)
{
#pragma unroll
  for(int i = 0; i < N / 2; ++i)
  {
    typename F::Raw row[N / 2];
#pragma unroll
    for(int j = 0; j < N / 2; ++j)
    {
      dst[i][j] = box<PRE>(src[2 * i][2 * j], src[2 * i][2 * j + 1], src[2 * i + 1][2 * j],
                           src[2 * i + 1][2 * j + 1], order);
      row[j]    = F::store(smem, dst[i][j]);
    }
    storeRow<F, N / 2>(texelPtr(img, dstLevel, dx, dy + i, F::kBytes), row);
  }
}

// ---------------------------------------------------------------------------
// Fast pipeline
// ---------------------------------------------------------------------------
template <class F, int LB>
__global__ void __launch_bounds__(256)
    fastPipelineKernel(ImageDesc img, uint32_t inLevel, uint32_t count, uint32_t tilesX, uint32_t numTiles)
{
  __shared__ typename F::Smem smem;
  F::fill(smem, 0.25f);
  __syncthreads();

  constexpr int BLK            = 1 << LB;
  const uint32_t teamLog2      = 2u * (count - (uint32_t)LB);  // lanes per team, log2 (0..6)
  const uint32_t unitLog2      = 6u - teamLog2;                // tiles per wave, log2
  const uint32_t lane          = threadIdx.x & 63u;
  const uint32_t laneInTeam    = lane & ((1u << teamLog2) - 1u);
  const uint32_t px            = compactEven(laneInTeam);
  const uint32_t py            = compactEven(laneInTeam >> 1);
  const uint32_t wavesPerBlock = blockDim.x >> 6;
  const uint32_t numUnits      = (numTiles + (1u << unitLog2) - 1u) >> unitLog2;
  const LevelDesc in           = img.lv[inLevel];

  for(uint32_t u = blockIdx.x * wavesPerBlock + (threadIdx.x >> 6); u < numUnits; u += gridDim.x * wavesPerBlock)
  {
    const uint32_t tile = (u << unitLog2) + (lane >> teamLog2);
    if(tile >= numTiles)
      continue;  // whole team
    const uint32_t tx = tile % tilesX, ty = tile / tilesX;
    const uint32_t x0 = (tx << count) + (px << LB), y0 = (ty << count) + (py << LB);

    typename F::Raw raw[BLK][BLK];
#pragma unroll
    for(int r = 0; r < BLK; ++r)
      loadRow<F, BLK>(img.base + in.offset + (uint64_t)(y0 + r) * in.pitch + (uint64_t)x0 * F::kBytes, raw[r]);

    // Level 1: LOAD_REDUCE4 as four loads, inputs premultiplied by 0.25.
    float4 l1[BLK / 2][BLK / 2];
    {
      const int order = fastOrder(count, 1u);
#pragma unroll
      for(int i = 0; i < BLK / 2; ++i)
      {
        typename F::Raw row[BLK / 2];
#pragma unroll
        for(int j = 0; j < BLK / 2; ++j)
        {
          l1[i][j] = box<true>(F::template load<true>(smem, raw[2 * i][2 * j]),
                               F::template load<true>(smem, raw[2 * i][2 * j + 1]),
                               F::template load<true>(smem, raw[2 * i + 1][2 * j]),
                               F::template load<true>(smem, raw[2 * i + 1][2 * j + 1]), order);
          row[j]   = F::store(smem, l1[i][j]);
        }
        storeRow<F, BLK / 2>(texelPtr(img, inLevel + 1u, x0 >> 1, (y0 >> 1) + i, F::kBytes), row);
      }
    }
    float4 v;
    if constexpr(LB == 1)
    {
      v = l1[0][0];
    }
    else if constexpr(LB == 2)
    {
      float4 l2[1][1];
      reduceGrid<F, false, 2>(smem, l1, l2, fastOrder(count, 2u), img, inLevel + 2u, x0 >> 2, y0 >> 2);
      v = l2[0][0];
    }
    else
    {
      float4 l2[2][2], l3[1][1];
      reduceGrid<F, false, 4>(smem, l1, l2, fastOrder(count, 2u), img, inLevel + 2u, x0 >> 2, y0 >> 2);
      reduceGrid<F, false, 2>(smem, l2, l3, fastOrder(count, 3u), img, inLevel + 3u, x0 >> 3, y0 >> 3);
      v = l3[0][0];
    }

    // Cross-lane levels: partners are lanes ^ (1 << 2k) (x) and ^ (2 << 2k) (y).
    for(uint32_t r = LB + 1u; r <= count; ++r)
    {
      const uint32_t k  = r - (uint32_t)LB - 1u;
      const int mx      = 1 << (2u * k);
      const int my      = 2 << (2u * k);
      const int order   = fastOrder(count, r);
      const int first   = order == 0 ? my : mx;
      const int second  = order == 0 ? mx : my;
      float4 s          = f4add(v, shflXor(v, first));
      v                 = f4mul(f4add(s, shflXor(s, second)), 0.25f);
      const uint32_t gm = (4u << (2u * k)) - 1u;
      if((laneInTeam & gm) == 0u)
      {
        const uint32_t sh = r - (uint32_t)LB;
        const uint32_t lx = (tx << (count - r)) + (px >> sh);
        const uint32_t ly = (ty << (count - r)) + (py >> sh);
        typename F::Raw t[1] = {F::store(smem, v)};
        storeRow<F, 1>(texelPtr(img, inLevel + r, lx, ly, F::kBytes), t);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// General pipeline
// ---------------------------------------------------------------------------
__device__ __forceinline__ int kernelSize(uint32_t dim) { return dim == 1u ? 1 : (int)(2u | (dim & 1u)); }

// reduceStoreSample_ (nvpro_pyramid.glsl:575-656) without the store: the sample
// of destination texel (dx, dy) of a (dw x dh) level from a kx x ky window whose
// texels `at(j, i)` returns (column j, row i).
template <class At>
__device__ __forceinline__ float4 generalSample(At at, int kx, int ky, uint32_t dw, uint32_t dh, uint32_t dx,
                                                uint32_t dy)
{
  float n   = (float)dh;
  float rcp = 1.0f / (2.0f * n + 1.0f);
  float w0  = rcp * (n - (float)dy);
  float w1  = rcp * n;
  float w2  = (1.0f - w0) - w1;
  float4 h[3];
#pragma unroll
  for(int j = 0; j < 3; ++j)
  {
    if(j < kx)
    {
      if(ky == 3)
        h[j] = reduce3(w0, at(j, 0), w1, at(j, 1), w2, at(j, 2));
      else if(ky == 2)
        h[j] = reduce2(at(j, 0), at(j, 1));
      else
        h[j] = at(j, 0);
    }
  }
  if(kx == 3)
  {
    n   = (float)dw;
    rcp = 1.0f / (2.0f * n + 1.0f);
    w0  = rcp * (n - (float)dx);
    w1  = rcp * n;
    w2  = (1.0f - w0) - w1;
    return reduce3(w0, h[0], w1, h[1], w2, h[2]);
  }
  if(kx == 2)
    return reduce2(h[0], h[1]);
  return h[0];
}

template <class F>
__global__ void __launch_bounds__(256) generalOneLevelKernel(ImageDesc img, uint32_t inLevel)
{
  __shared__ typename F::Smem smem;
  F::fill(smem, 1.0f);
  __syncthreads();
  const LevelDesc src = img.lv[inLevel];
  const LevelDesc dst = img.lv[inLevel + 1u];
  const int kx = kernelSize(src.width), ky = kernelSize(src.height);
  const uint32_t total = dst.width * dst.height;
  for(uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x)
  {
    const uint32_t dy = i / dst.width, dx = i - dy * dst.width;
    const uint8_t* p  = img.base + src.offset + (uint64_t)(2u * dy) * src.pitch + (uint64_t)(2u * dx) * F::kBytes;
    auto at           = [&](int j, int r) {
      typename F::Raw t[1];
      loadRow<F, 1>(p + (uint64_t)r * src.pitch + (uint64_t)j * F::kBytes, t);
      return F::template load<false>(smem, t[0]);
    };
    typename F::Raw o[1] = {F::store(smem, generalSample(at, kx, ky, dst.width, dst.height, dx, dy))};
    storeRow<F, 1>(texelPtr(img, inLevel + 1u, dx, dy, F::kBytes), o);
  }
}

template <class F>
__global__ void __launch_bounds__(256) generalTwoLevelKernel(ImageDesc img, uint32_t inLevel, uint32_t tilesX,
                                                            uint32_t numTiles)
{
  __shared__ typename F::Smem smem;
  __shared__ float4 mid[17][17];  // sharedLevel_ (nvpro_pyramid.glsl:555), [y][x]
  F::fill(smem, 1.0f);
  __syncthreads();
  const LevelDesc src = img.lv[inLevel];
  const LevelDesc l1  = img.lv[inLevel + 1u];
  const LevelDesc l2  = img.lv[inLevel + 2u];
  const int kx1 = kernelSize(src.width), ky1 = kernelSize(src.height);
  const int kx2 = kernelSize(l1.width), ky2 = kernelSize(l1.height);
  const uint32_t nx = kx2 == 3 ? 17u : 16u, ny = ky2 == 3 ? 17u : 16u;

  for(uint32_t tile = blockIdx.x; tile < numTiles; tile += gridDim.x)
  {
    const uint32_t ty = tile / tilesX, tx = tile - ty * tilesX;
    // First level: the tile of level inLevel+1 the 8x8 output tile needs
    // (fillIntermediateTile_, nvpro_pyramid.glsl:732-781), with bounds checks.
    for(uint32_t e = threadIdx.x; e < nx * ny; e += blockDim.x)
    {
      const uint32_t ly = e / nx, lx = e - ly * nx;
      const uint32_t gx = 16u * tx + lx, gy = 16u * ty + ly;
      if(gx < l1.width && gy < l1.height)
      {
        const uint8_t* p = img.base + src.offset + (uint64_t)(2u * gy) * src.pitch + (uint64_t)(2u * gx) * F::kBytes;
        auto at          = [&](int j, int r) {
          typename F::Raw t[1];
          loadRow<F, 1>(p + (uint64_t)r * src.pitch + (uint64_t)j * F::kBytes, t);
          return F::template load<false>(smem, t[0]);
        };
        float4 v             = generalSample(at, kx1, ky1, l1.width, l1.height, gx, gy);
        typename F::Raw o[1] = {F::store(smem, v)};
        storeRow<F, 1>(texelPtr(img, inLevel + 1u, gx, gy, F::kBytes), o);
        mid[ly][lx] = v;
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
        auto at              = [&](int j, int r) { return mid[2u * oy + r][2u * ox + j]; };
        typename F::Raw o[1] = {F::store(smem, generalSample(at, kx2, ky2, l2.width, l2.height, gx, gy))};
        storeRow<F, 1>(texelPtr(img, inLevel + 2u, gx, gy, F::kBytes), o);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static uint32_t capBlocks(uint64_t wanted)
{
  const uint64_t cap = (uint64_t)deviceComputeUnits() * 8u;
  if(wanted == 0)
    return 1;
  return (uint32_t)(wanted < cap ? wanted : cap);
}

template <class F>
static hipError_t launchTyped(hipStream_t stream, const ImageDesc& img, uint32_t pipeline, uint32_t inLevel,
                              uint32_t count)
{
  const LevelDesc& in = img.lv[inLevel];
  if(pipeline == NVPRO_PYRAMID_PIPELINE_FAST)
  {
    const uint32_t tilesX = in.width >> count, tilesY = in.height >> count;
    const uint32_t numTiles = tilesX * tilesY;
    const int lb            = count >= 6u ? 3 : (count >= 2u ? 2 : 1);
    const uint32_t teamLog2 = 2u * (count - (uint32_t)lb);
    const uint64_t units    = ((uint64_t)numTiles + (1u << (6u - teamLog2)) - 1u) >> (6u - teamLog2);
    const uint32_t blocks   = capBlocks((units + 3u) / 4u);
    if(lb == 1)
      hipLaunchKernelGGL((fastPipelineKernel<F, 1>), dim3(blocks), dim3(256), 0, stream, img, inLevel, count,
                         tilesX, numTiles);
    else if(lb == 2)
      hipLaunchKernelGGL((fastPipelineKernel<F, 2>), dim3(blocks), dim3(256), 0, stream, img, inLevel, count,
                         tilesX, numTiles);
    else
      hipLaunchKernelGGL((fastPipelineKernel<F, 3>), dim3(blocks), dim3(256), 0, stream, img, inLevel, count,
                         tilesX, numTiles);
  }
  else if(count == 1u)
  {
    const LevelDesc& d = img.lv[inLevel + 1u];
    const uint32_t blocks = capBlocks(((uint64_t)d.width * d.height + 255u) / 256u);
    hipLaunchKernelGGL((generalOneLevelKernel<F>), dim3(blocks), dim3(256), 0, stream, img, inLevel);
  }
  else
  {
    const LevelDesc& d = img.lv[inLevel + 2u];
    const uint32_t tilesX = (d.width + 7u) / 8u, tilesY = (d.height + 7u) / 8u;
    const uint32_t blocks = capBlocks((uint64_t)tilesX * tilesY);
    hipLaunchKernelGGL((generalTwoLevelKernel<F>), dim3(blocks), dim3(256), 0, stream, img, inLevel, tilesX,
                       tilesX * tilesY);
  }
  return hipGetLastError();
}

hipError_t launchStep(hipStream_t stream, const ImageDesc& img, uint32_t pipeline, uint32_t inputLevel,
                      uint32_t levelCount)
{
  switch(img.format)
  {
    case NVPRO_PYRAMID_FORMAT_SRGBA8:
      return launchTyped<Srgba8>(stream, img, pipeline, inputLevel, levelCount);
    case NVPRO_PYRAMID_FORMAT_RGBA16F:
      return launchTyped<Rgba16f>(stream, img, pipeline, inputLevel, levelCount);
    case NVPRO_PYRAMID_FORMAT_RGBA32F:
      return launchTyped<Rgba32f>(stream, img, pipeline, inputLevel, levelCount);
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

}  // namespace nvpro_pyramid
