// Measurement kernel, not part of the mip-chain library: the byte traffic of a
// full chain without its arithmetic.  It reads level 0 once and writes every
// texel of every level once, with values that are NOT mip levels -- the
// analogue of the reference's deliberately incorrect `baseline` pipeline
// (extras/general_pipelines/baseline/general_pipeline_alternative.glsl:1-26),
// whose time the reference quotes as the speed of light of the chain
// (docs/strategy.md.html:837-847).  Built into its own library
// (libnvpro_pyramid_sol.so) for the benchmark.
//
// Layout: the packed MipmapStorage chain of 4-byte texels
// (include/mipmap_storage.hpp:53-76), width and height multiples of 4.  A
// thread owns a 4x4 block of level 0 (four 16-byte row loads, consecutive
// threads on consecutive blocks of a row, so a wave's loads of one row are 1 KiB
// contiguous), writes the 2x2 of level 1 it covers (two 8-byte stores), its
// texel of level 2, and for every higher level the texel at its own block
// coordinates while those lie inside the level: every texel of every level is
// written exactly once.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

struct SolLevels
{
  uint64_t offset[16];  // bytes
  uint32_t width[16];
  uint32_t height[16];
  uint32_t count;
};

__global__ void __launch_bounds__(256) speedOfLightKernel(uint8_t* base, SolLevels lv)
{
  const uint32_t bw = lv.width[0] >> 2, bh = lv.height[0] >> 2;
  const uint64_t t  = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= (uint64_t)bw * bh)
    return;
  const uint32_t by = (uint32_t)(t / bw), bx = (uint32_t)(t - (uint64_t)by * bw);
  const uint8_t* src = base + lv.offset[0] + ((uint64_t)(4u * by) * lv.width[0] + 4u * bx) * 4u;
  uint4 r[4];
#pragma unroll
  for(int i = 0; i < 4; ++i)
    r[i] = *reinterpret_cast<const uint4*>(src + (uint64_t)i * lv.width[0] * 4u);
  // Level 1: one word per 2x2, combining all four words so no load is dead.
  const uint32_t a = r[0].x ^ r[0].y ^ r[1].x ^ r[1].y, b = r[0].z ^ r[0].w ^ r[1].z ^ r[1].w;
  const uint32_t c = r[2].x ^ r[2].y ^ r[3].x ^ r[3].y, d = r[2].z ^ r[2].w ^ r[3].z ^ r[3].w;
  uint8_t* l1 = base + lv.offset[1] + ((uint64_t)(2u * by) * lv.width[1] + 2u * bx) * 4u;
  *reinterpret_cast<uint2*>(l1)                                  = make_uint2(a, b);
  *reinterpret_cast<uint2*>(l1 + (uint64_t)lv.width[1] * 4u)     = make_uint2(c, d);
  const uint32_t v = a ^ b ^ c ^ d;
  for(uint32_t l = 2; l < lv.count; ++l)
  {
    if(bx >= lv.width[l] || by >= lv.height[l])
      return;
    *reinterpret_cast<uint32_t*>(base + lv.offset[l] + ((uint64_t)by * lv.width[l] + bx) * 4u) = v;
  }
}

}  // namespace

// Speed-of-light pass over a packed 4-byte-texel chain of `levels` levels
// (0: the full chain) at `base`; width and height multiples of 4.  Returns a
// hipError_t (hipErrorInvalidValue for an unsupported shape).
extern "C" int nvproPyramidSpeedOfLight(hipStream_t stream, void* base, uint32_t width, uint32_t height,
                                        uint32_t levels)
{
  if(!base || width < 4u || height < 4u || (width | height) % 4u)
    return (int)hipErrorInvalidValue;
  uint32_t full = 1;
  for(uint32_t m = width > height ? width : height; m > 1u; m >>= 1)
    ++full;
  if(levels == 0u)
    levels = full;
  if(levels < 3u || levels > full || levels > 16u)
    return (int)hipErrorInvalidValue;
  SolLevels lv{};
  lv.count    = levels;
  uint64_t at = 0;
  for(uint32_t l = 0; l < levels; ++l)
  {
    lv.width[l]  = width >> l ? width >> l : 1u;
    lv.height[l] = height >> l ? height >> l : 1u;
    lv.offset[l] = at;
    at += (uint64_t)lv.width[l] * lv.height[l] * 4u;
  }
  const uint64_t threads = (uint64_t)(width / 4u) * (height / 4u);
  const uint32_t blocks  = (uint32_t)((threads + 255u) / 256u);
  hipLaunchKernelGGL(speedOfLightKernel, dim3(blocks), dim3(256), 0, stream, static_cast<uint8_t*>(base), lv);
  return (int)hipGetLastError();
}
