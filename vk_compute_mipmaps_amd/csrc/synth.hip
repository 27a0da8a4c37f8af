// Seeded synthetic level 0 on the device: the workload generator of the
// benchmark and serving inputs, bit-identical to vk_compute_mipmaps_amd/synth.py
// (same counter-based hash, same integer value noise; RGBA16F radiance through
// FP64 exp, rounded to half).  Every texel is independent, so one thread per
// texel, grid-strided.
#include <hip/hip_fp16.h>

#include "pyramid_internal.hpp"

namespace nvpro_pyramid {

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// h(seed, s, i, j) = mix(mix(mix(seed ^ (s << 56)) ^ i) ^ j)
__device__ __forceinline__ uint64_t hash4(uint64_t seedMix, uint64_t i, uint64_t j)
{
  return mix64(mix64(seedMix ^ i) ^ j);
}

__global__ void __launch_bounds__(256) synthSrgba8Kernel(uint8_t* base, uint32_t width, uint32_t height,
                                                         uint32_t pitch, uint64_t seed)
{
  uint64_t sm[8];
#pragma unroll
  for(int s = 0; s < 8; ++s)
    sm[s] = mix64(seed ^ ((uint64_t)s << 56));
  const uint64_t n      = (uint64_t)width * height;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
  {
    const uint32_t y = (uint32_t)(t / width), x = (uint32_t)(t - (uint64_t)y * width);
    const uint64_t i = x >> 6, j = y >> 6, mx = x & 63u, my = y & 63u;
    const uint64_t D  = 1ull << 18;
    const uint64_t sx = mx * mx * (192u - 2u * mx), sy = my * my * (192u - 2u * my);
    uint32_t texel = 0;
#pragma unroll
    for(int c = 0; c < 3; ++c)
    {
      const uint64_t L00 = hash4(sm[c], i, j) >> 56, L10 = hash4(sm[c], i + 1, j) >> 56;
      const uint64_t L01 = hash4(sm[c], i, j + 1) >> 56, L11 = hash4(sm[c], i + 1, j + 1) >> 56;
      const uint64_t top = L00 * (D - sx) + L10 * sx, bot = L01 * (D - sx) + L11 * sx;
      const uint64_t num = top * (D - sy) + bot * sy;
      const int64_t v    = (int64_t)((num + (1ull << 35)) >> 36);
      const int64_t nz   = (int64_t)((hash4(sm[4 + c], x, y) >> 32) % 33u) - 16;
      int64_t q          = v + nz;
      q                  = q < 0 ? 0 : (q > 255 ? 255 : q);
      texel |= (uint32_t)q << (8 * c);
    }
    texel |= (uint32_t)(hash4(sm[7], x, y) >> 56) << 24;
    *reinterpret_cast<uint32_t*>(base + (uint64_t)y * pitch + (uint64_t)x * 4u) = texel;
  }
}

// Round-to-nearest-even straight from binary64 (no detour through binary32).
__device__ __forceinline__ uint16_t f64ToHalfBits(double v)
{
  return __builtin_bit_cast(uint16_t, (_Float16)v);
}

__global__ void __launch_bounds__(256) synthRgba16fKernel(uint8_t* base, uint32_t width, uint32_t height,
                                                          uint32_t pitch, uint64_t seed)
{
  uint64_t sm[4];
#pragma unroll
  for(int s = 0; s < 3; ++s)
    sm[s] = mix64(seed ^ ((uint64_t)s << 56));
  sm[3]                 = mix64(seed ^ (7ull << 56));
  const double lo       = -6.907755278982137;  // ln 1e-3
  const double span     = 13.815510557964274;  // ln 1e3 - ln 1e-3
  const double scale    = 0x1p-53;
  const uint64_t n      = (uint64_t)width * height;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
  {
    const uint32_t y = (uint32_t)(t / width), x = (uint32_t)(t - (uint64_t)y * width);
    uint16_t h[4];
#pragma unroll
    for(int c = 0; c < 3; ++c)
    {
      const double u = (double)(hash4(sm[c], x, y) >> 11) * scale;
      h[c]           = f64ToHalfBits(exp(lo + u * span));
    }
    h[3] = f64ToHalfBits((double)(hash4(sm[3], x, y) >> 11) * scale);
    uint2 v = make_uint2(h[0] | (uint32_t)h[1] << 16, h[2] | (uint32_t)h[3] << 16);
    *reinterpret_cast<uint2*>(base + (uint64_t)y * pitch + (uint64_t)x * 8u) = v;
  }
}

}  // namespace nvpro_pyramid

using namespace nvpro_pyramid;

extern "C" int nvproPyramidSynthesize(hipStream_t stream, uint32_t format, const NvproPyramidImage* image,
                                      uint64_t seed)
{
  const char* who = "nvproPyramidSynthesize";
  if(!image || !image->data)
    return hostFail(hipErrorInvalidValue, "%s: null image", who);
  if(format != NVPRO_PYRAMID_FORMAT_SRGBA8 && format != NVPRO_PYRAMID_FORMAT_RGBA16F)
    return hostFail(hipErrorInvalidValue, "%s: format %u has no generator", who, format);
  if(image->width == 0 || image->height == 0)
    return hostFail(hipErrorInvalidValue, "%s: zero image size", who);
  const uint32_t bpt   = nvproPyramidBytesPerTexel(format);
  const uint64_t pitch = image->rowPitches ? image->rowPitches[0] : (uint64_t)image->width * bpt;
  const uint64_t off   = image->levelOffsets ? image->levelOffsets[0] : 0u;
  if(pitch < (uint64_t)image->width * bpt || pitch > 0xffffffffull || (((uintptr_t)image->data + off) % bpt) != 0 ||
     pitch % bpt != 0)
    return hostFail(hipErrorInvalidValue, "%s: level 0 layout not aligned to %u bytes", who, bpt);
  uint8_t* base       = (uint8_t*)image->data + off;
  const uint64_t n    = (uint64_t)image->width * image->height;
  const uint64_t want = (n + 255u) / 256u;
  const uint32_t cap  = deviceComputeUnits() * 8u;
  const uint32_t grid = (uint32_t)(want < cap ? want : cap);
  if(format == NVPRO_PYRAMID_FORMAT_SRGBA8)
    hipLaunchKernelGGL(synthSrgba8Kernel, dim3(grid), dim3(256), 0, stream, base, image->width, image->height,
                       (uint32_t)pitch, seed);
  else
    hipLaunchKernelGGL(synthRgba16fKernel, dim3(grid), dim3(256), 0, stream, base, image->width, image->height,
                       (uint32_t)pitch, seed);
  hipError_t e = hipGetLastError();
  if(e != hipSuccess)
    return hostFail(e, "%s: launch failed: %s", who, hipGetErrorString(e));
  return 0;
}
