// Seeded synthetic level 0 on the device: the workload generator of the
// benchmark and serving inputs, bit-identical to vk_compute_mipmaps_amd/synth.py
// (same counter-based hash, same integer value noise; RGBA16F radiance through
// FP64 exp, rounded to half).  Every texel is independent: a workgroup writes
// chunks of kSynthChunk consecutive texels, one thread per texel; a batched
// launch covers the chunks of many images.
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

// Texel (x, y) of the seeded sRGBA8 image; sm[s] = mix64(seed ^ (s << 56)).
__device__ __forceinline__ uint32_t synthSrgba8Texel(const uint64_t (&sm)[8], uint32_t x, uint32_t y)
{
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
    const int64_t nz   = (int64_t)((uint32_t)(hash4(sm[4 + c], x, y) >> 32) % 33u) - 16;  // 32-bit remainder
    int64_t q          = v + nz;
    q                  = q < 0 ? 0 : (q > 255 ? 255 : q);
    texel |= (uint32_t)q << (8 * c);
  }
  return texel | (uint32_t)(hash4(sm[7], x, y) >> 56) << 24;
}

// Round-to-nearest-even straight from binary64 (no detour through binary32).
__device__ __forceinline__ uint16_t f64ToHalfBits(double v)
{
  return __builtin_bit_cast(uint16_t, (_Float16)v);
}

// Texel (x, y) of the seeded RGBA16F image; sm[0..2] the RGB streams, sm[7] alpha.
__device__ __forceinline__ uint2 synthRgba16fTexel(const uint64_t (&sm)[8], uint32_t x, uint32_t y)
{
  const double lo    = -6.907755278982137;  // ln 1e-3
  const double span  = 13.815510557964274;  // ln 1e3 - ln 1e-3
  const double scale = 0x1p-53;
  uint16_t h[4];
#pragma unroll
  for(int c = 0; c < 3; ++c)
  {
    const double u = (double)(hash4(sm[c], x, y) >> 11) * scale;
    h[c]           = f64ToHalfBits(exp(lo + u * span));
  }
  h[3] = f64ToHalfBits((double)(hash4(sm[7], x, y) >> 11) * scale);
  return make_uint2(h[0] | (uint32_t)h[1] << 16, h[2] | (uint32_t)h[3] << 16);
}

__device__ __forceinline__ void seedStreams(uint64_t seed, uint64_t (&sm)[8])
{
#pragma unroll
  for(int s = 0; s < 8; ++s)
    sm[s] = mix64(seed ^ ((uint64_t)s << 56));
}

// Lattice values of the sRGBA8 value noise a chunk needs, per channel: cell
// columns 0 .. ni-1 and rows j0 .. j0+nj-1 (one byte each).  Every texel of a
// 64x64 cell uses the same four corners, so a chunk computes them once.
constexpr uint32_t kLatticeMax = 4096;  // per channel; a chunk needs at most 3075 (3 cell rows x 1025 columns)
struct LatticeCache
{
  uint8_t v[3][kLatticeMax];
};

// Texel (x, y) of the seeded sRGBA8 image with the lattice from `lat`.
__device__ __forceinline__ uint32_t synthSrgba8TexelCached(const uint64_t (&sm)[8], const LatticeCache& lat,
                                                           uint32_t ni, uint32_t j0, uint32_t x, uint32_t y)
{
  const uint64_t i = x >> 6, j = (y >> 6) - j0, mx = x & 63u, my = y & 63u;
  const uint64_t D  = 1ull << 18;
  const uint64_t sx = mx * mx * (192u - 2u * mx), sy = my * my * (192u - 2u * my);
  const uint32_t e  = (uint32_t)(j * ni + i);
  uint32_t texel    = 0;
#pragma unroll
  for(int c = 0; c < 3; ++c)
  {
    const uint64_t L00 = lat.v[c][e], L10 = lat.v[c][e + 1u], L01 = lat.v[c][e + ni], L11 = lat.v[c][e + ni + 1u];
    const uint64_t top = L00 * (D - sx) + L10 * sx, bot = L01 * (D - sx) + L11 * sx;
    const uint64_t num = top * (D - sy) + bot * sy;
    const int64_t v    = (int64_t)((num + (1ull << 35)) >> 36);
    const int64_t nz   = (int64_t)((uint32_t)(hash4(sm[4 + c], x, y) >> 32) % 33u) - 16;  // 32-bit remainder
    int64_t q          = v + nz;
    q                  = q < 0 ? 0 : (q > 255 ? 255 : q);
    texel |= (uint32_t)q << (8 * c);
  }
  return texel | (uint32_t)(hash4(sm[7], x, y) >> 56) << 24;
}

// Row-major texel t = y * w + x, walked by a workgroup: each thread starts at
// first + threadIdx.x and steps by blockDim.x, with (x, y) kept incrementally
// (one division per thread and chunk; a step adds q rows and r columns, where
// blockDim.x = q * w + r).  A level 0 has fewer than 2^32 texels.
struct TexelWalk
{
  uint32_t x, y, q, r, w;
  __device__ TexelWalk(uint32_t first, uint32_t width)
  {
    const uint32_t t = first + threadIdx.x;
    w                = width;
    y                = t / width;
    x                = t - y * width;
    q                = blockDim.x / width;
    r                = blockDim.x - q * width;
  }
  __device__ void step()
  {
    x += r;
    y += q;
    if(x >= w)
    {
      x -= w;
      ++y;
    }
  }
};

// Texels [first, end) of `job` by the whole workgroup (uniform arguments;
// contains barriers).
__device__ __forceinline__ void synthTexels(LatticeCache& lat, const SynthJob& job, uint64_t first, uint64_t end,
                                            const uint64_t (&sm)[8])
{
  const uint32_t w = job.width;
  if(job.format == NVPRO_PYRAMID_FORMAT_SRGBA8)
  {
    const uint32_t y0 = (uint32_t)(first / w), y1 = (uint32_t)((end - 1u) / w);
    const uint32_t j0 = y0 >> 6, nj = (y1 >> 6) - j0 + 2u, ni = ((w - 1u) >> 6) + 2u;
    if(ni * nj <= kLatticeMax)
    {
      __syncthreads();  // the previous chunk's texels have read the cache
      for(uint32_t e = threadIdx.x; e < 3u * ni * nj; e += blockDim.x)
      {
        const uint32_t c = e / (ni * nj), k = e - c * ni * nj;
        const uint32_t jj = k / ni, ii = k - jj * ni;
        lat.v[c][k]       = (uint8_t)(hash4(sm[c], ii, j0 + jj) >> 56);
      }
      __syncthreads();
      TexelWalk p((uint32_t)first, w);
      for(uint64_t t = first + threadIdx.x; t < end; t += blockDim.x, p.step())
        *reinterpret_cast<uint32_t*>(job.base + (uint64_t)p.y * job.pitch + (uint64_t)p.x * 4u) =
            synthSrgba8TexelCached(sm, lat, ni, j0, p.x, p.y);
      return;
    }
  }
  TexelWalk p((uint32_t)first, w);
  for(uint64_t t = first + threadIdx.x; t < end; t += blockDim.x, p.step())
  {
    uint8_t* row = job.base + (uint64_t)p.y * job.pitch;
    if(job.format == NVPRO_PYRAMID_FORMAT_SRGBA8)
      *reinterpret_cast<uint32_t*>(row + (uint64_t)p.x * 4u) = synthSrgba8Texel(sm, p.x, p.y);
    else
      *reinterpret_cast<uint2*>(row + (uint64_t)p.x * 8u) = synthRgba16fTexel(sm, p.x, p.y);
  }
}

// One image: blocks stride over chunks of kSynthChunk texels.
__global__ void __launch_bounds__(256) synthKernel(SynthJob job)
{
  __shared__ LatticeCache lat;
  uint64_t sm[8];
  seedStreams(job.seed, sm);
  const uint64_t n = (uint64_t)job.width * job.height;
  for(uint64_t first = (uint64_t)blockIdx.x * kSynthChunk; first < n; first += (uint64_t)gridDim.x * kSynthChunk)
    synthTexels(lat, job, first, first + kSynthChunk < n ? first + kSynthChunk : n, sm);
}

// Many images: chunks[c] is texels [first, first + count) of jobs[chunks[c].job].
__global__ void __launch_bounds__(256) synthBatchKernel(const SynthJob* __restrict__ jobs,
                                                        const SynthChunk* __restrict__ chunks, uint32_t numChunks)
{
  __shared__ LatticeCache lat;
  for(uint32_t c = blockIdx.x; c < numChunks; c += gridDim.x)
  {
    const SynthChunk ch = chunks[c];
    const SynthJob job  = jobs[ch.job];
    uint64_t sm[8];
    seedStreams(job.seed, sm);
    synthTexels(lat, job, ch.first, ch.first + ch.count, sm);
  }
}

hipError_t launchSynthBatch(hipStream_t stream, const SynthJob* jobs, const SynthChunk* chunks, uint32_t numChunks)
{
  if(numChunks == 0)
    return hipSuccess;
  const uint32_t cap  = deviceComputeUnits() * 8u;
  const uint32_t grid = numChunks < cap ? numChunks : cap;
  hipLaunchKernelGGL(synthBatchKernel, dim3(grid), dim3(256), 0, stream, jobs, chunks, numChunks);
  return hipGetLastError();
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
  SynthJob job{(uint8_t*)image->data + off, seed, image->width, image->height, (uint32_t)pitch, format};
  const uint64_t n    = (uint64_t)image->width * image->height;
  const uint64_t want = (n + kSynthChunk - 1u) / kSynthChunk;
  const uint32_t cap  = deviceComputeUnits() * 8u;
  const uint32_t grid = (uint32_t)(want < cap ? want : cap);
  hipLaunchKernelGGL(synthKernel, dim3(grid), dim3(256), 0, stream, job);
  hipError_t e = hipGetLastError();
  if(e != hipSuccess)
    return hostFail(e, "%s: launch failed: %s", who, hipGetErrorString(e));
  return 0;
}
