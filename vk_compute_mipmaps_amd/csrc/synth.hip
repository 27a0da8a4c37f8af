// Seeded synthetic level 0 on the device: the workload generator of the
// benchmark and serving inputs, bit-identical to vk_compute_mipmaps_amd/synth.py
// (same counter-based hash, same integer value noise; RGBA16F radiance through
// FP64 exp, rounded to half).  Every texel is independent: a workgroup writes
// chunks (strips of up to 256 columns by a block of rows), one thread per
// column; a batched launch covers the chunks of many images.
#include <hip/hip_fp16.h>

#include <vector>

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

// Round-to-nearest-even straight from binary64 (no detour through binary32).
__device__ __forceinline__ uint16_t f64ToHalfBits(double v)
{
  return __builtin_bit_cast(uint16_t, (_Float16)v);
}

__device__ __forceinline__ void seedStreams(uint64_t seed, uint64_t (&sm)[8])
{
#pragma unroll
  for(int s = 0; s < 8; ++s)
    sm[s] = mix64(seed ^ ((uint64_t)s << 56));
}

// Lattice values of the sRGBA8 value noise a strip chunk needs, per channel:
// cell columns i0 .. i0+ni-1 and rows j0 .. j0+nj-1 (one byte each).  Every
// texel of a 64x64 cell uses the same four corners, so a chunk computes them
// once.
constexpr uint32_t kLatticeMax = 4096;  // cells per channel (a strip is at most 6 cells wide)
struct LatticeCache
{
  uint8_t v[3][kLatticeMax];
};

// The texels of one chunk of a job (SynthChunk: a strip of at most 256 columns
// from x0, `rows` rows from y0).  A thread owns one column of the strip, or one
// column of one of the strip's row groups when the strip is narrower than the
// workgroup (cols columns, 256 / cols rows per pass), and walks down it, so
// the first half of every per-texel hash, mix(mix(seed ^ s << 56) ^ x), is
// computed once per thread: four of the eight mixes of a texel.  Each pass of
// the workgroup writes whole row segments (coalesced).  Uniform arguments;
// contains barriers.
__device__ __forceinline__ void synthChunk(LatticeCache& lat, const SynthJob& job, const SynthChunk& ch,
                                           const uint64_t (&sm)[8])
{
  const uint32_t w = job.width, x0 = ch.x0, y0 = ch.y0, y1 = ch.y0 + ch.rows;
  const uint32_t cols = w - x0 < 256u ? w - x0 : 256u, rpp = blockDim.x / cols;
  const uint32_t dx = threadIdx.x % cols, dy = threadIdx.x / cols;
  const bool active = dy < rpp;
  const uint32_t x  = x0 + dx;
  if(job.format == NVPRO_PYRAMID_FORMAT_SRGBA8)
  {
    const uint32_t i0 = x0 >> 6, ni = ((x0 + cols - 1u) >> 6) - i0 + 2u;
    const uint32_t j0 = y0 >> 6, nj = ((y1 - 1u) >> 6) - j0 + 2u;
    __syncthreads();  // the previous chunk's texels have read the cache
    for(uint32_t e = threadIdx.x; e < 3u * ni * nj; e += blockDim.x)
    {
      const uint32_t c = e / (ni * nj), k = e - c * ni * nj;
      const uint32_t jj = k / ni, ii = k - jj * ni;
      lat.v[c][k]       = (uint8_t)(hash4(sm[c], i0 + ii, j0 + jj) >> 56);
    }
    __syncthreads();
    if(!active)
      return;
    uint64_t cm[4];  // streams 4, 5, 6 (noise) and 7 (alpha), column half
#pragma unroll
    for(int s = 0; s < 4; ++s)
      cm[s] = mix64(sm[4 + s] ^ (uint64_t)x);
    // Value noise with the exact integer smoothstep: weights over 2^18, the
    // column weight constant per thread; every product fits 32 x 32 -> 64 bits.
    const uint32_t D = 1u << 18, mx = x & 63u, sx = mx * mx * (192u - 2u * mx);
    const uint32_t ie = (x >> 6) - i0;
    for(uint32_t y = y0 + dy; y < y1; y += rpp)
    {
      const uint32_t my = y & 63u, sy = my * my * (192u - 2u * my);
      const uint32_t e  = ((y >> 6) - j0) * ni + ie;
      uint32_t texel    = 0;
#pragma unroll
      for(int c = 0; c < 3; ++c)
      {
        const uint32_t L00 = lat.v[c][e], L10 = lat.v[c][e + 1u], L01 = lat.v[c][e + ni], L11 = lat.v[c][e + ni + 1u];
        const uint32_t top = L00 * (D - sx) + L10 * sx, bot = L01 * (D - sx) + L11 * sx;  // < 2^26
        const uint64_t num = (uint64_t)top * (D - sy) + (uint64_t)bot * sy;
        const int32_t v    = (int32_t)((num + (1ull << 35)) >> 36);
        const int32_t nz   = (int32_t)((uint32_t)(mix64(cm[c] ^ (uint64_t)y) >> 32) % 33u) - 16;
        int32_t q          = v + nz;
        q                  = q < 0 ? 0 : (q > 255 ? 255 : q);
        texel |= (uint32_t)q << (8 * c);
      }
      texel |= (uint32_t)(mix64(cm[3] ^ (uint64_t)y) >> 56) << 24;
      *reinterpret_cast<uint32_t*>(job.base + (uint64_t)y * job.pitch + (uint64_t)x * 4u) = texel;
    }
    return;
  }
  if(!active)
    return;
  // RGBA16F: RGB = exp(U) with U uniform on [ln 1e-3, ln 1e3) from streams 0..2,
  // alpha uniform on [0, 1) from stream 7; FP64, rounded to binary16 once.
  const double lo    = -6.907755278982137;  // ln 1e-3
  const double span  = 13.815510557964274;  // ln 1e3 - ln 1e-3
  const double scale = 0x1p-53;
  uint64_t cm[4];
#pragma unroll
  for(int s = 0; s < 3; ++s)
    cm[s] = mix64(sm[s] ^ (uint64_t)x);
  cm[3] = mix64(sm[7] ^ (uint64_t)x);
  for(uint32_t y = y0 + dy; y < y1; y += rpp)
  {
    uint16_t h[4];
#pragma unroll
    for(int c = 0; c < 3; ++c)
    {
      const double u = (double)(mix64(cm[c] ^ (uint64_t)y) >> 11) * scale;
      h[c]           = f64ToHalfBits(exp(lo + u * span));
    }
    h[3] = f64ToHalfBits((double)(mix64(cm[3] ^ (uint64_t)y) >> 11) * scale);
    *reinterpret_cast<uint2*>(job.base + (uint64_t)y * job.pitch + (uint64_t)x * 8u) =
        make_uint2(h[0] | (uint32_t)h[1] << 16, h[2] | (uint32_t)h[3] << 16);
  }
}

// Rows per chunk for `texels` texels per chunk of a job: whole passes of the
// workgroup, and few enough rows that the strip's lattice fits the cache.
__host__ __device__ inline uint32_t synthChunkRows(uint32_t width, uint64_t texels)
{
  const uint32_t cols = width < 256u ? width : 256u;
  uint64_t rows       = texels / cols;
  rows                = rows < 1u ? 1u : rows;
  return rows > 16384u ? 16384u : (uint32_t)rows;  // 16384 rows: 258 cell rows x 6 columns per channel
}

// One image: blocks stride over its chunks (strips x row blocks).
__global__ void __launch_bounds__(256) synthKernel(SynthJob job, uint32_t rowsPerChunk)
{
  __shared__ LatticeCache lat;
  uint64_t sm[8];
  seedStreams(job.seed, sm);
  const uint32_t strips = (job.width + 255u) / 256u, blocks = (job.height + rowsPerChunk - 1u) / rowsPerChunk;
  for(uint32_t c = blockIdx.x; c < strips * blocks; c += gridDim.x)
  {
    const uint32_t b = c / strips, st = c - b * strips;
    const uint32_t y0 = b * rowsPerChunk, rows = job.height - y0 < rowsPerChunk ? job.height - y0 : rowsPerChunk;
    synthChunk(lat, job, SynthChunk{0u, st * 256u, y0, rows}, sm);
  }
}

// Many images: chunks[c] is a strip chunk of jobs[chunks[c].job].
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
    synthChunk(lat, job, ch, sm);
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

uint32_t appendSynthChunks(std::vector<SynthChunk>& chunks, uint32_t job, uint32_t width, uint32_t height,
                           uint64_t texelsPerChunk)
{
  const uint32_t rows = synthChunkRows(width, texelsPerChunk);
  uint32_t n          = 0;
  for(uint32_t y0 = 0; y0 < height; y0 += rows)
    for(uint32_t x0 = 0; x0 < width; x0 += 256u, ++n)
      chunks.push_back(SynthChunk{job, x0, y0, height - y0 < rows ? height - y0 : rows});
  return n;
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
  const uint32_t rows = synthChunkRows(image->width, kSynthChunk);
  const uint64_t want = (uint64_t)((image->width + 255u) / 256u) * ((image->height + rows - 1u) / rows);
  const uint32_t cap  = deviceComputeUnits() * 8u;
  const uint32_t grid = (uint32_t)(want < cap ? want : cap);
  hipLaunchKernelGGL(synthKernel, dim3(grid), dim3(256), 0, stream, job, rows);
  hipError_t e = hipGetLastError();
  if(e != hipSuccess)
    return hostFail(e, "%s: launch failed: %s", who, hipGetErrorString(e));
  return 0;
}
