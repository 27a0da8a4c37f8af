// Helpers shared by the HIP kernels: vector row access, level addressing and
// the general pipeline's sample filter.
#pragma once

#include "functors.hpp"
#include "pyramid_internal.hpp"

namespace nvpro_pyramid {

template <int NW>
struct __attribute__((aligned(4))) Words
{
  uint32_t w[NW];
};

// Load / store NBYTES (a multiple of 4) at a 4-byte aligned address with the
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

// loadRow for data read once: 16-byte loads with the non-temporal hint (rows
// whose size is not a multiple of 16 bytes load as loadRow does).
template <class F, int N>
__device__ __forceinline__ void loadRowStream(const uint8_t* p, typename F::Raw (&out)[N])
{
  constexpr int NBYTES = N * F::kBytes;
  if constexpr(NBYTES % 16 == 0)
  {
    typedef uint32_t U4 __attribute__((ext_vector_type(4)));
    uint32_t w[NBYTES / 4];
#pragma unroll
    for(int i = 0; i < NBYTES / 16; ++i)
    {
      const U4 v = __builtin_nontemporal_load(reinterpret_cast<const U4*>(p) + i);
      w[4 * i] = v.x, w[4 * i + 1] = v.y, w[4 * i + 2] = v.z, w[4 * i + 3] = v.w;
    }
    __builtin_memcpy(out, w, sizeof w);
  }
  else
    loadRow<F, N>(p, out);
}

template <class F, int N>
__device__ __forceinline__ void storeRow(uint8_t* p, const typename F::Raw (&in)[N])
{
  uint32_t w[N * F::kBytes / 4];
  __builtin_memcpy(w, in, sizeof w);
  storeWords<N * F::kBytes>(p, w);
}

template <class F>
__device__ __forceinline__ typename F::Raw loadTexel(const uint8_t* p)
{
  typename F::Raw t[1];
  loadRow<F, 1>(p, t);
  return t[0];
}

template <class F>
__device__ __forceinline__ void storeTexel(uint8_t* p, typename F::Raw v)
{
  typename F::Raw t[1] = {v};
  storeRow<F, 1>(p, t);
}

// Reads a word of every level descriptor, so that the scalar cache holds the
// whole level table (kernel-argument memory) before a kernel addresses any
// level: otherwise each level's first use that falls in a new 64-byte line
// misses and stalls the wave there (about 0.3 us on MI355X), which a kernel
// that walks many small levels pays once per line.  Call at kernel start and
// pass the result to levelsReady() after the table fill's barrier: the loads
// then all miss at once, in one round, before the first level is addressed.
__device__ __forceinline__ uint32_t prefetchLevels(const ImageDesc& img)
{
  uint32_t acc = img.levelCount;
#pragma unroll
  for(uint32_t k = 0; k < NVPRO_PYRAMID_MAX_LEVELS; ++k)
    acc ^= img.lv[k].pitch;
  return acc;
}
__device__ __forceinline__ void levelsReady(uint32_t acc) { __asm__ volatile("" ::"s"(acc)); }

__device__ __forceinline__ uint8_t* texelPtr(const ImageDesc& img, uint32_t level, uint32_t x, uint32_t y,
                                             uint32_t bytes)
{
  const LevelDesc& d = img.lv[level];
  return img.base + d.offset + (uint64_t)y * d.pitch + (uint64_t)x * bytes;
}

// A texel store that other workgroups can read once this wave's stores have
// completed (s_waitcnt): agent-scope relaxed atomic stores of 4 or 8 bytes,
// which are written through to the device's point of coherence instead of
// staying dirty in the L2 of one XCD.
template <class F>
__device__ __forceinline__ void storeTexelCoherent(uint8_t* p, typename F::Raw v)
{
  if constexpr(F::kBytes == 4)
  {
    uint32_t w;
    __builtin_memcpy(&w, &v, 4);
    __hip_atomic_store(reinterpret_cast<uint32_t*>(p), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  else
  {
    uint64_t w[F::kBytes / 8];
    __builtin_memcpy(w, &v, sizeof w);
#pragma unroll
    for(uint32_t i = 0; i < F::kBytes / 8; ++i)
      __hip_atomic_store(reinterpret_cast<uint64_t*>(p) + i, w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <class F, bool COHERENT>
__device__ __forceinline__ void storeTexelTo(uint8_t* p, typename F::Raw v)
{
  if constexpr(COHERENT)
    storeTexelCoherent<F>(p, v);
  else
    storeTexel<F>(p, v);
}

// The last workgroup of the grid to get here returns true, the others false.
// Counters as kArriveLeaves describes.
// Every thread of every workgroup calls it once, after its last store of the
// values the last workgroup will read; those stores must be storeTexelCoherent.
// The last workgroup then sees them: the counts are agent-scope atomics taken
// after every wave's stores have completed, and the winner drops its caches'
// lines (agent-scope acquire) before it reads.  No workgroup waits for
// another, so a grid of any size cannot hang here.  `counters` (kArriveWords)
// are zero at launch and zero again when the last workgroup returns, ready for
// the next launch on the same stream.  `flag`: one LDS word.  Contains barriers.
__device__ __forceinline__ bool arriveLast(uint32_t* counters, uint32_t groups, uint32_t* flag)
{
  __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if(threadIdx.x == 0)
  {
    // Up to kArriveLeaves workgroups count at the root alone (one atomic's
    // latency on the way to the continuation instead of two).
    bool last = true;
    if(groups > kArriveLeaves)
    {
      const uint32_t leaf    = blockIdx.x % kArriveLeaves;
      const uint32_t inLeaf  = (groups - leaf + kArriveLeaves - 1u) / kArriveLeaves;
      uint32_t* const leafCt = counters + kArriveStride * (1u + leaf);
      last = __hip_atomic_fetch_add(leafCt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u == inLeaf;
      if(last)
        __hip_atomic_store(leafCt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if(last)
    {
      const uint32_t atRoot = groups > kArriveLeaves ? kArriveLeaves : groups;
      last = __hip_atomic_fetch_add(counters, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u == atRoot;
      if(last)
        __hip_atomic_store(counters, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    *flag = last ? 1u : 0u;
  }
  __syncthreads();
  const bool last = *reinterpret_cast<volatile uint32_t*>(flag) != 0u;
  __syncthreads();  // `flag` may be scratch the caller reuses
  if(last)
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return last;
}

// Waits (one thread) until an agent-scope counter reaches `need`.  Bounded: after
// spinLimit polls it sets kBatchErrorSpin in *error and returns, so a kernel
// always ends.
__device__ __forceinline__ void waitCounter(uint32_t* counter, uint32_t need, uint32_t spinLimit, uint32_t* error)
{
  uint32_t n = 0;
  while(__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need)
  {
    if(++n >= spinLimit)
    {
      __hip_atomic_fetch_or(error, kBatchErrorSpin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    __builtin_amdgcn_s_sleep(2);
  }
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

// kernelSizeFromInputSize_ (nvpro_pyramid.glsl:557-561).
__host__ __device__ __forceinline__ int kernelSize(uint32_t dim) { return dim == 1u ? 1 : (int)(2u | (dim & 1u)); }

// Weights of reduceStoreSample_ (nvpro_pyramid.glsl:581-586) along an axis of a
// destination level n texels long: rcp = 1/(2n+1) once per level and axis, then
// for destination coordinate d: w0 = rcp (n - d), w1 = rcp n, w2 = (1 - w0) - w1.
struct Axis3
{
  float nf, rcp;
};
__device__ __forceinline__ Axis3 axis3(uint32_t n)
{
  const float nf = (float)n;
  return Axis3{nf, 1.0f / (2.0f * nf + 1.0f)};
}
struct Weights3
{
  float w0, w1, w2;
};
__device__ __forceinline__ Weights3 weights3(const Axis3& a, uint32_t d)
{
  const float w0 = a.rcp * (a.nf - (float)d);
  const float w1 = a.rcp * a.nf;
  return Weights3{w0, w1, (1.0f - w0) - w1};
}

// The two halves of reduceStoreSample_ (nvpro_pyramid.glsl:575-656): the
// vertical reduction of one column of a k-tall window (REDUCE of 3 with the
// destination row's weights, REDUCE2, or a copy), then the horizontal one of the
// column results.
template <class F>
__device__ __forceinline__ typename F::Val verticalTap(const Weights3& wy, int ky, typename F::Val a0,
                                                       typename F::Val a1, typename F::Val a2)
{
  if(ky == 3)
    return F::reduce3(wy.w0, a0, wy.w1, a1, wy.w2, a2);
  if(ky == 2)
    return reduce2<F>(a0, a1);
  return a0;
}

template <class F>
__device__ __forceinline__ typename F::Val horizontalTap(const Axis3& ax, uint32_t dx, int kx, typename F::Val h0,
                                                         typename F::Val h1, typename F::Val h2)
{
  if(kx == 3)
  {
    const Weights3 wx = weights3(ax, dx);
    return F::reduce3(wx.w0, h0, wx.w1, h1, wx.w2, h2);
  }
  if(kx == 2)
    return reduce2<F>(h0, h1);
  return h0;
}

// reduceStoreSample_ without the store: the sample of destination texel (dx, dy)
// of a level with axes ax, ay (axis3 of its width and height) from a kx x ky
// window whose texel at (column j, row i) `at(j, i)` returns.  Vertical first,
// then horizontal.
template <class F, class At>
__device__ __forceinline__ typename F::Val generalSample(At at, int kx, int ky, const Axis3& ax, const Axis3& ay,
                                                         uint32_t dx, uint32_t dy)
{
  using V           = typename F::Val;
  const Weights3 wy = weights3(ay, dy);
  V h[3];
#pragma unroll
  for(int j = 0; j < 3; ++j)
  {
    if(j < kx)
    {
      const V a0 = at(j, 0);
      const V a1 = ky >= 2 ? at(j, 1) : a0;
      const V a2 = ky == 3 ? at(j, 2) : a0;
      h[j]       = verticalTap<F>(wy, ky, a0, a1, a2);
    }
  }
  return horizontalTap<F>(ax, dx, kx, h[0], h[1], h[2]);
}

// generalSample with the kernel sizes as constants: kx and ky are the same for
// every texel of a level, so this is a uniform branch to straight-line code
// (no selects between the 2- and 3-tap forms).  Sides of 1 take the general
// code.
template <class F, class At>
__device__ __forceinline__ typename F::Val generalSampleAny(At at, int kx, int ky, const Axis3& ax, const Axis3& ay,
                                                            uint32_t dx, uint32_t dy)
{
  if(kx == 3 && ky == 3)
    return generalSample<F>(at, 3, 3, ax, ay, dx, dy);
  if(kx == 2 && ky == 2)
    return generalSample<F>(at, 2, 2, ax, ay, dx, dy);
  if(kx == 3 && ky == 2)
    return generalSample<F>(at, 3, 2, ax, ay, dx, dy);
  if(kx == 2 && ky == 3)
    return generalSample<F>(at, 2, 3, ax, ay, dx, dy);
  return generalSample<F>(at, kx, ky, ax, ay, dx, dy);
}

// Cross-lane moves of any value made of 32-bit words.
// DPP row shift-left: lane i reads lane i + N of its 16-lane row (bound_ctrl: 0
// beyond the row).  Only the lowest lane of each reduced group is consumed.
template <int N, class V>
__device__ __forceinline__ V dppShl(V v)
{
  static_assert(sizeof(V) % 4 == 0, "32-bit words");
  uint32_t w[sizeof(V) / 4];
  __builtin_memcpy(w, &v, sizeof v);
#pragma unroll
  for(int k = 0; k < (int)(sizeof(V) / 4); ++k)
    w[k] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)w[k], 0x100 + N, 0xf, 0xf, true);
  V r;
  __builtin_memcpy(&r, w, sizeof r);
  return r;
}

template <class V>
__device__ __forceinline__ V shflXor(V v, int mask)
{
  uint32_t w[sizeof(V) / 4];
  __builtin_memcpy(w, &v, sizeof v);
#pragma unroll
  for(int k = 0; k < (int)(sizeof(V) / 4); ++k)
    w[k] = (uint32_t)__shfl_xor((int)w[k], mask, 64);
  V r;
  __builtin_memcpy(&r, w, sizeof r);
  return r;
}

// One cross-lane 2x2 reduction level: the group leader (lowest lane) combines its
// value with lanes +mx (x neighbour), +my (y neighbour) and +mx+my.
template <class F, int MX>
__device__ __forceinline__ typename F::Val crossLaneBox(typename F::Val v, int order)
{
  constexpr int MY = 2 * MX;
  if constexpr(MY < 16)
  {
    if(order == 0)
    {
      const typename F::Val s = F::add(v, dppShl<MY>(v));
      return F::mul(F::add(s, dppShl<MX>(s)), 0.25f);
    }
    const typename F::Val s = F::add(v, dppShl<MX>(v));
    return F::mul(F::add(s, dppShl<MY>(s)), 0.25f);
  }
  else
  {
    const int first = order == 0 ? MY : MX, second = order == 0 ? MX : MY;
    const typename F::Val s = F::add(v, shflXor(v, first));
    return F::mul(F::add(s, shflXor(s, second)), 0.25f);
  }
}

}  // namespace nvpro_pyramid
