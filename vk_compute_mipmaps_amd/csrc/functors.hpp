// Device functors: the load / reduce / store customisation of the schedule.
//
// The reference expresses these as GLSL macros (NVPRO_PYRAMID_TYPE, _LOAD,
// _REDUCE, _REDUCE2, _REDUCE4, _LOAD_REDUCE4, _STORE, nvpro_pyramid.glsl:27-120)
// and ships one set, sRGBA8 (srgba8_mipmap_preamble.glsl:19-128).  Here each set
// is a struct of static device functions; the kernels are templates over it.
//
//   Raw                  storage texel (what a level holds in memory)
//   Val                  NVPRO_PYRAMID_TYPE: the value reductions operate on
//   Tables               per-workgroup LDS tables (may be empty)
//   fill<BS>(t, scale)   workgroup-cooperative set-up by BS threads; `scale`
//                        premultiplies decoded values (0.25 for the 2x2 box)
//   DualTables,          tables serving load<true> (times 0.25) and
//   fillDual<BS>(t)      load<false> (times 1) alike
//   load<PRE>(t, raw)    NVPRO_PYRAMID_LOAD: Raw -> Val (times 0.25 if PRE)
//   store(t, v)          NVPRO_PYRAMID_STORE: Val -> Raw
//   add, mul, reduce3    the arithmetic REDUCE, REDUCE2 and REDUCE4 are built
//                        from (box, reduce2, reduce3 below)
//   kPrescale            whether load<true> then box<true> equals box<false>
//                        of plain loads bit for bit (exact scaling by 0.25)
//
// Every FP32 operation is written out; the library is compiled with
// -ffp-contract=off so none is fused behind our back.
#pragma once

#include <stdint.h>

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "srgb_tables.inc"

namespace nvpro_pyramid {

static __constant__ __attribute__((aligned(16))) uint32_t gSrgbDecode[256]                    = {NVPRO_SRGB_DECODE_BITS};
static __constant__ __attribute__((aligned(16))) uint32_t gSrgbEncode[NVPRO_SRGB_ENC_BUCKETS] = {NVPRO_SRGB_ENCODE_BUCKETS};

__device__ __forceinline__ float4 f4add(float4 a, float4 b)
{
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float4 f4mul(float4 a, float s)
{
  return make_float4(a.x * s, a.y * s, a.z * s, a.w * s);
}

// NVPRO_PYRAMID_REDUCE of srgba8_mipmap_preamble.glsl:24-25, a0*v0 + a1*v1 + a2*v2,
// with the contraction fixed as fma(a2, v2, fma(a1, v1, a0 * v0)).
__device__ __forceinline__ float fma3(float w0, float v0, float w1, float v1, float w2, float v2)
{
  return __builtin_fmaf(w2, v2, __builtin_fmaf(w1, v1, w0 * v0));
}

// Round an FP32 value to binary16 (nearest even) after it has been rounded to
// FP32.  The empty asm keeps the compiler from merging the producing FMA and the
// conversion into one mixed-precision instruction (v_fma_mixlo_f16), which
// would round once instead of twice.
__device__ __forceinline__ __half toHalf(float f)
{
  __asm__ volatile("" : "+v"(f));
  return __float2half_rn(f);
}
__device__ __forceinline__ float4 fma3(float w0, float4 v0, float w1, float4 v1, float w2, float4 v2)
{
  return make_float4(fma3(w0, v0.x, w1, v1.x, w2, v2.x), fma3(w0, v0.y, w1, v1.y, w2, v2.y),
                     fma3(w0, v0.z, w1, v1.z, w2, v2.z), fma3(w0, v0.w, w1, v1.w, w2, v2.w));
}

// FP32 arithmetic on float4 values, shared by the FP32 functors.
struct Fp32Arith
{
  using Val                          = float4;
  static constexpr bool kPrescale    = true;
  __device__ static Val add(Val a, Val b) { return f4add(a, b); }
  __device__ static Val mul(Val a, float s) { return f4mul(a, s); }
  __device__ static Val reduce3(float w0, Val v0, float w1, Val v1, float w2, Val v2)
  {
    return fma3(w0, v0, w1, v1, w2, v2);
  }
};

// NVPRO_PYRAMID_REDUCE2 of srgba8_mipmap_preamble.glsl:35: 0.5 * (v0 + v1).
template <class F>
__device__ __forceinline__ typename F::Val reduce2(typename F::Val v0, typename F::Val v1)
{
  return F::mul(F::add(v0, v1), 0.5f);
}

// NVPRO_PYRAMID_REDUCE4 of srgba8_mipmap_preamble.glsl:37-38,
// 0.25 * ((v00 + v01) + (v10 + v11)).  The reference passes the four quadrants
// of a 2x2 square in one of two orders (nvpro_pyramid.glsl:179-188 vs
// :364-392); `order` selects which pairs are added first:
//   order 0: (UL + LL) + (UR + LR)      order 1: (UL + UR) + (LL + LR)
// PRE: the inputs were already multiplied by 0.25 (exactly), so the sum is the
// result.
template <class F, bool PRE>
__device__ __forceinline__ typename F::Val box(typename F::Val ul, typename F::Val ur, typename F::Val ll,
                                               typename F::Val lr, int order)
{
  const typename F::Val s =
      order == 0 ? F::add(F::add(ul, ll), F::add(ur, lr)) : F::add(F::add(ul, ur), F::add(ll, lr));
  return PRE ? s : F::mul(s, 0.25f);
}

// Summation order of relative level r of a fast dispatch filling `count` levels
// (nvpro_pyramid.glsl:272-532): LOAD_REDUCE4 and the 6-level in-thread stage use
// order 0, the subgroup-shuffle stages order 1, the shared-memory stages order 0.
__host__ __device__ __forceinline__ int fastOrder(uint32_t count, uint32_t r)
{
  if(r == 1u)
    return 0;
  if(count <= 3u)
    return 1;
  if(count == 6u)
    return (r == 3u || r == 4u) ? 1 : 0;
  return r <= 3u ? 1 : 0;
}

// ---------------------------------------------------------------------------
// sRGBA8: srgba8_mipmap_preamble.glsl:19-128 with exact conversions.
// ---------------------------------------------------------------------------
struct Srgba8 : Fp32Arith
{
  using Raw                        = uint32_t;
  static constexpr uint32_t kBytes = 4;

  // Alpha is carried in code units: the reference decodes the UNORM code as
  // c / 255 and encodes uint(a * 255 + 0.5) (preamble:125), a linear map whose
  // scale commutes with every reduction, so the value here is c itself (times
  // the fill scale) and the encode is floor(a + 0.5).  Box filters of codes are
  // then exact in FP32 (a 2x2 sum of 4^k codes stays below 2^24), and every
  // .5 tie rounds up as exact arithmetic does.
  //
  // LDS tables: the decoded value of every RGB code (times the fill scale) and
  // the encode buckets.
  struct __attribute__((aligned(16))) Tables
  {
    float dec[256];
    uint32_t enc[NVPRO_SRGB_ENC_BUCKETS];  // code at bucket start << 16 | low 16 bits of the step's float - 1
  };
  static_assert(NVPRO_SRGB_ENC_BUCKETS % 4u == 0u, "16-byte table loads");
  static_assert(23u - NVPRO_SRGB_ENC_MANT_BITS == 16u, "the floats of a bucket differ in their low 16 bits");

  template <int BS>
  __device__ static void fill(Tables& t, float scale)
  {
    // 16-byte loads, every global load issued before any LDS store.
    constexpr int kEnc = (NVPRO_SRGB_ENC_BUCKETS / 4 + BS - 1) / BS;
    static_assert(BS >= 64, "the decode table is one 16-byte load per thread of the first 64");
    const uint4* genc = reinterpret_cast<const uint4*>(gSrgbEncode);
    uint4 e[kEnc];
#pragma unroll
    for(int k = 0; k < kEnc; ++k)
    {
      const uint32_t i = threadIdx.x + k * BS;
      e[k] = i < NVPRO_SRGB_ENC_BUCKETS / 4u ? genc[i] : make_uint4(0, 0, 0, 0);
    }
    const uint4 d = threadIdx.x < 64u ? reinterpret_cast<const uint4*>(gSrgbDecode)[threadIdx.x] : make_uint4(0, 0, 0, 0);
    if(threadIdx.x < 64u)
      reinterpret_cast<float4*>(t.dec)[threadIdx.x] =
          make_float4(__uint_as_float(d.x) * scale, __uint_as_float(d.y) * scale, __uint_as_float(d.z) * scale,
                      __uint_as_float(d.w) * scale);
#pragma unroll
    for(int k = 0; k < kEnc; ++k)
    {
      const uint32_t i = threadIdx.x + k * BS;
      if(i < NVPRO_SRGB_ENC_BUCKETS / 4u)
        reinterpret_cast<uint4*>(t.enc)[i] = e[k];
    }
  }

  // Tables for kernels that run premultiplied (fast) and plain (general, tail)
  // loads alike: the unscaled tables, with load<true> multiplying by 0.25 after
  // the lookup.  Scaling by 0.25 is exact for every decoded value (none is
  // subnormal), so this equals the lookup in a table filled at scale 0.25, bit
  // for bit, and the LDS holds one set of tables instead of two.
  struct DualTables : Tables
  {
  };

  template <int BS>
  __device__ static void fillDual(DualTables& t)
  {
    fill<BS>(t, 1.0f);
  }

  // texelFetch of the sRGB sampler (preamble:21-22): sRGB decode of RGB by table
  // lookup; alpha in code units (see Tables).  With plain Tables, PRE says the
  // tables were filled at 0.25 (the fast pipeline) and alpha gets the same
  // exact scale; plain loads (tables filled at 1) take the code as it is.
  template <bool PRE>
  __device__ static float4 load(const Tables& t, uint32_t texel)
  {
    const float a = (float)(texel >> 24) * (PRE ? 0.25f : 1.0f);
    return make_float4(t.dec[texel & 0xffu], t.dec[(texel >> 8) & 0xffu], t.dec[(texel >> 16) & 0xffu], a);
  }
  template <bool PRE>
  __device__ static float4 load(const DualTables& t, uint32_t texel)
  {
    const float4 v = load<false>(static_cast<const Tables&>(t), texel);
    return PRE ? f4mul(v, 0.25f) : v;
  }

  // srgbFromLinear (preamble:110-121), exact: x is clamped to [2^-13, 1) (every
  // smaller value, and NaN, encodes to 0, every larger one to 255); the bucket
  // of its exponent and top mantissa bits holds the code at the bucket start
  // and where in the bucket the code steps up.
  __device__ static uint32_t encodeChannel(const Tables& t, float x)
  {
    const uint32_t b = __float_as_uint(fminf(fmaxf(x, 0x1p-13f), 0x1.fffffep-1f));
    const uint32_t e = t.enc[(b >> (23u - NVPRO_SRGB_ENC_MANT_BITS)) - (NVPRO_SRGB_ENC_MIN_EXP << NVPRO_SRGB_ENC_MANT_BITS)];
    return (e >> 16) + ((b & 0xffffu) > (e & 0xffffu) ? 1u : 0u);
  }

  // srgbFromLinearVec (preamble:123-128); alpha (code units) = floor(a + 0.5)
  // clamped to [0, 255].
  __device__ static uint32_t store(const Tables& t, float4 v)
  {
    uint32_t a = (uint32_t)fminf(fmaxf(v.w + 0.5f, 0.0f), 255.0f);
    return encodeChannel(t, v.x) | encodeChannel(t, v.y) << 8 | encodeChannel(t, v.z) << 16 | a << 24;
  }
};

// ---------------------------------------------------------------------------
// RGBA32F: linear float4 texels, identity load and store.
// ---------------------------------------------------------------------------
struct Rgba32f : Fp32Arith
{
  struct __attribute__((aligned(4))) Raw
  {
    float x, y, z, w;
  };
  static constexpr uint32_t kBytes = 16;
  struct Tables
  {
    int unused;
  };
  using DualTables = Tables;
  template <int BS>
  __device__ static void fill(Tables&, float)
  {
  }
  template <int BS>
  __device__ static void fillDual(Tables&)
  {
  }
  template <bool PRE>
  __device__ static float4 load(const Tables&, Raw t)
  {
    float4 v = make_float4(t.x, t.y, t.z, t.w);
    return PRE ? f4mul(v, 0.25f) : v;
  }
  __device__ static Raw store(const Tables&, float4 v)
  {
    return Raw{v.x, v.y, v.z, v.w};
  }
};

// ---------------------------------------------------------------------------
// RGBA16F: linear half4 texels; FP32 arithmetic, round-to-nearest-even on store.
// ---------------------------------------------------------------------------
struct Rgba16f : Fp32Arith
{
  struct __attribute__((aligned(4))) Raw
  {
    uint32_t lo, hi;  // (r | g << 16), (b | a << 16) as IEEE half bits
  };
  static constexpr uint32_t kBytes = 8;
  struct Tables
  {
    int unused;
  };
  using DualTables = Tables;
  template <int BS>
  __device__ static void fill(Tables&, float)
  {
  }
  template <int BS>
  __device__ static void fillDual(Tables&)
  {
  }
  __device__ static float h2f(uint32_t bits) { return __half2float(__ushort_as_half((unsigned short)bits)); }
  __device__ static uint32_t f2h(float f) { return __half_as_ushort(toHalf(f)); }
  template <bool PRE>
  __device__ static float4 load(const Tables&, Raw t)
  {
    float4 v = make_float4(h2f(t.lo & 0xffffu), h2f(t.lo >> 16), h2f(t.hi & 0xffffu), h2f(t.hi >> 16));
    return PRE ? f4mul(v, 0.25f) : v;
  }
  __device__ static Raw store(const Tables&, float4 v)
  {
    return Raw{f2h(v.x) | f2h(v.y) << 16, f2h(v.z) | f2h(v.w) << 16};
  }
};

// ---------------------------------------------------------------------------
// RGBA16F with PACKED_FP16 arithmetic: 2x2 boxes and 2-tap reductions run in
// binary16 on (r,g) and (b,a) pairs (v_pk_add_f16 / v_pk_mul_f16); 3-tap
// reductions widen to FP32 for the weights and round the result to binary16.
// ---------------------------------------------------------------------------
struct Rgba16fPacked
{
  using Raw                        = Rgba16f::Raw;
  static constexpr uint32_t kBytes = 8;
  // Scaling a binary16 subnormal by 0.25 can round, so no premultiplied loads.
  static constexpr bool kPrescale = false;
  struct Val
  {
    __half2 rg, ba;
  };
  using Tables     = Rgba16f::Tables;
  using DualTables = Tables;
  template <int BS>
  __device__ static void fill(Tables&, float)
  {
  }
  template <int BS>
  __device__ static void fillDual(Tables&)
  {
  }
  template <bool PRE>
  __device__ static Val load(const Tables&, Raw t)
  {
    static_assert(!PRE, "binary16 values are not premultiplied");
    return Val{__builtin_bit_cast(__half2, t.lo), __builtin_bit_cast(__half2, t.hi)};
  }
  __device__ static Raw store(const Tables&, Val v)
  {
    return Raw{__builtin_bit_cast(uint32_t, v.rg), __builtin_bit_cast(uint32_t, v.ba)};
  }
  __device__ static Val add(Val a, Val b) { return Val{__hadd2(a.rg, b.rg), __hadd2(a.ba, b.ba)}; }
  __device__ static Val mul(Val a, float s)
  {
    const __half2 k = __float2half2_rn(s);  // 0.25 and 0.5 are exact in binary16
    return Val{__hmul2(a.rg, k), __hmul2(a.ba, k)};
  }
  __device__ static Val reduce3(float w0, Val v0, float w1, Val v1, float w2, Val v2)
  {
    const float2 a0 = __half22float2(v0.rg), a1 = __half22float2(v1.rg), a2 = __half22float2(v2.rg);
    const float2 b0 = __half22float2(v0.ba), b1 = __half22float2(v1.ba), b2 = __half22float2(v2.ba);
    return Val{__halves2half2(toHalf(fma3(w0, a0.x, w1, a1.x, w2, a2.x)), toHalf(fma3(w0, a0.y, w1, a1.y, w2, a2.y))),
               __halves2half2(toHalf(fma3(w0, b0.x, w1, b1.x, w2, b2.x)), toHalf(fma3(w0, b0.y, w1, b1.y, w2, b2.y)))};
  }
};

}  // namespace nvpro_pyramid
