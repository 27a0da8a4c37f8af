/*
 * orc.c — CPU oracle for the nvpro_pyramid mip-chain path.
 *
 * TEST INFRASTRUCTURE ONLY: the product never links or calls this file.  Built
 * with -ffp-contract=off so every FP32 operation rounds exactly where written.
 * Function-level comments cite the reference file:line each one restates
 * (paths relative to the nvpro-samples/vk_compute_mipmaps repository).
 */
#include "orc.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Layout                                                                     */
/* ------------------------------------------------------------------------ */

/* nvpro_pyramid_dispatch.hpp:122-132: count shifts until both edges reach 0. */
uint32_t orc_full_levels(uint32_t w, uint32_t h)
{
  uint32_t n = 0;
  while(w != 0 || h != 0)
  {
    w >>= 1;
    h >>= 1;
    ++n;
  }
  return n;
}

/* mipmap_storage.hpp:67-69: halve rounding down, never below 1. */
void orc_level_dims(uint32_t w, uint32_t h, uint32_t level, uint32_t* lw, uint32_t* lh)
{
  for(uint32_t i = 0; i < level; ++i)
  {
    w = (w >> 1) | (w == 1u);
    h = (h >> 1) | (h == 1u);
  }
  *lw = w;
  *lh = h;
}

/* mipmap_storage.hpp:53-76: level l starts after the texels of levels < l. */
uint64_t orc_packed_offsets(uint32_t w, uint32_t h, uint32_t levels, uint64_t* offsets)
{
  uint64_t off = 0;
  for(uint32_t l = 0; l < levels; ++l)
  {
    uint32_t lw, lh;
    orc_level_dims(w, h, l, &lw, &lh);
    if(offsets)
      offsets[l] = off;
    off += (uint64_t)lw * lh;
  }
  return off;
}

/* ------------------------------------------------------------------------ */
/* Planner: nvproCmdPyramidDispatch's loop (nvpro_pyramid_dispatch.hpp:109-188) */
/* with the default fast (:195-242) and general (:247-292) dispatchers.        */
/* ------------------------------------------------------------------------ */
int orc_plan(uint32_t w, uint32_t h, uint32_t mipLevels, int fastAvailable, uint32_t div,
             uint32_t fastMax, orc_step* out, uint32_t cap)
{
  if(mipLevels == 0)
    mipLevels = orc_full_levels(w, h);
  uint32_t level = 0, remaining = mipLevels - 1u, x = w, y = h;
  int n = 0;
  while(remaining != 0)
  {
    orc_step s;
    memset(&s, 0, sizeof s);
    uint32_t done = 0;
    if(fastAvailable && (x % div) == 0 && (y % div) == 0)
    {
      uint32_t tx = x, ty = y;
      while(tx % 2u == 0 && ty % 2u == 0 && done < remaining && done < fastMax)
      {
        tx /= 2u;
        ty /= 2u;
        done++;
      }
      uint64_t samples = (uint64_t)x * y;
      s.pipeline    = ORC_PIPE_FAST;
      s.groupCountX = (uint32_t)(done > 5 ? (samples + 4095u) >> 12 : (samples + 1023u) >> 10);
    }
    if(done == 0)
    {
      done           = remaining >= 2u ? 2u : remaining;
      uint32_t dw    = x >> done ? x >> done : 1u;
      uint32_t dh    = y >> done ? y >> done : 1u;
      s.pipeline     = ORC_PIPE_GENERAL;
      s.groupCountX  = done == 1u ? (uint32_t)(((uint64_t)dw * dh + 127u) / 128u) :
                                    ((dw + 7u) / 8u) * ((dh + 7u) / 8u);
    }
    if(done == 0 || done > remaining)
      return -1;
    s.inputLevel   = level;
    s.levelCount   = done;
    s.pushConstant = level << 5 | done;
    s.inputWidth   = x;
    s.inputHeight  = y;
    if((uint32_t)n < cap)
      out[n] = s;
    ++n;
    level += done;
    remaining -= done;
    x >>= done;
    x = x ? x : 1u;
    y >>= done;
    y = y ? y : 1u;
  }
  return n;
}

/* ------------------------------------------------------------------------ */
/* sRGB conversions                                                           */
/* ------------------------------------------------------------------------ */

/* shaders/srgb.h:18-28, evaluated in double. */
double orc_srgb_decode_f64(uint32_t code)
{
  if(code > 255u)
    code = 255u;
  double u = (double)code / 255.0;
  return u <= 0.04045 ? u * (25.0 / 323.0) : pow((200.0 * u + 11.0) / 211.0, 2.4);
}

/* FP32 rounding of the exact decode. */
float orc_srgb_decode_f32(uint32_t code) { return (float)orc_srgb_decode_f64(code); }

/* Alpha in code units: the UNORM decode c / 255 of shaders/srgb.h:55-61 with
 * the factor 255 moved into the encode (a linear scale commutes with every
 * reduction of the schedule). */
float orc_alpha_decode_f32(uint32_t code) { return (float)code; }

/* srgba8_mipmap_preamble.glsl:110-121 in exact arithmetic (double is exact for
 * every float input to well below one code; pinned by the threshold tests). */
uint32_t orc_srgb_encode_exact(double x)
{
  if(!(x > 0.0))
    return 0u;
  double s = x <= 0.0031308 ? 12.92 * x : 1.055 * pow(x, 1.0 / 2.4) - 0.055;
  double u = s * 255.0 + 0.5;
  if(u <= 0.0)
    return 0u;
  if(u >= 255.0)
    return 255u;
  return (uint32_t)floor(u);
}

/* srgba8_mipmap_preamble.glsl:125, uint(a * 255 + 0.5) clamped, on a value in
 * code units: floor(a + 0.5). */
uint32_t orc_alpha_encode_f32(float a)
{
  float t = a + 0.5f;
  if(!(t > 0.0f))
    return 0u;
  if(t >= 255.0f)
    return 255u;
  return (uint32_t)t;
}

uint32_t orc_alpha_encode_f64(double a)
{
  double t = a + 0.5;
  if(!(t > 0.0))
    return 0u;
  if(t >= 255.0)
    return 255u;
  return (uint32_t)floor(t);
}

static float bits_to_float(uint32_t b)
{
  float f;
  memcpy(&f, &b, 4);
  return f;
}

float orc_srgb_threshold(uint32_t k)
{
  /* Bisection over the bit patterns of [0, 1]: encode is non-decreasing. */
  uint32_t lo = 0u, hi = 0x3f800000u; /* encode(lo) < k <= encode(hi) */
  if(k == 0u)
    return 0.0f;
  while(hi - lo > 1u)
  {
    uint32_t mid = lo + (hi - lo) / 2u;
    if(orc_srgb_encode_exact((double)bits_to_float(mid)) >= k)
      hi = mid;
    else
      lo = mid;
  }
  return bits_to_float(hi);
}

/* ------------------------------------------------------------------------ */
/* Schedule emulators                                                         */
/* ------------------------------------------------------------------------ */

typedef struct { float c[4]; } f4;
typedef struct { double c[4]; } d4;

static float g_decode_f32[256];
static double g_decode_f64[256];
static pthread_once_t g_tables_once = PTHREAD_ONCE_INIT;

static void init_tables(void)
{
  for(uint32_t i = 0; i < 256u; ++i)
  {
    g_decode_f64[i] = orc_srgb_decode_f64(i);
    g_decode_f32[i] = (float)g_decode_f64[i];
  }
}

/* Which 2x2 summation order the reference uses for relative level r of a fast
 * dispatch of `count` levels.  0 = ((UL+LL)+(UR+LR)), 1 = ((UL+UR)+(LL+LR)).
 *  r = 1            : LOAD_REDUCE4, REDUCE4(v00, v01=(0,1), v10=(1,0), v11)   glsl:179-188
 *  count = 6, r = 2 : REDUCE4(s00, s01=(0,2), s10=(2,0), s11) in-thread      glsl:317-342
 *  shuffle stages   : v01 = xor 1 (x neighbour), v10 = xor 2 (y neighbour)    glsl:364-392
 *  shared stages    : in01 = smem[+2] (y), in10 = smem[+1] (x)               glsl:487-526 */
int orc_debug_flags = 0;

static int fast_order_ref(uint32_t count, uint32_t r)
{
  if(r == 1u)
    return 0;
  if(count <= 3u)
    return 1;
  if(count == 6u)
    return (r == 3u || r == 4u) ? 1 : 0;
  /* count 4 or 5: three levels in the 16-thread sub-teams, then shared memory */
  return r <= 3u ? 1 : 0;
}

/* ORC_DEBUG_FLIP_ORDER plants a summation-order mistake (tests only). */
static int fast_order(uint32_t count, uint32_t r)
{
  int o = fast_order_ref(count, r);
  return (orc_debug_flags & ORC_DEBUG_FLIP_ORDER) ? 1 - o : o;
}

static f4 box_f32(f4 ul, f4 ur, f4 ll, f4 lr, int order)
{
  f4 o;
  for(int c = 0; c < 4; ++c)
    o.c[c] = order == 0 ? 0.25f * ((ul.c[c] + ll.c[c]) + (ur.c[c] + lr.c[c])) :
                          0.25f * ((ul.c[c] + ur.c[c]) + (ll.c[c] + lr.c[c]));
  return o;
}

static d4 box_f64(d4 ul, d4 ur, d4 ll, d4 lr)
{
  d4 o;
  for(int c = 0; c < 4; ++c)
    o.c[c] = 0.25 * ((ul.c[c] + ll.c[c]) + (ur.c[c] + lr.c[c]));
  return o;
}

/* kernelSizeFromInputSize_ (glsl:557-561). */
static int ksize(uint32_t dim) { return dim == 1u ? 1 : (int)(2u | (dim & 1u)); }

/* REDUCE of srgba8_mipmap_preamble.glsl:24-25 with contraction fixed. */
static float reduce3_f32(float w0, float v0, float w1, float v1, float w2, float v2)
{
  float t = w0 * v0;
  t = fmaf(w1, v1, t);
  return fmaf(w2, v2, t);
}

/* reduceStoreSample_ (glsl:575-656) in FP32: vertical per column, then horizontal. */
static f4 general_f32(const f4* src, uint32_t sw, uint32_t dw, uint32_t dh, uint32_t x, uint32_t y,
                      int kx, int ky)
{
  float n   = (float)dh;
  float rcp = 1.0f / (2.0f * n + 1.0f);
  float w0  = rcp * (n - (float)y);
  float w1  = rcp * n;
  float w2  = (orc_debug_flags & ORC_DEBUG_CPU_WEIGHTS) ? rcp * (float)(y + 1u) : (1.0f - w0) - w1;
  f4 hcol[3];
  for(int j = 0; j < kx; ++j)
  {
    const f4* col = src + (uint64_t)(2u * y) * sw + 2u * x + (uint32_t)j;
    for(int c = 0; c < 4; ++c)
    {
      if(ky == 3)
        hcol[j].c[c] = reduce3_f32(w0, col[0].c[c], w1, col[sw].c[c], w2, col[2u * sw].c[c]);
      else if(ky == 2)
        hcol[j].c[c] = 0.5f * (col[0].c[c] + col[sw].c[c]);
      else
        hcol[j].c[c] = col[0].c[c];
    }
  }
  f4 o;
  if(kx == 3)
  {
    n   = (float)dw;
    rcp = 1.0f / (2.0f * n + 1.0f);
    w0  = rcp * (n - (float)x);
    w1  = rcp * n;
    w2  = (1.0f - w0) - w1;
    for(int c = 0; c < 4; ++c)
      o.c[c] = reduce3_f32(w0, hcol[0].c[c], w1, hcol[1].c[c], w2, hcol[2].c[c]);
  }
  else if(kx == 2)
  {
    for(int c = 0; c < 4; ++c)
      o.c[c] = 0.5f * (hcol[0].c[c] + hcol[1].c[c]);
  }
  else
    o = hcol[0];
  return o;
}

/* The same filter in FP64 with the exact weights (n-y, n, y+1) / (2n+1) of the
 * NP2 paper cited at mipmap_storage.hpp:337-338. */
static d4 general_f64(const d4* src, uint32_t sw, uint32_t dw, uint32_t dh, uint32_t x, uint32_t y,
                      int kx, int ky)
{
  double n = dh, den = 2.0 * n + 1.0;
  double w0 = (n - y) / den, w1 = n / den, w2 = (y + 1.0) / den;
  d4 hcol[3];
  for(int j = 0; j < kx; ++j)
  {
    const d4* col = src + (uint64_t)(2u * y) * sw + 2u * x + (uint32_t)j;
    for(int c = 0; c < 4; ++c)
    {
      if(ky == 3)
        hcol[j].c[c] = w0 * col[0].c[c] + w1 * col[sw].c[c] + w2 * col[2u * sw].c[c];
      else if(ky == 2)
        hcol[j].c[c] = 0.5 * (col[0].c[c] + col[sw].c[c]);
      else
        hcol[j].c[c] = col[0].c[c];
    }
  }
  d4 o;
  if(kx == 3)
  {
    n = dw;
    den = 2.0 * n + 1.0;
    w0 = (n - x) / den;
    w1 = n / den;
    w2 = (x + 1.0) / den;
    for(int c = 0; c < 4; ++c)
      o.c[c] = w0 * hcol[0].c[c] + w1 * hcol[1].c[c] + w2 * hcol[2].c[c];
  }
  else if(kx == 2)
  {
    for(int c = 0; c < 4; ++c)
      o.c[c] = 0.5 * (hcol[0].c[c] + hcol[1].c[c]);
  }
  else
    o = hcol[0];
  return o;
}

static void store_srgba8_f32(uint8_t* t, f4 v)
{
  for(int c = 0; c < 3; ++c)
    t[c] = (uint8_t)orc_srgb_encode_exact((double)v.c[c]);
  t[3] = (uint8_t)orc_alpha_encode_f32(v.c[3]);
}

static void store_srgba8_f64(uint8_t* t, d4 v)
{
  for(int c = 0; c < 3; ++c)
    t[c] = (uint8_t)orc_srgb_encode_exact(v.c[c]);
  t[3] = (uint8_t)orc_alpha_encode_f64(v.c[3]);
}

static int check_plan(uint32_t w, uint32_t h, uint32_t mipLevels, const orc_step* steps, uint32_t nSteps)
{
  uint32_t next = 0;
  for(uint32_t i = 0; i < nSteps; ++i)
  {
    if(steps[i].inputLevel != next || steps[i].levelCount == 0)
      return -1;
    next += steps[i].levelCount;
  }
  (void)w;
  (void)h;
  return next + 1u == mipLevels ? 0 : -1;
}

int orc_pyramid_srgba8_f32(uint8_t* chain, uint32_t w, uint32_t h, uint32_t mipLevels,
                           const orc_step* steps, uint32_t nSteps)
{
  pthread_once(&g_tables_once, init_tables);
  if(mipLevels == 0)
    mipLevels = orc_full_levels(w, h);
  if(check_plan(w, h, mipLevels, steps, nSteps))
    return -1;
  uint64_t offs[32];
  orc_packed_offsets(w, h, mipLevels, offs);
  for(uint32_t s = 0; s < nSteps; ++s)
  {
    uint32_t L = steps[s].inputLevel, count = steps[s].levelCount;
    uint32_t sw, sh;
    orc_level_dims(w, h, L, &sw, &sh);
    f4* prev = (f4*)malloc(sizeof(f4) * (uint64_t)sw * sh);
    const uint8_t* in = chain + offs[L] * 4u;
    for(uint64_t i = 0; i < (uint64_t)sw * sh; ++i)
    {
      for(int c = 0; c < 3; ++c)
        prev[i].c[c] = g_decode_f32[in[4 * i + c]];
      prev[i].c[3] = orc_alpha_decode_f32(in[4 * i + 3]);
    }
    for(uint32_t r = 1; r <= count; ++r)
    {
      uint32_t dw, dh;
      orc_level_dims(w, h, L + r, &dw, &dh);
      f4* cur = (f4*)malloc(sizeof(f4) * (uint64_t)dw * dh);
      uint8_t* out = chain + offs[L + r] * 4u;
      if(steps[s].pipeline == ORC_PIPE_FAST)
      {
        int order = fast_order(count, r);
        for(uint32_t y = 0; y < dh; ++y)
          for(uint32_t x = 0; x < dw; ++x)
          {
            const f4* p = prev + (uint64_t)(2u * y) * sw + 2u * x;
            cur[(uint64_t)y * dw + x] = box_f32(p[0], p[1], p[sw], p[sw + 1], order);
          }
      }
      else
      {
        int kx = ksize(sw), ky = ksize(sh);
        for(uint32_t y = 0; y < dh; ++y)
          for(uint32_t x = 0; x < dw; ++x)
            cur[(uint64_t)y * dw + x] = general_f32(prev, sw, dw, dh, x, y, kx, ky);
      }
      for(uint64_t i = 0; i < (uint64_t)dw * dh; ++i)
        store_srgba8_f32(out + 4 * i, cur[i]);
      free(prev);
      prev = cur;
      sw   = dw;
      sh   = dh;
    }
    free(prev);
  }
  return 0;
}

int orc_pyramid_srgba8_f64(uint8_t* chain, uint32_t w, uint32_t h, uint32_t mipLevels,
                           const orc_step* steps, uint32_t nSteps)
{
  pthread_once(&g_tables_once, init_tables);
  if(mipLevels == 0)
    mipLevels = orc_full_levels(w, h);
  if(check_plan(w, h, mipLevels, steps, nSteps))
    return -1;
  uint64_t offs[32];
  orc_packed_offsets(w, h, mipLevels, offs);
  for(uint32_t s = 0; s < nSteps; ++s)
  {
    uint32_t L = steps[s].inputLevel, count = steps[s].levelCount;
    uint32_t sw, sh;
    orc_level_dims(w, h, L, &sw, &sh);
    d4* prev = (d4*)malloc(sizeof(d4) * (uint64_t)sw * sh);
    const uint8_t* in = chain + offs[L] * 4u;
    for(uint64_t i = 0; i < (uint64_t)sw * sh; ++i)
    {
      for(int c = 0; c < 3; ++c)
        prev[i].c[c] = g_decode_f64[in[4 * i + c]];
      prev[i].c[3] = (double)in[4 * i + 3];
    }
    for(uint32_t r = 1; r <= count; ++r)
    {
      uint32_t dw, dh;
      orc_level_dims(w, h, L + r, &dw, &dh);
      d4* cur = (d4*)malloc(sizeof(d4) * (uint64_t)dw * dh);
      uint8_t* out = chain + offs[L + r] * 4u;
      int kx = ksize(sw), ky = ksize(sh);
      for(uint32_t y = 0; y < dh; ++y)
        for(uint32_t x = 0; x < dw; ++x)
        {
          if(kx == 2 && ky == 2)
          {
            const d4* p = prev + (uint64_t)(2u * y) * sw + 2u * x;
            cur[(uint64_t)y * dw + x] = box_f64(p[0], p[1], p[sw], p[sw + 1]);
          }
          else
            cur[(uint64_t)y * dw + x] = general_f64(prev, sw, dw, dh, x, y, kx, ky);
        }
      for(uint64_t i = 0; i < (uint64_t)dw * dh; ++i)
        store_srgba8_f64(out + 4 * i, cur[i]);
      free(prev);
      prev = cur;
      sw   = dw;
      sh   = dh;
    }
    free(prev);
  }
  return 0;
}

void orc_pyramid_linear_f64(double* chain, uint32_t w, uint32_t h, uint32_t mipLevels)
{
  if(mipLevels == 0)
    mipLevels = orc_full_levels(w, h);
  uint64_t offs[32];
  orc_packed_offsets(w, h, mipLevels, offs);
  for(uint32_t l = 1; l < mipLevels; ++l)
  {
    uint32_t sw, sh, dw, dh;
    orc_level_dims(w, h, l - 1, &sw, &sh);
    orc_level_dims(w, h, l, &dw, &dh);
    const d4* src = (const d4*)(chain + offs[l - 1] * 4u);
    d4* dst = (d4*)(chain + offs[l] * 4u);
    int kx = ksize(sw), ky = ksize(sh);
    for(uint32_t y = 0; y < dh; ++y)
      for(uint32_t x = 0; x < dw; ++x)
        dst[(uint64_t)y * dw + x] = general_f64(src, sw, dw, dh, x, y, kx, ky);
  }
}

/* ------------------------------------------------------------------------ */
/* Premultiplied alpha while staging (include/scoped_image.hpp:233-252)       */
/* ------------------------------------------------------------------------ */

static float ref_linear_from_srgb(uint32_t arg);
static uint32_t ref_srgb_from_linear(float arg);

/* alpha = a * (1/255); rgb = srgbFromLinear(linearFromSrgb(c) * alpha); alpha
 * byte kept.  exact != 0: the FP32 rounding of the exact decode and the exact
 * encode (the library's conversions); exact == 0: the reference's FP32 powf
 * formulas of shaders/srgb.h:18-41. */
void orc_premultiply_srgba8(uint8_t* texels, uint64_t count, int exact)
{
  for(uint64_t i = 0; i < count; ++i)
  {
    uint8_t* t        = texels + 4u * i;
    const float alpha = (float)t[3] * (1.0f / 255.0f);
    for(int c = 0; c < 3; ++c)
    {
      if(exact)
      {
        const float v = orc_srgb_decode_f32(t[c]) * alpha;
        t[c]          = (uint8_t)orc_srgb_encode_exact((double)v);
      }
      else
        t[c] = (uint8_t)ref_srgb_from_linear(ref_linear_from_srgb(t[c]) * alpha);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* The reference CPU generator (include/mipmap_storage.hpp:207-414).          */
/* ------------------------------------------------------------------------ */

/* shaders/srgb.h:18-28, FP32 with powf. */
static float ref_linear_from_srgb(uint32_t arg)
{
  arg     = arg > 255u ? 255u : arg;
  float u = (float)arg * (1.0F / 255.0F);
  return u <= 0.04045F ? u * (25.0F / 323.0F) : powf((200.0F * u + 11.0F) * (1.0F / 211.0F), 2.4F);
}

/* shaders/srgb.h:30-41 with bias 0.5. */
static uint32_t ref_srgb_from_linear(float arg)
{
  float srgb = arg <= 0.0031308F ? (323.0F / 25.0F) * arg : 1.055F * powf(arg, 1.0F / 2.4F) - 0.055F;
  float t    = srgb * 255.0F + 0.5F;
  t          = t < 0.F ? 0.F : (t > 255.F ? 255.F : t);
  return (uint32_t)t;
}

typedef struct ref_level_job {
  uint8_t* chain;
  const uint64_t* offs;
  uint32_t w, h, level, y0, y1;
} ref_level_job;

static void ref_load(const uint8_t* src, uint32_t sw, uint32_t x, uint32_t y, float s[4])
{
  const uint8_t* t = src + 4u * ((uint64_t)y * sw + x);
  s[0] = ref_linear_from_srgb(t[0]);
  s[1] = ref_linear_from_srgb(t[1]);
  s[2] = ref_linear_from_srgb(t[2]);
  s[3] = t[3] * (1.F / 255.F);
}

static void ref_mac(float lhs[4], const float rhs[4], float wt)
{
  for(int c = 0; c < 4; ++c)
    lhs[c] += rhs[c] * wt;
}

/* generateLevel<SrcWidthEven, SrcHeightEven> (mipmap_storage.hpp:246-392). */
static void ref_generate_rows(const ref_level_job* j)
{
  uint32_t sw, sh, dw, dh;
  orc_level_dims(j->w, j->h, j->level - 1u, &sw, &sh);
  orc_level_dims(j->w, j->h, j->level, &dw, &dh);
  const int swe = !(sw & 1u), she = !(sh & 1u);
  const uint8_t* src = j->chain + j->offs[j->level - 1u] * 4u;
  uint8_t* dst       = j->chain + j->offs[j->level] * 4u;
  for(uint32_t y = j->y0; y < j->y1; ++y)
  {
    for(uint32_t x = 0; x < dw; ++x)
    {
      float s00[4], s10[4], s20[4], s01[4], s11[4], s21[4], s02[4], s12[4], s22[4];
      const int haveX1 = swe || sw != 1u, haveY1 = she || sh != 1u;
      ref_load(src, sw, 2 * x, 2 * y, s00);
      if(haveX1)
        ref_load(src, sw, 2 * x + 1, 2 * y, s10);
      if(haveY1)
      {
        ref_load(src, sw, 2 * x, 2 * y + 1, s01);
        if(haveX1)
          ref_load(src, sw, 2 * x + 1, 2 * y + 1, s11);
      }
      if(!swe && sw != 1u)
      {
        ref_load(src, sw, 2 * x + 2, 2 * y, s20);
        if(haveY1) /* the reference also loads it when sh == 1; the value is unused */
          ref_load(src, sw, 2 * x + 2, 2 * y + 1, s21);
        if(!she && sh != 1u)
          ref_load(src, sw, 2 * x + 2, 2 * y + 2, s22);
      }
      if(!she && sh != 1u)
      {
        ref_load(src, sw, 2 * x, 2 * y + 2, s02);
        ref_load(src, sw, 2 * x + 1, 2 * y + 2, s12);
      }

      float s0[4] = {0, 0, 0, 0}, s1[4] = {0, 0, 0, 0}, s2[4] = {0, 0, 0, 0};
      if(she)
      {
        ref_mac(s0, s00, 0.5F);
        ref_mac(s0, s01, 0.5F);
        if(haveX1)
        {
          ref_mac(s1, s10, 0.5F);
          ref_mac(s1, s11, 0.5F);
        }
        if(!swe && sw != 1u)
        {
          ref_mac(s2, s20, 0.5F);
          ref_mac(s2, s21, 0.5F);
        }
      }
      else if(sh == 1u)
      {
        memcpy(s0, s00, sizeof s0);
        if(haveX1)
          memcpy(s1, s10, sizeof s1);
        if(!swe && sw != 1u)
          memcpy(s2, s20, sizeof s2);
      }
      else
      {
        const float n   = (float)dh;
        const float rcp = 1.0F / (2 * n + 1);
        const float w0  = rcp * (n - (float)y);
        const float w1  = rcp * n;
        const float w2  = rcp * (float)(1 + y);
        ref_mac(s0, s00, w0);
        ref_mac(s0, s01, w1);
        ref_mac(s0, s02, w2);
        if(haveX1)
        {
          ref_mac(s1, s10, w0);
          ref_mac(s1, s11, w1);
          ref_mac(s1, s12, w2);
        }
        if(!swe && sw != 1u) /* the reference reads unloaded samples when sw == 1; unused */
        {
          ref_mac(s2, s20, w0);
          ref_mac(s2, s21, w1);
          ref_mac(s2, s22, w2);
        }
      }

      float r[4] = {0, 0, 0, 0};
      if(swe)
      {
        ref_mac(r, s0, 0.5F);
        ref_mac(r, s1, 0.5F);
      }
      else if(sw == 1u)
      {
        memcpy(r, s0, sizeof r);
      }
      else
      {
        const float n   = (float)dw;
        const float rcp = 1.0F / (2 * n + 1);
        const float w0  = rcp * (n - (float)x);
        const float w1  = rcp * n;
        const float w2  = rcp * (float)(1 + x);
        ref_mac(r, s0, w0);
        ref_mac(r, s1, w1);
        ref_mac(r, s2, w2);
      }

      uint8_t* o = dst + 4u * ((uint64_t)y * dw + x);
      float a    = r[3] * 255.F;
      a          = a < 0.F ? 0.F : (a > 255.F ? 255.F : a);
      o[0]       = (uint8_t)ref_srgb_from_linear(r[0]);
      o[1]       = (uint8_t)ref_srgb_from_linear(r[1]);
      o[2]       = (uint8_t)ref_srgb_from_linear(r[2]);
      o[3]       = (uint8_t)a; /* truncation: mipmap_storage.hpp:403-404 */
    }
  }
}

static void* ref_thread(void* p)
{
  ref_generate_rows((const ref_level_job*)p);
  return NULL;
}

void orc_reference_cpu_srgba8_mt(uint8_t* chain, uint32_t w, uint32_t h, uint32_t threads)
{
  uint32_t levels = orc_full_levels(w, h);
  uint64_t offs[32];
  orc_packed_offsets(w, h, levels, offs);
  if(threads == 0)
    threads = 1;
  if(threads > 256)
    threads = 256;
  for(uint32_t l = 1; l < levels; ++l)
  {
    uint32_t dw, dh;
    orc_level_dims(w, h, l, &dw, &dh);
    uint32_t nt = threads < dh ? threads : dh;
    ref_level_job jobs[256];
    pthread_t tid[256];
    for(uint32_t t = 0; t < nt; ++t)
    {
      jobs[t].chain = chain;
      jobs[t].offs  = offs;
      jobs[t].w     = w;
      jobs[t].h     = h;
      jobs[t].level = l;
      jobs[t].y0    = (uint32_t)((uint64_t)dh * t / nt);
      jobs[t].y1    = (uint32_t)((uint64_t)dh * (t + 1) / nt);
    }
    if(nt == 1)
      ref_generate_rows(&jobs[0]);
    else
    {
      for(uint32_t t = 0; t < nt; ++t)
        pthread_create(&tid[t], NULL, ref_thread, &jobs[t]);
      for(uint32_t t = 0; t < nt; ++t)
        pthread_join(tid[t], NULL);
    }
  }
}

void orc_reference_cpu_srgba8(uint8_t* chain, uint32_t w, uint32_t h)
{
  orc_reference_cpu_srgba8_mt(chain, w, h, 1);
}

/* MipmapStorage::compare (mipmap_storage.hpp:159-202). */
uint32_t orc_compare_srgba8(const uint8_t* a, const uint8_t* b, uint32_t w, uint32_t h, uint32_t levels,
                            uint32_t* worstX, uint32_t* worstY, uint32_t* worstLevel,
                            uint32_t* worstChannel, uint64_t* diffTexels, uint64_t* totalTexels)
{
  if(levels == 0)
    levels = orc_full_levels(w, h);
  uint64_t offs[32];
  orc_packed_offsets(w, h, levels, offs);
  uint32_t worst = 0, wx = 0, wy = 0, wl = 0, wc = 0;
  uint64_t diff = 0, total = 0;
  for(uint32_t l = 1; l < levels; ++l)
  {
    uint32_t dw, dh;
    orc_level_dims(w, h, l, &dw, &dh);
    const uint8_t* pa = a + offs[l] * 4u;
    const uint8_t* pb = b + offs[l] * 4u;
    for(uint32_t y = 0; y < dh; ++y)
      for(uint32_t x = 0; x < dw; ++x)
      {
        uint64_t i = 4u * ((uint64_t)y * dw + x);
        int any    = 0;
        for(uint32_t c = 0; c < 4; ++c)
        {
          uint32_t d = pa[i + c] > pb[i + c] ? pa[i + c] - pb[i + c] : pb[i + c] - pa[i + c];
          any |= d != 0;
          if(d > worst)
          {
            worst = d;
            wx    = x;
            wy    = y;
            wl    = l;
            wc    = c;
          }
        }
        diff += (uint64_t)any;
        total++;
      }
  }
  if(worstX)
    *worstX = wx;
  if(worstY)
    *worstY = wy;
  if(worstLevel)
    *worstLevel = wl;
  if(worstChannel)
    *worstChannel = wc;
  if(diffTexels)
    *diffTexels = diff;
  if(totalTexels)
    *totalTexels = total;
  return worst;
}

/* ------------------------------------------------------------------------ */
/* Float-texel emulator (the RGBA32F functors: TYPE vec4, LOAD/STORE identity). */
/* ------------------------------------------------------------------------ */
int orc_pyramid_rgba32f_f32(float* chain, uint32_t w, uint32_t h, uint32_t mipLevels,
                            const orc_step* steps, uint32_t nSteps)
{
  if(mipLevels == 0)
    mipLevels = orc_full_levels(w, h);
  if(check_plan(w, h, mipLevels, steps, nSteps))
    return -1;
  uint64_t offs[32];
  orc_packed_offsets(w, h, mipLevels, offs);
  for(uint32_t s = 0; s < nSteps; ++s)
  {
    uint32_t L = steps[s].inputLevel, count = steps[s].levelCount;
    uint32_t sw, sh;
    orc_level_dims(w, h, L, &sw, &sh);
    const f4* prev = (const f4*)(chain + offs[L] * 4u);
    for(uint32_t r = 1; r <= count; ++r)
    {
      uint32_t dw, dh;
      orc_level_dims(w, h, L + r, &dw, &dh);
      f4* cur = (f4*)(chain + offs[L + r] * 4u);
      if(steps[s].pipeline == ORC_PIPE_FAST)
      {
        int order = fast_order(count, r);
        for(uint32_t y = 0; y < dh; ++y)
          for(uint32_t x = 0; x < dw; ++x)
          {
            const f4* p = prev + (uint64_t)(2u * y) * sw + 2u * x;
            cur[(uint64_t)y * dw + x] = box_f32(p[0], p[1], p[sw], p[sw + 1], order);
          }
      }
      else
      {
        int kx = ksize(sw), ky = ksize(sh);
        for(uint32_t y = 0; y < dh; ++y)
          for(uint32_t x = 0; x < dw; ++x)
            cur[(uint64_t)y * dw + x] = general_f32(prev, sw, dw, dh, x, y, kx, ky);
      }
      prev = cur;
      sw   = dw;
      sh   = dh;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* RGBA16F: IEEE binary16 texels                                              */
/* ------------------------------------------------------------------------ */

/* binary16 -> double, exact. */
double orc_half_to_double(uint16_t h)
{
  const int sign = h >> 15, exp = (h >> 10) & 31, man = h & 1023;
  double v;
  if(exp == 0)
    v = ldexp((double)man, -24);
  else if(exp == 31)
    v = man ? NAN : INFINITY;
  else
    v = ldexp((double)(man | 1024), exp - 25);
  return sign ? -v : v;
}

/* double -> binary16, round to nearest even, with subnormals and overflow to inf. */
uint16_t orc_half_from_double(double x)
{
  if(isnan(x))
    return 0x7e00;
  uint16_t sign = signbit(x) ? 0x8000 : 0;
  double a = fabs(x);
  if(a >= 65520.0) /* rounds to inf */
    return sign | 0x7c00;
  if(a < ldexp(1.0, -25)) /* below half the smallest subnormal (ties to even: 0) */
    return sign;
  int e;
  frexp(a, &e); /* a = m * 2^e, m in [0.5, 1) */
  int q = e - 11; /* quantum exponent for an 11-bit significand */
  if(q < -24)
    q = -24; /* subnormal quantum */
  double scaled = ldexp(a, -q); /* exact: power-of-two scaling */
  double r      = nearbyint(scaled); /* default rounding mode: to nearest even */
  uint32_t m    = (uint32_t)r;
  if(q == -24 && m < 1024)
    return sign | (uint16_t)m; /* subnormal (or rounds up to the smallest normal below) */
  if(m == 2048)
  {
    m = 1024;
    q += 1;
  }
  int biased = q + 25; /* m in [1024, 2047], value = m * 2^q */
  if(biased >= 31)
    return sign | 0x7c00;
  return sign | (uint16_t)(biased << 10) | (uint16_t)(m & 1023);
}

float orc_half_to_float(uint16_t h) { return (float)orc_half_to_double(h); }

/* A binary16 operation: exact in double (two halves' sum, product with 0.25 or
 * 0.5, are exact in double), then one rounding to half. */
static double h16(double x) { return orc_half_to_double(orc_half_from_double(x)); }

typedef struct { double c[4]; } h4; /* values that are exactly representable halves */

static h4 box_h16(h4 ul, h4 ur, h4 ll, h4 lr, int order)
{
  h4 o;
  for(int c = 0; c < 4; ++c)
  {
    double s = order == 0 ? h16(h16(ul.c[c] + ll.c[c]) + h16(ur.c[c] + lr.c[c]))
                          : h16(h16(ul.c[c] + ur.c[c]) + h16(ll.c[c] + lr.c[c]));
    o.c[c] = h16(s * 0.25);
  }
  return o;
}

/* reduceStoreSample_ in PACKED_FP16 arithmetic: 2-tap stages in half precision,
 * 3-tap stages widened to FP32 (weights as in general_f32), rounded to half. */
static h4 general_h16(const h4* src, uint32_t sw, uint32_t dw, uint32_t dh, uint32_t x, uint32_t y, int kx, int ky)
{
  float n   = (float)dh;
  float rcp = 1.0f / (2.0f * n + 1.0f);
  float w0  = rcp * (n - (float)y);
  float w1  = rcp * n;
  float w2  = (1.0f - w0) - w1;
  h4 hcol[3];
  for(int j = 0; j < kx; ++j)
  {
    const h4* col = src + (uint64_t)(2u * y) * sw + 2u * x + (uint32_t)j;
    for(int c = 0; c < 4; ++c)
    {
      if(ky == 3)
        hcol[j].c[c] = h16((double)reduce3_f32(w0, (float)col[0].c[c], w1, (float)col[sw].c[c], w2,
                                               (float)col[2u * sw].c[c]));
      else if(ky == 2)
        hcol[j].c[c] = h16(h16(col[0].c[c] + col[sw].c[c]) * 0.5);
      else
        hcol[j].c[c] = col[0].c[c];
    }
  }
  h4 o;
  if(kx == 3)
  {
    n   = (float)dw;
    rcp = 1.0f / (2.0f * n + 1.0f);
    w0  = rcp * (n - (float)x);
    w1  = rcp * n;
    w2  = (1.0f - w0) - w1;
    for(int c = 0; c < 4; ++c)
      o.c[c] = h16((double)reduce3_f32(w0, (float)hcol[0].c[c], w1, (float)hcol[1].c[c], w2, (float)hcol[2].c[c]));
  }
  else if(kx == 2)
  {
    for(int c = 0; c < 4; ++c)
      o.c[c] = h16(h16(hcol[0].c[c] + hcol[1].c[c]) * 0.5);
  }
  else
    o = hcol[0];
  return o;
}

int orc_pyramid_rgba16f(uint16_t* chain, uint32_t w, uint32_t h, uint32_t mipLevels, const orc_step* steps,
                        uint32_t nSteps, int packedFp16)
{
  if(mipLevels == 0)
    mipLevels = orc_full_levels(w, h);
  if(check_plan(w, h, mipLevels, steps, nSteps))
    return -1;
  uint64_t offs[32];
  orc_packed_offsets(w, h, mipLevels, offs);
  for(uint32_t s = 0; s < nSteps; ++s)
  {
    uint32_t L = steps[s].inputLevel, count = steps[s].levelCount;
    uint32_t sw, sh;
    orc_level_dims(w, h, L, &sw, &sh);
    const uint64_t n0 = (uint64_t)sw * sh;
    f4* prevf         = packedFp16 ? NULL : (f4*)malloc(sizeof(f4) * n0);
    h4* prevh         = packedFp16 ? (h4*)malloc(sizeof(h4) * n0) : NULL;
    const uint16_t* in = chain + offs[L] * 4u;
    for(uint64_t i = 0; i < n0; ++i)
      for(int c = 0; c < 4; ++c)
      {
        if(packedFp16)
          prevh[i].c[c] = orc_half_to_double(in[4 * i + c]);
        else
          prevf[i].c[c] = orc_half_to_float(in[4 * i + c]);
      }
    for(uint32_t r = 1; r <= count; ++r)
    {
      uint32_t dw, dh;
      orc_level_dims(w, h, L + r, &dw, &dh);
      const uint64_t n1 = (uint64_t)dw * dh;
      uint16_t* out     = chain + offs[L + r] * 4u;
      const int kx = ksize(sw), ky = ksize(sh);
      if(packedFp16)
      {
        h4* cur = (h4*)malloc(sizeof(h4) * n1);
        for(uint32_t y = 0; y < dh; ++y)
          for(uint32_t x = 0; x < dw; ++x)
          {
            h4 v;
            if(steps[s].pipeline == ORC_PIPE_FAST)
            {
              const h4* p = prevh + (uint64_t)(2u * y) * sw + 2u * x;
              v           = box_h16(p[0], p[1], p[sw], p[sw + 1], fast_order(count, r));
            }
            else
              v = general_h16(prevh, sw, dw, dh, x, y, kx, ky);
            cur[(uint64_t)y * dw + x] = v;
            for(int c = 0; c < 4; ++c)
              out[4 * ((uint64_t)y * dw + x) + c] = orc_half_from_double(v.c[c]);
          }
        free(prevh);
        prevh = cur;
      }
      else
      {
        f4* cur = (f4*)malloc(sizeof(f4) * n1);
        for(uint32_t y = 0; y < dh; ++y)
          for(uint32_t x = 0; x < dw; ++x)
          {
            f4 v;
            if(steps[s].pipeline == ORC_PIPE_FAST)
            {
              const f4* p = prevf + (uint64_t)(2u * y) * sw + 2u * x;
              v           = box_f32(p[0], p[1], p[sw], p[sw + 1], fast_order(count, r));
            }
            else
              v = general_f32(prevf, sw, dw, dh, x, y, kx, ky);
            cur[(uint64_t)y * dw + x] = v;
            for(int c = 0; c < 4; ++c)
              out[4 * ((uint64_t)y * dw + x) + c] = orc_half_from_double((double)v.c[c]);
          }
        free(prevf);
        prevf = cur;
      }
      sw = dw;
      sh = dh;
    }
    free(prevf);
    free(prevh);
  }
  return 0;
}
