/*
 * orc.h — CPU oracle for the nvpro_pyramid mip-chain path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark do, and only as the checker.
 *
 * It restates, in plain C, the algorithms of nvpro-samples/vk_compute_mipmaps
 * that the product must reproduce:
 *   - the dispatch planner    nvpro_pyramid/nvpro_pyramid_dispatch.hpp:109-292
 *   - the sRGB conversions    shaders/srgb.h:18-41, nvpro_pyramid/srgba8_mipmap_preamble.glsl:110-128
 *   - the GPU schedule's arithmetic (FP32, per-stage summation order)
 *                             nvpro_pyramid/nvpro_pyramid.glsl:272-532, :557-882
 *   - an FP64 filter oracle with the same quantisation points
 *   - the reference's own CPU generator and comparer
 *                             include/mipmap_storage.hpp:53-76, :159-202, :246-414
 *
 * Arithmetic the reference leaves to the Vulkan implementation is fixed here as
 * follows (each choice is inside the reference's implementation-defined range):
 *   - sRGB decode of a code c: FP32 rounding of the exact value of srgb.h:18-28.
 *   - alpha in code units: the reference decodes c / 255 (srgb.h:55-61) and
 *     encodes uint(a * 255 + 0.5) (preamble:123-128), a linear scale that
 *     commutes with every reduction; here the value is c itself and the encode
 *     floor(a + 0.5) clamped, so box filters of alpha are exact and .5 ties
 *     round up as in exact arithmetic.
 *   - sRGB encode: the exact real value of preamble:110-121 (no pow error).
 *   - 3-tap REDUCE (preamble:24-25): fmaf(w2, v2, fmaf(w1, v1, w0 * v0)).
 *   - reciprocal 1 / (2n + 1) (glsl:583): IEEE division.
 *   - LOAD_REDUCE4 without the bilinear sampler (preamble:40 with
 *     USE_BILINEAR_SAMPLING=0): 4 loads, REDUCE4 of glsl:179-188.
 */
#ifndef ORC_H_
#define ORC_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_PIPE_GENERAL 0u
#define ORC_PIPE_FAST 1u

typedef struct orc_step {
  uint32_t pipeline, inputLevel, levelCount, pushConstant, groupCountX, inputWidth, inputHeight,
      reserved;
} orc_step;

/* Planted mistakes for tests of the tests: flip every fast-pipeline summation
 * order, or use the reference CPU's w2 = rcp * (y + 1) in the FP32 emulator. */
#define ORC_DEBUG_FLIP_ORDER 1
#define ORC_DEBUG_CPU_WEIGHTS 2
extern int orc_debug_flags;

/* ---- layout ---- */
uint32_t orc_full_levels(uint32_t w, uint32_t h);
void orc_level_dims(uint32_t w, uint32_t h, uint32_t level, uint32_t* lw, uint32_t* lh);
/* Offsets in texels of `levels` packed levels; returns the total texel count. */
uint64_t orc_packed_offsets(uint32_t w, uint32_t h, uint32_t levels, uint64_t* offsets);

/* ---- planner ---- */
/* Returns the number of steps (may exceed cap; only cap are written), or -1 if
 * a dispatcher would violate the reference's asserts. */
int orc_plan(uint32_t w, uint32_t h, uint32_t mipLevels, int fastAvailable, uint32_t fastDivisibility,
             uint32_t fastMaxLevels, orc_step* out, uint32_t cap);

/* ---- sRGB ---- */
double orc_srgb_decode_f64(uint32_t code);
float orc_srgb_decode_f32(uint32_t code);
float orc_alpha_decode_f32(uint32_t code);
uint32_t orc_srgb_encode_exact(double x);
void orc_premultiply_srgba8(uint8_t* texels, uint64_t count, int exact);
uint32_t orc_alpha_encode_f32(float a);
uint32_t orc_alpha_encode_f64(double a);
/* Smallest float whose exact encode is >= k, for k in 1..255. */
float orc_srgb_threshold(uint32_t k);

/* ---- schedule emulators over packed sRGBA8 chains (level 0 filled) ---- */
/* FP32, reference summation order per stage; steps from orc_plan. */
int orc_pyramid_srgba8_f32(uint8_t* chain, uint32_t w, uint32_t h, uint32_t mipLevels,
                           const orc_step* steps, uint32_t nSteps);
/* FP64 filter; quantise to sRGBA8 at every stored level, carry FP64 inside a step. */
int orc_pyramid_srgba8_f64(uint8_t* chain, uint32_t w, uint32_t h, uint32_t mipLevels,
                           const orc_step* steps, uint32_t nSteps);
/* FP64 filter on linear float4 texels with no quantisation at all (levels packed,
 * 4 doubles per texel): used for energy-conservation and box-filter checks. */
void orc_pyramid_linear_f64(double* chain, uint32_t w, uint32_t h, uint32_t mipLevels);

/* The reference CPU generator: cpuGenerateMipmaps_sRGBA (mipmap_storage.hpp:395-414). */
void orc_reference_cpu_srgba8(uint8_t* chain, uint32_t w, uint32_t h);
/* Same, levels computed by `threads` threads splitting rows (0 = 1 thread). */
void orc_reference_cpu_srgba8_mt(uint8_t* chain, uint32_t w, uint32_t h, uint32_t threads);

/* MipmapStorage::compare (mipmap_storage.hpp:159-202): worst |a-b| over levels >= 1.
 * Also counts differing channels / texels and returns them through the pointers. */
uint32_t orc_compare_srgba8(const uint8_t* a, const uint8_t* b, uint32_t w, uint32_t h,
                            uint32_t levels, uint32_t* worstX, uint32_t* worstY, uint32_t* worstLevel,
                            uint32_t* worstChannel, uint64_t* diffTexels, uint64_t* totalTexels);

/* ---- RGBA16F / RGBA32F emulators (linear float texels) ---- */
/* FP32 emulation of the float functors; texels are 4 floats.  Same order as sRGBA8. */
int orc_pyramid_rgba32f_f32(float* chain, uint32_t w, uint32_t h, uint32_t mipLevels,
                            const orc_step* steps, uint32_t nSteps);

/* binary16 conversions (round to nearest even, subnormals, overflow to inf). */
double orc_half_to_double(uint16_t h);
float orc_half_to_float(uint16_t h);
uint16_t orc_half_from_double(double x);
/* RGBA16F chains: FP32 arithmetic (packedFp16 = 0; loads widen, stores round)
 * or PACKED_FP16 arithmetic (2-tap / 2x2 stages in binary16, 3-tap stages in
 * FP32 then rounded to binary16). */
int orc_pyramid_rgba16f(uint16_t* chain, uint32_t w, uint32_t h, uint32_t mipLevels, const orc_step* steps,
                        uint32_t nSteps, int packedFp16);

#ifdef __cplusplus
}
#endif
#endif
