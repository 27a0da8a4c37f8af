/*
 * nvpro_pyramid_hip.h — C-ABI of the MI355X (gfx950) mip-chain generator.
 *
 * Drop-in for the nvpro_pyramid host dispatcher of nvpro-samples/vk_compute_mipmaps.
 * Each entry point below names the reference interface it replaces (file:line in
 * that repository).  The library records nothing into a Vulkan command buffer;
 * it launches hand-written HIP kernels on the caller's stream instead, making
 * exactly the level-splitting decisions of nvproCmdPyramidDispatch.
 *
 * Conventions
 *  - Every entry point that can fail returns an int status: 0 (hipSuccess) or a
 *    hipError_t value (hipErrorInvalidValue for bad arguments).  A message for the
 *    last failure on the calling thread is available from nvproPyramidGetLastError().
 *  - No entry point synchronises the host with the device.  Work is stream
 *    ordered; the caller orders work before and after (as the reference leaves
 *    barriers before and after to the caller, nvpro_pyramid_dispatch.hpp:48-53).
 *  - No entry point of the plain path allocates device memory.
 *  - Images live in linear device memory.  By default the levels are packed
 *    back to back, row-major, tight rows: the MipmapStorage layout of
 *    include/mipmap_storage.hpp:53-76.
 */
#ifndef NVPRO_PYRAMID_HIP_H_
#define NVPRO_PYRAMID_HIP_H_

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Push-constant layout: pc = inputLevel << 5 | levelCount.
 * nvpro_pyramid_dispatch.hpp:77, nvpro_pyramid.glsl:157-162. */
#define NVPRO_PYRAMID_INPUT_LEVEL_SHIFT 5u
#define NVPRO_PYRAMID_LEVEL_COUNT_MASK 31u

/* Maximum mip levels of an image: the sRGBA8 functor's imageMipLevels[16]
 * (srgba8_mipmap_preamble.glsl:13, scoped_image.hpp:47).  Edges < 65536. */
#define NVPRO_PYRAMID_MAX_LEVELS 16u

/* Pipelines of nvpro_pyramid.glsl (NVPRO_PYRAMID_IS_FAST_PIPELINE 0 / 1). */
#define NVPRO_PYRAMID_PIPELINE_GENERAL 0u
#define NVPRO_PYRAMID_PIPELINE_FAST 1u

/* Texel formats: each selects one compiled set of load/reduce/store functors,
 * the counterpart of the NVPRO_PYRAMID_{TYPE,LOAD,REDUCE,STORE} macros
 * (nvpro_pyramid.glsl:27-120).
 *  SRGBA8  : the sRGBA8 functors of srgba8_mipmap_preamble.glsl:19-128
 *            (4 bytes/texel, sRGB-encoded RGB, linear alpha, FP32 arithmetic).
 *  RGBA16F : linear half4 texels (8 bytes/texel).
 *  RGBA32F : linear float4 texels (16 bytes/texel). */
#define NVPRO_PYRAMID_FORMAT_SRGBA8 0u
#define NVPRO_PYRAMID_FORMAT_RGBA16F 1u
#define NVPRO_PYRAMID_FORMAT_RGBA32F 2u

/* Arithmetic of the RGBA16F functors.
 *  FP32        : loads widen to FP32, all reductions in FP32.
 *  PACKED_FP16 : 2x2 box reductions in packed FP16, 3-tap weights in FP32. */
#define NVPRO_PYRAMID_ARITH_FP32 0u
#define NVPRO_PYRAMID_ARITH_PACKED_FP16 1u

/* Mirrors struct NvproPyramidPipelines (nvpro_pyramid_dispatch.hpp:29-35).
 * The general pipeline is always present.  fastPipeline == 0 is the
 * VK_NULL_HANDLE case: the fast dispatcher is never tried (:147), which is also
 * what minimal_mipmaps.cpp's -force-no-fast-pipeline produces (:266-270). */
typedef struct NvproPyramidPipelines {
  uint32_t format;       /* NVPRO_PYRAMID_FORMAT_* */
  uint32_t fastPipeline; /* nonzero: fast pipeline available */
  uint32_t arithmetic;   /* NVPRO_PYRAMID_ARITH_* (RGBA16F only; else 0) */
  uint32_t reserved;     /* must be 0 */
} NvproPyramidPipelines;

/* The image the pipelines read and write; replaces the descriptor sets the
 * caller binds before nvproCmdPyramidDispatch (srgba8_mipmap_preamble.glsl:10-13,
 * minimal_mipmaps.cpp:201-206).  Level 0 is input; levels 1.. are written. */
typedef struct NvproPyramidImage {
  void* data;                   /* device pointer; byte offsets are relative to it */
  uint32_t width, height;       /* level 0 size */
  uint32_t levelCount;          /* levels present in storage; 0 = full chain */
  uint32_t reserved;            /* must be 0 */
  const uint64_t* levelOffsets; /* host array [levelCount] of byte offsets, or NULL for
                                   the packed MipmapStorage layout */
  const uint32_t* rowPitches;   /* host array [levelCount] of row pitches in bytes, or
                                   NULL for tight rows */
} NvproPyramidImage;

/* Mirrors struct NvproPyramidState (nvpro_pyramid_dispatch.hpp:63-75). */
typedef struct NvproPyramidState {
  uint32_t currentLevel;    /* input level of the next dispatch */
  uint32_t remainingLevels; /* levels still to fill; never 0 when passed to a dispatcher */
  uint32_t currentX, currentY;
} NvproPyramidState;

/* One recorded dispatch.  pipeline/pushConstant/groupCountX are exactly what
 * the reference records (vkCmdBindPipeline, vkCmdPushConstants, vkCmdDispatch). */
typedef struct NvproPyramidStep {
  uint32_t pipeline;     /* NVPRO_PYRAMID_PIPELINE_* */
  uint32_t inputLevel;   /* pushConstant >> 5 */
  uint32_t levelCount;   /* pushConstant & 31 */
  uint32_t pushConstant;
  uint32_t groupCountX;  /* workgroups the reference dispatches for this step */
  uint32_t inputWidth, inputHeight;
  uint32_t reserved;
} NvproPyramidStep;

/* The "command buffer" a dispatcher records into (VkCommandBuffer analogue). */
typedef struct NvproPyramidCmd NvproPyramidCmd;

/* Mirrors nvpro_pyramid_dispatcher_t (nvpro_pyramid_dispatch.hpp:80-104): record
 * at most one dispatch for `state` and return the levels it fills.  A fast
 * dispatcher may return 0 (declined); a general dispatcher must not. */
typedef uint32_t (*nvpro_pyramid_dispatcher_t)(NvproPyramidCmd* cmd, const NvproPyramidState* state);

/* Bind + push constant + dispatch, as a dispatcher does with vkCmdBindPipeline,
 * vkCmdPushConstants and vkCmdDispatch (nvpro_pyramid_dispatch.hpp:211-238,
 * :260-290).  groupCountX is the reference's workgroup count, recorded for
 * reporting; the HIP kernels choose their own launch shape.  Returns 0 or an
 * error (bad pipeline, level range outside the image, fast divisibility unmet). */
int nvproPyramidCmdDispatch(NvproPyramidCmd* cmd, uint32_t pipeline, uint32_t pushConstant,
                            uint32_t groupCountX);

/* nvproPyramidDefaultFastDispatcher<4, 6> (nvpro_pyramid_dispatch.hpp:195-242). */
uint32_t nvproPyramidDefaultFastDispatcher(NvproPyramidCmd* cmd, const NvproPyramidState* state);

/* The template itself: nvproPyramidDefaultFastDispatcher<divisibility, maxLevels>
 * (nvpro_pyramid_dispatch.hpp:195-208, :218-228).  divisibility must be even and
 * nonzero, maxLevels in 1..6; returns 0 otherwise. */
uint32_t nvproPyramidFastDispatcherEx(NvproPyramidCmd* cmd, const NvproPyramidState* state,
                                      uint32_t divisibility, uint32_t maxLevels);

/* nvproPyramidDefaultGeneralDispatcher (nvpro_pyramid_dispatch.hpp:247-292). */
uint32_t nvproPyramidDefaultGeneralDispatcher(NvproPyramidCmd* cmd, const NvproPyramidState* state);

/* nvproCmdPyramidDispatch(cmdBuf, pipelines, baseWidth, baseHeight, mipLevels = 0)
 * (nvpro_pyramid_dispatch.hpp:54-59, :294-304).  Fills levels 1..mipLevels-1 of
 * `image` on `stream`.  mipLevels 0 means the full chain
 * (floor(log2(max(w, h))) + 1, :122-132); a value above the image's level count
 * is an error.  mipLevels 1 fills nothing and returns 0 (the reference asserts
 * there, :169). */
int nvproPyramidDispatch(hipStream_t stream, const NvproPyramidPipelines* pipelines,
                         const NvproPyramidImage* image, uint32_t mipLevels);

/* The overload with explicit dispatchers (nvpro_pyramid_dispatch.hpp:109-188).
 * fastDispatcher is tried first when pipelines->fastPipeline is nonzero; when it
 * returns 0 the general dispatcher runs.  A general dispatcher returning 0, or
 * any dispatcher returning more than the remaining levels, is an error
 * (the reference's asserts, :169, :172); nothing is launched in that case. */
int nvproPyramidDispatchCustom(hipStream_t stream, const NvproPyramidPipelines* pipelines,
                               const NvproPyramidImage* image, uint32_t mipLevels,
                               nvpro_pyramid_dispatcher_t generalDispatcher,
                               nvpro_pyramid_dispatcher_t fastDispatcher);

/* Host-only planning: runs the same dispatcher loop and returns the dispatches
 * instead of launching them.  Null dispatchers select the defaults.  *stepCount
 * receives the number of steps even when it exceeds `capacity`. */
int nvproPyramidPlan(const NvproPyramidPipelines* pipelines, uint32_t width, uint32_t height,
                     uint32_t mipLevels, nvpro_pyramid_dispatcher_t generalDispatcher,
                     nvpro_pyramid_dispatcher_t fastDispatcher, NvproPyramidStep* steps,
                     uint32_t capacity, uint32_t* stepCount);

/* floor(log2(max(w, h))) + 1 (nvpro_pyramid_dispatch.hpp:122-132). */
uint32_t nvproPyramidFullLevelCount(uint32_t width, uint32_t height);

/* Packed MipmapStorage layout (mipmap_storage.hpp:53-76): byte offset of each of
 * `levelCount` levels (0 = full chain) and the total size.  Either output may be
 * NULL.  offsets must hold levelCount entries. */
int nvproPyramidPackedLayout(uint32_t width, uint32_t height, uint32_t levelCount,
                             uint32_t bytesPerTexel, uint64_t* offsets, uint64_t* totalBytes);

/* Bytes per texel of a format, 0 if unknown. */
uint32_t nvproPyramidBytesPerTexel(uint32_t format);

/* Message for the last failed call on this thread ("" if none). */
const char* nvproPyramidGetLastError(void);

#ifdef __cplusplus
}
#endif

#endif /* NVPRO_PYRAMID_HIP_H_ */
