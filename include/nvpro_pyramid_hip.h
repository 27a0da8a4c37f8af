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
 * there, :169).
 *
 * Launches: one per step of the plan, except that the small end of a plan runs
 * in fewer: the steps from the first whose first output level has at most 1024
 * texels run in one launch, and a fast step (of at most 64 or at least 1024
 * workgroups) whose successor is a fast step of one workgroup, or the rest of
 * the plan with at most 128 texels in its first output level, runs that in its
 * own launch, in whichever workgroup finishes last.  Results are the same bytes
 * either way.  This fusion needs a 16.6 KB block of arrival counters per
 * stream, taken from a static device array of 64 blocks per device (nothing is
 * allocated; a stream that fuses keeps its block until nvproPyramidStreamRelease
 * or the end of the process, and while 64 streams of a device hold one, others
 * do not fuse); it is off while the stream is being
 * captured into a graph, and when the environment variable NVPRO_PYRAMID_FUSE
 * is "0".  NVPRO_PYRAMID_THIN=1 / =0 forces the general steps' thin (31x3) or
 * full (31x7) tiles; by default a step runs thin tiles when all of them fit on
 * the GPU at once.  Neither changes the bytes written. */
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

/* Kernel launches made by this thread's last successful nvproPyramidDispatch or
 * nvproPyramidDispatchCustom (a diagnostic with no reference counterpart: it
 * shows whether a dispatch was fused). */
uint32_t nvproPyramidLastLaunchCount(void);

/* Fused dispatches (a step and its small successor in one launch) count
 * arrivals in a per-stream block of a static device array with 64 blocks per
 * device; a stream keeps its block, and streams beyond the 64th dispatch
 * unfused (same bytes, one more launch).  A caller that creates and destroys
 * streams calls this before destroying one, once it enqueues no more work on
 * it: it waits for the stream's work and hands its block to later streams.
 * Other threads may dispatch meanwhile.  No reference counterpart (Vulkan
 * command buffers carry no such state). */
int nvproPyramidStreamRelease(hipStream_t stream);

/* Workload generator (no reference counterpart; the reference's benchmark loads
 * image files instead, demo_app/mipmaps_app.cpp:556-879): writes the seeded
 * synthetic level 0 of `image` on `stream` (SRGBA8: smooth value noise with a
 * 64-px period plus uniform noise, uniform alpha; RGBA16F: log-uniform radiance
 * in [1e-3, 1e3), uniform alpha), bit-identical to vk_compute_mipmaps_amd/synth.py. */
int nvproPyramidSynthesize(hipStream_t stream, uint32_t format, const NvproPyramidImage* image, uint64_t seed);

/* Premultiplied alpha on level 0, in place (the staging step of
 * include/scoped_image.hpp:233-252, "-do-premultiply-alpha" of
 * minimal_mipmaps.cpp:316-367, moved to the device).  SRGBA8: alpha = a/255,
 * rgb = srgbFromLinear(linearFromSrgb(c) * alpha) with the exact conversions of
 * the sRGBA8 functor; RGBA16F: rgb * a in FP32, rounded to binary16. */
int nvproPyramidPremultiplyAlpha(hipStream_t stream, uint32_t format, const NvproPyramidImage* image);

/* writeMipmapsTga (include/mipmap_storage.hpp:441-479): every level of a packed
 * sRGBA8 chain in host memory to its own TGA file, level 0 to baseFilename and
 * level l to baseFilename with "l." inserted before the extension
 * ("out.tga" -> "out.1.tga").  Uncompressed 32-bit, top-left origin.
 * levelCount 0: full chain. */
int nvproPyramidWriteTga(const void* hostChain, uint32_t width, uint32_t height, uint32_t levelCount,
                         const char* baseFilename);

/* Reads a 24/32-bit true-colour TGA (uncompressed or run-length encoded, either
 * origin) into tight RGBA8 rows.  rgba NULL: only *width and *height. */
int nvproPyramidReadTga(const char* filename, void* rgba, uint64_t capacity, uint32_t* width, uint32_t* height);

/* ------------------------------------------------------------------------- */
/* Batched chains.  No reference counterpart: the reference loops over images  */
/* and calls nvproCmdPyramidDispatch once per image (demo_app/mipmaps_app.cpp: */
/* 672-722).  A batch runs the same per-image plans (default dispatchers), so   */
/* every image's levels are identical to what nvproPyramidDispatch writes, but  */
/* the steps of all images go through one persistent work-queue launch (QUEUE)  */
/* or one launch per round of independent steps (ROUNDS).                       */
/* ------------------------------------------------------------------------- */

#define NVPRO_PYRAMID_BATCH_MODE_ROUNDS 0u
#define NVPRO_PYRAMID_BATCH_MODE_QUEUE 1u

/* Error flags reported by nvproPyramidBatchReadErrors. */
#define NVPRO_PYRAMID_BATCH_ERROR_SPIN 1u /* a queue dependency wait gave up (results invalid) */

typedef struct NvproPyramidBatchConfig {
  uint32_t mode;            /* NVPRO_PYRAMID_BATCH_MODE_* */
  uint32_t workgroupsPerCU; /* persistent workgroups per compute unit; 0: as many as fit */
  uint32_t itemBytes;       /* input bytes per work item (a run of groups of one step); 0: 65536 */
  uint32_t tailTexels;      /* a step whose first output level has at most this many texels, and
                               the rest of its image's plan, run in one workgroup; 0: never;
                               at most 128 */
  uint32_t useGraph;        /* nonzero: nvproPyramidBatchLaunch replays a hipGraph built at creation */
  uint32_t spinLimit;       /* queue mode: polls (each >= 50 ns) before a dependency wait gives up and
                               sets NVPRO_PYRAMID_BATCH_ERROR_SPIN; 0: 2^20.  Small values are for
                               tests of the error path. */
  uint32_t reserved[2];     /* must be 0 */
} NvproPyramidBatchConfig;

typedef struct NvproPyramidBatchInfo {
  uint32_t imageCount, segmentCount, itemCount, roundCount;
  uint32_t launchCount;      /* kernel launches per nvproPyramidBatchLaunch */
  uint32_t workgroups;       /* workgroups of the largest launch */
  uint64_t algorithmicBytes; /* every base level read once + every output level written once */
  uint64_t deviceBytes;      /* the batch's device work-list size */
  NvproPyramidBatchConfig config;
} NvproPyramidBatchInfo;

typedef struct NvproPyramidBatch NvproPyramidBatch;

/* Defaults: QUEUE, as many workgroups as fit, 64 KiB items, 128-texel tails, graph on. */
int nvproPyramidBatchDefaultConfig(NvproPyramidBatchConfig* config);

/* Plans `imageCount` images of one format (mipLevels: per-image level counts,
 * NULL or 0 entries for full chains), allocates and uploads the work list
 * (synchronously, on a private stream) and, with useGraph, builds the graph.
 * config NULL: defaults.  The images must stay allocated while the batch is
 * launched; their contents may change between launches. */
int nvproPyramidBatchCreate(const NvproPyramidPipelines* pipelines, const NvproPyramidImage* images,
                            const uint32_t* mipLevels, uint32_t imageCount,
                            const NvproPyramidBatchConfig* config, NvproPyramidBatch** batch);

/* Enqueues the whole batch on `stream`.  Launches of one batch object must not
 * overlap in time (they share its counters): order them on one stream. */
int nvproPyramidBatchLaunch(NvproPyramidBatch* batch, hipStream_t stream);

int nvproPyramidBatchGetInfo(const NvproPyramidBatch* batch, NvproPyramidBatchInfo* info);

/* Synchronises with the device, returns the batch's sticky error flags and clears them. */
int nvproPyramidBatchReadErrors(NvproPyramidBatch* batch, uint32_t* flags);

int nvproPyramidBatchDestroy(NvproPyramidBatch* batch);

/* Project autotuner: times each candidate configuration (mode, workgroups per
 * CU, item size, tail size; graph on) over `repetitions` launches of these
 * images on `stream` and returns the fastest.  It writes the images' output
 * levels (with the same values every time).  Results are cached for the life
 * of the process per device, format, arithmetic and exact list of image sizes
 * and level counts. */
int nvproPyramidBatchAutotune(const NvproPyramidPipelines* pipelines, const NvproPyramidImage* images,
                              const uint32_t* mipLevels, uint32_t imageCount, hipStream_t stream,
                              uint32_t repetitions, NvproPyramidBatchConfig* best, float* bestMicroseconds);


/* ------------------------------------------------------------------------- */
/* Batched chains without allocation.  No counterpart in the reference.  For
 * callers that cannot allocate device or pinned memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc): one batch in
 * ROUNDS mode, planned and launched in one call, with exactly the launches
 * nvproPyramidBatchCreate + nvproPyramidBatchLaunch make for that
 * configuration (so the same bytes).  The work list goes to a static
 * per-device scratch array of NVPRO_PYRAMID_NO_ALLOC_SCRATCH_BYTES (each GPU
 * the library runs on holds it until the process ends) through one
 * host-to-device copy from ordinary host memory on `stream`; a batch whose work
 * list is larger is rejected with hipErrorInvalidValue before anything is
 * enqueued (nvproPyramidGetLastError names the size).  Every call on a GPU
 * shares that GPU's array, which keeps the last call's work list: calls that
 * overlap in time on one GPU race, so order them on one stream.  config NULL:
 * defaults with mode ROUNDS; mode must be ROUNDS; useGraph is ignored. */
#define NVPRO_PYRAMID_NO_ALLOC_SCRATCH_BYTES (4u << 20)

int nvproPyramidBatchDispatchNoAlloc(hipStream_t stream, const NvproPyramidPipelines* pipelines,
                                     const NvproPyramidImage* images, const uint32_t* mipLevels,
                                     uint32_t imageCount, const NvproPyramidBatchConfig* config);

/* ------------------------------------------------------------------------- */
/* Serving loop.  No reference counterpart (the reference regenerates one       */
/* image's chain per frame, demo_app/mipmaps_app.cpp:362-369).  A server admits */
/* mip-chain requests against an HBM budget, keeps each chain resident in its   */
/* arena until released, and batches requests as they arrive: every poll       */
/* launches the admitted requests (one batch per format, on a free in-flight    */
/* slot with its own stream) while earlier batches are still running.  Each     */
/* chain is planned as nvproPyramidDispatch plans it, so a result is the same   */
/* whatever it was batched with.  All entry points are thread-safe.             */
/* ------------------------------------------------------------------------- */

typedef struct NvproPyramidServerConfig {
  uint64_t budgetBytes;         /* device arena for resident chains; default 80 GB */
  uint32_t maxInFlight;         /* batches on the device at once (1..16); default 4 */
  uint32_t maxBatchImages;      /* images per batch; default 1024 */
  uint32_t arithmetic16f;       /* NVPRO_PYRAMID_ARITH_* of RGBA16F requests; default FP32 */
  uint32_t reserved[2];         /* must be 0 */
  NvproPyramidBatchConfig batch; /* work-list configuration (default: rounds mode); useGraph must be 0 */
} NvproPyramidServerConfig;

typedef struct NvproPyramidRequest {
  uint32_t format;     /* NVPRO_PYRAMID_FORMAT_SRGBA8 or _RGBA16F */
  uint32_t width, height;
  uint32_t mipLevels;  /* 0: full chain */
  uint64_t seed;       /* level 0 is the seeded synthetic image (nvproPyramidSynthesize) ... */
  const void* level0;  /* ... unless this points at tight host texels, copied at launch; they must
                          stay valid until the request completes */
} NvproPyramidRequest;

typedef struct NvproPyramidCompletion {
  uint64_t ticket;
  int status;               /* 0 or an error */
  uint32_t format;
  NvproPyramidImage image;  /* the resident chain: packed layout in the server's arena */
  uint64_t bytes;
  double latencyUs;         /* submit to completion seen by a poll */
  double waitUs;            /* submit to admission */
} NvproPyramidCompletion;

typedef struct NvproPyramidServerStats {
  uint64_t submitted, admitted, completed, released, batches, batchedImages, failedBatches;
  uint64_t budgetBytes, residentBytes, peakResidentBytes;
  uint32_t waiting, inFlightBatches;
} NvproPyramidServerStats;

typedef struct NvproPyramidServer NvproPyramidServer;

int nvproPyramidServerDefaultConfig(NvproPyramidServerConfig* config);
/* Allocates the arena and the in-flight slots.  config NULL: defaults. */
int nvproPyramidServerCreate(const NvproPyramidServerConfig* config, NvproPyramidServer** server);
/* Queues a request; hipErrorOutOfMemory if its chain can never fit the budget. */
int nvproPyramidServerSubmit(NvproPyramidServer* server, const NvproPyramidRequest* request, uint64_t* ticket);
/* Makes progress (retire, admit, launch) and returns up to `capacity` completions. */
int nvproPyramidServerPoll(NvproPyramidServer* server, NvproPyramidCompletion* completions, uint32_t capacity,
                           uint32_t* count);
/* Frees a completed request's resident chain. */
int nvproPyramidServerRelease(NvproPyramidServer* server, uint64_t ticket);
int nvproPyramidServerGetStats(NvproPyramidServer* server, NvproPyramidServerStats* stats);
/* Waits for in-flight batches and frees everything. */
int nvproPyramidServerDestroy(NvproPyramidServer* server);

#ifdef __cplusplus
}
#endif

#endif /* NVPRO_PYRAMID_HIP_H_ */
