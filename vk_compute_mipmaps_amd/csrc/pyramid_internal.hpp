// Internal declarations shared by the C-ABI host code and the HIP kernels.
#pragma once

#include <vector>

#include <stdint.h>

#include <hip/hip_runtime.h>

#include "nvpro_pyramid_hip.h"

namespace nvpro_pyramid {

// One mip level in linear device memory.
struct LevelDesc
{
  uint64_t offset;  // bytes from ImageDesc::base
  uint32_t width, height;
  uint32_t pitch;   // bytes per row
  uint32_t pad;
};

// A bound image: the kernels' view of NvproPyramidImage.
struct ImageDesc
{
  uint8_t* base;
  uint32_t levelCount;
  uint32_t format;
  uint32_t arithmetic;
  uint32_t pad;
  LevelDesc lv[NVPRO_PYRAMID_MAX_LEVELS];
};

// The steps a single tail launch runs (see tail_pipeline.hip).
struct TailStep
{
  uint32_t pipeline, inputLevel, levelCount, pad;
};
struct TailPlan
{
  uint32_t count, image, pad[2];  // image: index into the batch's ImageDesc array
  TailStep steps[NVPRO_PYRAMID_MAX_LEVELS];
};

// A step whose first output level has at most kTailMaxTexels texels, and every
// step after it, runs in the one-workgroup tail kernel.
constexpr uint32_t kTailMaxTexels = 1024;

// Batched work (batch_pipeline.hip): the steps of many images as segments, cut
// into items of consecutive groups, ordered so that every item comes after the
// items it depends on.
enum SegKind : uint32_t
{
  kSegFast1     = 1,  // kinds 1..6: fast step filling that many levels
  kSegGeneral1  = 7,
  kSegGeneral2  = 8,
  kSegTail      = 9,  // inputLevel holds the index of the TailPlan
};
constexpr uint32_t kNoSlot = 0xffffffffu;
struct QueueSeg
{
  uint32_t image, kind, inputLevel, tilesX, numTiles;
  uint32_t depSlot, depCount;  // wait until counter depSlot reaches depCount (kNoSlot: no wait)
  uint32_t doneSlot;           // counter bumped once per finished item (kNoSlot: none)
};
struct QueueItem
{
  uint32_t seg, firstGroup, groupCount, pad;
};
// Counter block of a batch: head of the queue, (unused), sticky error flags,
// then one completion counter per slot, then the arrival counters
// (kArriveWords, see below) of the queue kernel's exit.
constexpr uint32_t kCounterHead = 0, kCounterError = 2, kCounterSlots = 4;
constexpr uint32_t kBatchErrorSpin = 1u;  // a dependency wait gave up
// Largest first output level a tail may have inside the batched kernel (its LDS
// ping-pong overlays the other bodies' scratch).
constexpr uint32_t kBatchTailTexels = 128;

// Arrival counters of a fused launch (arriveLast in kernel_common.hpp): a root
// and kArriveLeaves leaves, each on its own 256-byte line.  In a grid of more
// than kArriveLeaves workgroups, workgroup b counts at leaf b % kArriveLeaves
// and the last workgroup of a leaf counts at the root; a smaller grid counts at
// the root only.
// One counter taken by every workgroup of a large grid serialises its atomics
// at one memory channel (microseconds for a few thousand workgroups); the
// leaves spread them over many.
constexpr uint32_t kArriveLeaves = 64;
constexpr uint32_t kArriveStride = 64;  // words between counters
constexpr uint32_t kArriveWords  = (1u + kArriveLeaves) * kArriveStride;

// Offset of a batch's exit arrival counters in its counter block, and the
// block's size in words.
__host__ __device__ inline uint32_t batchExitCounters(uint32_t numSlots)
{
  return (kCounterSlots + numSlots + kArriveStride - 1u) / kArriveStride * kArriveStride;
}
__host__ __device__ inline uint32_t batchCounterWords(uint32_t numSlots)
{
  return batchExitCounters(numSlots) + kArriveWords;
}

// Arrival counter blocks of fused launches: kFusedSlots blocks of kArriveWords
// in a zero-initialised device array (fast_pipeline.hip), one per stream that
// fuses, assigned by the dispatcher.  The block of `slot` on the current
// device, or null (the symbol's address is unavailable).
constexpr uint32_t kFusedSlots = 64;
uint32_t* fusedCounterBlock(uint32_t slot);

// What a fused launch runs in its last workgroup after its own step: a fast
// step of one group (kNextFast: inLevel, count, tilesX, numTiles) or the tail
// plan passed beside it (kNextTail).  arrive: kArriveWords zeroed counters that
// no other launch uses while this one runs.
constexpr uint32_t kNextFast = 1, kNextTail = 2;
struct NextStep
{
  uint32_t* arrive;
  uint32_t kind;
  uint32_t inLevel, count, tilesX, numTiles;
};
// Largest first output level of a tail that a fused fast launch runs (its
// ping-pong overlays the fast step's level-2 scratch).
constexpr uint32_t kFastTailTexels = 128;

// Per-pipeline launchers.  The caller has validated every step.  With `next`
// a fast launch also runs what it describes in its last workgroup (`tail`: the
// plan of a kNextTail).
hipError_t launchFast(hipStream_t stream, const ImageDesc& img, uint32_t inputLevel, uint32_t levelCount,
                      uint32_t wavesPerCU, const NextStep* next = nullptr, const TailPlan* tail = nullptr);
hipError_t launchGeneral(hipStream_t stream, const ImageDesc& img, uint32_t inputLevel, uint32_t levelCount);
hipError_t launchTail(hipStream_t stream, const ImageDesc& img, const TailPlan& plan);

// Batched kernel over items [itemBegin, itemEnd) of a work list in device
// memory.  queue: one persistent launch that dequeues items and waits on their
// dependencies (spinLimit bounds each wait); otherwise a launch of independent
// items (one round).  `blocks` workgroups.
struct BatchArgs
{
  const ImageDesc* imgs;
  const QueueSeg* segs;
  const QueueItem* items;
  const TailPlan* tails;
  uint32_t* counters;
  uint32_t* errorHost;  // queue mode: pinned host word the last workgroup copies the error flags to (or null)
  uint32_t itemBegin, itemEnd, numSlots, spinLimit;
};
hipError_t launchBatch(hipStream_t stream, uint32_t format, uint32_t arithmetic, bool queue, const BatchArgs& args,
                       uint32_t blocks);
// Resident workgroups per compute unit of the batched kernel.
uint32_t batchMaxWorkgroupsPerCU(uint32_t format, uint32_t arithmetic, bool queue);
// Copies copyBytes from pinned host memory `src` to device memory `dst` and
// zero-fills dst up to totalBytes, in one kernel on `stream` (both sizes
// multiples of 16, copyBytes <= totalBytes).  Used for work lists instead of
// hipMemcpyAsync, whose host-to-device copies of a few tens of KB were
// measured to block the calling thread for about 10 ms now and then while
// other streams had work in flight.
hipError_t launchStageCopy(hipStream_t stream, void* dst, const void* src, size_t copyBytes, size_t totalBytes);

// Seeded synthetic level 0 (synth.hip): one image per job; a batched launch
// writes chunks of many jobs.  A chunk is a strip of the columns [x0, x0 + 256)
// (or to the edge) by `rows` rows from y0.
struct SynthJob
{
  uint8_t* base;  // level 0
  uint64_t seed;
  uint32_t width, height, pitch, format;
};
struct SynthChunk
{
  uint32_t job, x0, y0, rows;
};
constexpr uint64_t kSynthChunk = 4096;  // texels per chunk of one image: enough workgroups for 1 MP
hipError_t launchSynthBatch(hipStream_t stream, const SynthJob* jobs, const SynthChunk* chunks, uint32_t numChunks);
// Appends the chunks of job `job` (about texelsPerChunk texels each); returns how many.
uint32_t appendSynthChunks(std::vector<SynthChunk>& chunks, uint32_t job, uint32_t width, uint32_t height,
                           uint64_t texelsPerChunk);

// Launch a validated plan: per-step kernels for the large steps, then one tail
// launch for the small remainder.  With `arrive` (kArriveWords zeroed counters
// of the stream), a fast step whose successor is a fast step of one group, or
// a tail of at most kFastTailTexels first outputs, runs that successor in its
// own launch.  *launches: the kernel launches made.
hipError_t executePlan(hipStream_t stream, const ImageDesc& img, const NvproPyramidStep* steps, uint32_t count,
                       uint32_t* arrive, uint32_t* launches);

// Number of compute units of the current device (cached per device).
uint32_t deviceComputeUnits();

// Host helpers of the dispatcher (nvpro_pyramid_dispatch.cpp).
// A plan has at most one step per level.
constexpr uint32_t kMaxPlanSteps = NVPRO_PYRAMID_MAX_LEVELS;
// Records a message for nvproPyramidGetLastError and returns code.
int hostFail(int code, const char* fmt, ...);
void hostClearError();
// Validates pipelines and image, binds the image and runs the dispatcher loop:
// *desc and steps[0, *stepCount) (capacity kMaxPlanSteps).  0 or an error.
int prepareImage(const NvproPyramidPipelines* pipelines, const NvproPyramidImage* image, uint32_t mipLevels,
                 nvpro_pyramid_dispatcher_t generalDispatcher, nvpro_pyramid_dispatcher_t fastDispatcher,
                 ImageDesc* desc, NvproPyramidStep* steps, uint32_t* stepCount, const char* who);

}  // namespace nvpro_pyramid
