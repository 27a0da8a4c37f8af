// Batched kernel: the steps of many images' plans in one launch (queue mode) or
// in one launch per round of independent steps (rounds mode).
//
// The host cuts every step of every image into items (runs of consecutive
// groups of one step) and orders them so that an item comes after every item
// it depends on: step s of an image needs all of step s-1.  A workgroup runs any
// item kind with the group bodies of groups.hpp: fast steps of 1..6 levels,
// general steps of 1 or 2 levels, and whole tails.  The bodies' LDS scratch is
// overlaid; the decode tables (DualTables) serve the fast bodies' premultiplied
// loads and the general ones' plain loads with exactly the values their
// per-step kernels see.
//
// Queue mode: workgroups are persistent and take the next item from an atomic
// head counter.  Before an item of step s a workgroup waits for the completion
// counter of step s-1 to reach its item count (then drops stale cache lines);
// an item stores its step's last level with write-through stores, waits for
// them to complete and bumps the step's counter.  Items are taken in list
// order, so the unfinished item with the smallest index never waits on
// anything unfinished: the launch completes whatever the number of resident
// workgroups.  Each wait is bounded all the same (spinLimit); a wait that gives
// up sets an error flag the host can read, and the workgroup carries on, so a
// launch always ends.  The last workgroup to exit (arriveLast) clears the head
// and completion counters for the next launch.
//
// Rounds mode: the items of one round are independent; workgroups take items
// blockIdx.x, blockIdx.x + gridDim.x, ... and stream order separates rounds.
#include "groups.hpp"

namespace nvpro_pyramid {

template <class F>
struct BatchScratch
{
  using V                     = typename F::Val;
  static constexpr size_t a   = sizeof(V) * kGroupBS;                            // fast: level2
  static constexpr size_t b   = sizeof(V) * gen2MidTexels<F>();                 // general: mid
  static constexpr size_t c   = sizeof(V) * 2 * kBatchTailTexels;               // tail: ping-pong
  static constexpr size_t ab  = a > b ? a : b;
  static constexpr size_t kBytes = ab > c ? ab : c;
};

template <class F>
struct BatchShared
{
  typename F::DualTables tab;
  alignas(16) unsigned char scratch[BatchScratch<F>::kBytes];
};

// The work list is read-only during a launch and never aliases texel memory;
// the restrict-qualified pointers let its fields live in scalar registers.
#define NVPRO_BATCH_PARAMS                                                                                          \
  const ImageDesc *__restrict__ imgs, const QueueSeg *__restrict__ segs, const QueueItem *__restrict__ items,       \
      const TailPlan *__restrict__ tails, uint32_t *__restrict__ counters, uint32_t *__restrict__ errorHost,        \
      uint32_t itemBegin, uint32_t itemEnd, uint32_t numSlots, uint32_t spinLimit

// CO: the step's last level is stored coherently, for the queue mode's
// successors in the same launch.
template <class F, bool CO>
__device__ __forceinline__ void runItem(BatchShared<F>& sh, const ImageDesc* __restrict__ imgs,
                                        const TailPlan* __restrict__ tails, const QueueSeg& seg,
                                        const QueueItem& item, uint32_t& iter)
{
  using V              = typename F::Val;
  V* scratch           = reinterpret_cast<V*>(sh.scratch);
  const ImageDesc& img = imgs[seg.image];
  const uint32_t end   = item.firstGroup + item.groupCount;
  for(uint32_t g = item.firstGroup; g < end; ++g, ++iter)
  {
    switch(seg.kind)
    {
      case 1: fastGroup<F, 1, CO, false, false, false>(sh.tab, scratch, img, seg.inputLevel, seg.tilesX, seg.numTiles, g, iter); break;
      case 2: fastGroup<F, 2, CO, false, false, false>(sh.tab, scratch, img, seg.inputLevel, seg.tilesX, seg.numTiles, g, iter); break;
      case 3: fastGroup<F, 3, CO, false, false, false>(sh.tab, scratch, img, seg.inputLevel, seg.tilesX, seg.numTiles, g, iter); break;
      case 4: fastGroup<F, 4, CO, false, false, false>(sh.tab, scratch, img, seg.inputLevel, seg.tilesX, seg.numTiles, g, iter); break;
      case 5: fastGroup<F, 5, CO, false, false, false>(sh.tab, scratch, img, seg.inputLevel, seg.tilesX, seg.numTiles, g, iter); break;
      case 6: fastGroup<F, 6, CO, false, false, false>(sh.tab, scratch, img, seg.inputLevel, seg.tilesX, seg.numTiles, g, iter); break;
      case kSegGeneral1: generalOneGroup<F, CO>(sh.tab, img, seg.inputLevel, g); break;
      case kSegGeneral2: generalTwoGroupAny<F, CO>(sh.tab, scratch, img, seg.inputLevel, seg.tilesX, g); break;
      case kSegTail:
        tailRun<F, kGroupBS>(sh.tab, scratch, scratch + kBatchTailTexels, img, tails[seg.inputLevel]);
        break;
      default: break;
    }
  }
}

template <class F>
__global__ void __launch_bounds__(kGroupBS) __attribute__((amdgpu_waves_per_eu(F::kBytes > 8u ? 4 : 5, 8)))
batchRoundKernel(NVPRO_BATCH_PARAMS)
{
  __shared__ BatchShared<F> sh;
  F::template fillDual<kGroupBS>(sh.tab);
  __syncthreads();
  uint32_t iter = 0;
  for(uint32_t it = itemBegin + blockIdx.x; it < itemEnd; it += gridDim.x)
  {
    const QueueItem item = items[it];
    const QueueSeg seg   = segs[item.seg];
    runItem<F, false>(sh, imgs, tails, seg, item, iter);
    __syncthreads();  // the next item may overlay this one's scratch
  }
}

template <class F>
__global__ void __launch_bounds__(kGroupBS) __attribute__((amdgpu_waves_per_eu(F::kBytes > 8u ? 4 : 5, 8)))
batchQueueKernel(NVPRO_BATCH_PARAMS)
{
  __shared__ BatchShared<F> sh;
  __shared__ uint32_t sNext, sFlag;
  uint32_t* head  = counters + kCounterHead;
  uint32_t* error = counters + kCounterError;
  uint32_t* slots = counters + kCounterSlots;
  F::template fillDual<kGroupBS>(sh.tab);
  if(threadIdx.x == 0)
    sNext = itemBegin + __hip_atomic_fetch_add(head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  uint32_t iter = 0;
  uint32_t it   = __builtin_amdgcn_readfirstlane(sNext);
  while(it < itemEnd)
  {
    const QueueItem item = items[it];
    const QueueSeg seg   = segs[item.seg];
    uint32_t next        = 0;
    if(threadIdx.x == 0)
    {
      // Take the next item now so its latency overlaps this one.  Items are taken
      // in order and this one is finished before the next starts, so holding it
      // early cannot block anything.
      next = itemBegin + __hip_atomic_fetch_add(head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if(seg.depSlot != kNoSlot)
        waitCounter(slots + seg.depSlot, seg.depCount, spinLimit, error);
    }
    if(seg.depSlot != kNoSlot)
    {
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // drop stale lines before reading the previous step
    }
    runItem<F, true>(sh, imgs, tails, seg, item, iter);
    // Publish: the step's last level went out with write-through stores; once
    // every wave's stores have completed the item is counted.  (An agent-scope
    // release fence here would write back the whole L2 of the XCD per wave and
    // item, which costs milliseconds over a batch.)
    if(seg.doneSlot != kNoSlot)
      __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if(threadIdx.x == 0)
    {
      if(seg.doneSlot != kNoSlot)
        __hip_atomic_fetch_add(slots + seg.doneSlot, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      sNext = next;
    }
    __syncthreads();
    it = __builtin_amdgcn_readfirstlane(sNext);
  }
  // Exit: the last workgroup out resets the counters for the next launch.  Every
  // other workgroup has made its last head and slot update by then.
  if(arriveLast(counters + batchExitCounters(numSlots), gridDim.x, &sFlag))
  {
    for(uint32_t i = threadIdx.x; i < numSlots; i += kGroupBS)
      __hip_atomic_store(slots + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if(threadIdx.x == 0)
    {
      __hip_atomic_store(head, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // Every wait has ended by now; the flags go straight to the host (a
      // system-scope store), saving a copy after the launch.
      if(errorHost)
        __hip_atomic_store(errorHost, __hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

template <class F>
static hipError_t launchBatchTyped(hipStream_t stream, bool queue, const BatchArgs& a, uint32_t blocks)
{
  if(queue)
    hipLaunchKernelGGL((batchQueueKernel<F>), dim3(blocks), dim3(kGroupBS), 0, stream, a.imgs, a.segs, a.items,
                       a.tails, a.counters, a.errorHost, a.itemBegin, a.itemEnd, a.numSlots, a.spinLimit);
  else
    hipLaunchKernelGGL((batchRoundKernel<F>), dim3(blocks), dim3(kGroupBS), 0, stream, a.imgs, a.segs, a.items,
                       a.tails, a.counters, a.errorHost, a.itemBegin, a.itemEnd, a.numSlots, a.spinLimit);
  return hipGetLastError();
}

hipError_t launchBatch(hipStream_t stream, uint32_t format, uint32_t arithmetic, bool queue, const BatchArgs& a,
                       uint32_t blocks)
{
  if(blocks == 0)
    return hipErrorInvalidValue;
  switch(format)
  {
    case NVPRO_PYRAMID_FORMAT_SRGBA8:
      return launchBatchTyped<Srgba8>(stream, queue, a, blocks);
    case NVPRO_PYRAMID_FORMAT_RGBA16F:
      return arithmetic == NVPRO_PYRAMID_ARITH_PACKED_FP16 ? launchBatchTyped<Rgba16fPacked>(stream, queue, a, blocks)
                                                           : launchBatchTyped<Rgba16f>(stream, queue, a, blocks);
    case NVPRO_PYRAMID_FORMAT_RGBA32F:
      return launchBatchTyped<Rgba32f>(stream, queue, a, blocks);
    default:
      return hipErrorInvalidValue;
  }
}

// The host buffer is pinned memory the host wrote before the launch.  It is
// read with system-scope loads, which do not hit lines a device cache may
// still hold from an earlier batch's copy of the same buffer.
__global__ void __launch_bounds__(256) stageCopyKernel(uint2* __restrict__ dst, const uint64_t* __restrict__ src,
                                                       uint32_t copyWords, uint32_t totalWords)
{
  for(uint32_t i = blockIdx.x * 256u + threadIdx.x; i < totalWords; i += gridDim.x * 256u)
  {
    const uint64_t v = i < copyWords ? __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) : 0u;
    dst[i]           = make_uint2((uint32_t)v, (uint32_t)(v >> 32));
  }
}

hipError_t launchStageCopy(hipStream_t stream, void* dst, const void* src, size_t copyBytes, size_t totalBytes)
{
  if(copyBytes > totalBytes || (copyBytes | totalBytes) % 16u != 0 || totalBytes / 8u > 0xffffffffull)
    return hipErrorInvalidValue;
  const uint32_t total = (uint32_t)(totalBytes / 8u);
  if(total == 0)
    return hipSuccess;
  const uint32_t blocks = std::min<uint32_t>((total + 255u) / 256u, 64u);
  hipLaunchKernelGGL(stageCopyKernel, dim3(blocks), dim3(256), 0, stream, (uint2*)dst, (const uint64_t*)src,
                     (uint32_t)(copyBytes / 8u), total);
  return hipGetLastError();
}

template <class F>
static uint32_t maxWorkgroupsTyped(bool queue)
{
  int n = 0;
  hipError_t e = queue ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batchQueueKernel<F>, kGroupBS, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batchRoundKernel<F>, kGroupBS, 0);
  return (e == hipSuccess && n > 0) ? (uint32_t)n : 1u;
}

uint32_t batchMaxWorkgroupsPerCU(uint32_t format, uint32_t arithmetic, bool queue)
{
  switch(format)
  {
    case NVPRO_PYRAMID_FORMAT_SRGBA8:
      return maxWorkgroupsTyped<Srgba8>(queue);
    case NVPRO_PYRAMID_FORMAT_RGBA16F:
      return arithmetic == NVPRO_PYRAMID_ARITH_PACKED_FP16 ? maxWorkgroupsTyped<Rgba16fPacked>(queue)
                                                           : maxWorkgroupsTyped<Rgba16f>(queue);
    case NVPRO_PYRAMID_FORMAT_RGBA32F:
      return maxWorkgroupsTyped<Rgba32f>(queue);
    default:
      return 1u;
  }
}

}  // namespace nvpro_pyramid
