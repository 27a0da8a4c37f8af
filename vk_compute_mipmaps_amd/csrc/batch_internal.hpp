// Host-side internals of batched chains, shared by the batch C-ABI and the
// serving loop: the work-list planner and a reusable device work list.
#pragma once

#include <stdint.h>

#include <vector>

#include "pyramid_internal.hpp"

namespace nvpro_pyramid {

// The work list of a batch (see batch_pipeline.hip).
struct BatchPlan
{
  std::vector<ImageDesc> imgs;
  std::vector<QueueSeg> segs;
  std::vector<QueueItem> items;
  std::vector<TailPlan> tails;
  std::vector<uint32_t> roundEnd;  // items [roundEnd[r-1], roundEnd[r]) form round r
  uint32_t numSlots         = 0;
  uint64_t algorithmicBytes = 0;
};

void defaultBatchConfig(NvproPyramidBatchConfig* config);
int checkBatchConfig(const NvproPyramidBatchConfig& config, const char* who);
// Plans every image (default dispatchers) and cuts the steps into items.
int buildBatchPlan(const NvproPyramidPipelines* pipelines, const NvproPyramidImage* images,
                   const uint32_t* mipLevels, uint32_t count, const NvproPyramidBatchConfig& config,
                   BatchPlan& plan, const char* who);

// A device copy of one work list, reused from batch to batch: a grow-only
// device buffer and pinned staging.  stage() writes the staging and enqueues
// its copy on `stream`; it must not be called again before the copy and every
// launch of the previous work list have finished.
class BatchRunner
{
public:
  BatchRunner() = default;
  BatchRunner(const BatchRunner&) = delete;
  BatchRunner& operator=(const BatchRunner&) = delete;
  ~BatchRunner();
  // `extra` (extraBytes, may be 0) is staged in the same copy, at extraDevice().
  hipError_t stage(const BatchPlan& plan, const NvproPyramidPipelines& pipes, const NvproPyramidBatchConfig& cfg,
                   hipStream_t stream, const void* extra = nullptr, size_t extraBytes = 0);
  void* extraDevice() const { return extra_; }
  // Queue mode: the launch's last workgroup stores the error flags at `host`
  // (pinned, or null for none).
  void setErrorHost(uint32_t* host) { args_.errorHost = host; }
  hipError_t launch(hipStream_t stream) const;
  uint32_t* errorWord() const;
  // Grows the buffers to at least `bytes` now, so that work lists up to that
  // size never allocate (hipFree of an outgrown buffer waits for the device).
  hipError_t reserve(size_t bytes);
  // Rounds mode without any allocation: the work list (which has no counters
  // in rounds mode) goes to caller-owned device memory `dev` of at least
  // roundsWorkListBytes(plan) bytes, through an ordinary host buffer `host`
  // (grown as needed) and one host-to-device copy on `stream`.
  static size_t roundsWorkListBytes(const BatchPlan& plan);
  hipError_t stageRoundsInto(const BatchPlan& plan, const NvproPyramidPipelines& pipes,
                             const NvproPyramidBatchConfig& cfg, hipStream_t stream, void* dev,
                             std::vector<uint8_t>& host);

  uint32_t launches = 0, blocks = 0;
  size_t devBytes   = 0;

private:
  void bind(const BatchPlan& plan, const NvproPyramidPipelines& pipes, const NvproPyramidBatchConfig& cfg,
            uint8_t* base, size_t oImg, size_t oSeg, size_t oItm, size_t oTl, size_t oCnt, size_t size);
  void* dev_ = nullptr;
  void* pinned_ = nullptr;
  void* extra_  = nullptr;
  size_t devCap_ = 0, pinnedCap_ = 0;
  BatchArgs args_{};
  uint32_t format_ = 0, arithmetic_ = 0, items_ = 0;
  bool queue_ = true;
  std::vector<uint32_t> roundEnd_;
};

}  // namespace nvpro_pyramid
