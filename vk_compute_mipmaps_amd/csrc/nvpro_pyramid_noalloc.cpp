// nvproPyramidBatchDispatchNoAlloc: a batch in rounds mode whose work list
// lives in a static per-device array instead of memory the library allocates.
//
// `scratch` is a file-scope __device__ array; every GPU the library runs on
// holds its own copy at its own address.  In host code the array's name is only
// a host-side stand-in, so the device address of the current device's copy is
// looked up once per device with hipGetSymbolAddress and kept; kernels and
// copies only ever see that address.  Rounds-mode kernels read the work list
// and nothing else of it (they use no counters), so nothing in it has to start
// at zero.
#include <string.h>

#include <map>
#include <mutex>
#include <vector>

#include "batch_internal.hpp"

using namespace nvpro_pyramid;

static __device__ __align__(256) uint8_t scratch[NVPRO_PYRAMID_NO_ALLOC_SCRATCH_BYTES];

namespace {

// The device address of the current device's `scratch`, looked up on the first
// call on that device (with that device current).
hipError_t scratchAddress(void** out, const char* who)
{
  static std::mutex lock;
  static std::map<int, void*> addr;  // by device
  int device = 0;
  if(hipError_t e = hipGetDevice(&device))
    return (hipError_t)hostFail(e, "%s: hipGetDevice: %s", who, hipGetErrorString(e));
  std::lock_guard<std::mutex> guard(lock);
  auto it = addr.find(device);
  if(it == addr.end())
  {
    void* p = nullptr;
    if(hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)))
      return (hipError_t)hostFail(e, "%s: scratch address on device %d: %s", who, device, hipGetErrorString(e));
    it = addr.emplace(device, p).first;
  }
  *out = it->second;
  return hipSuccess;
}

}  // namespace

extern "C" int nvproPyramidBatchDispatchNoAlloc(hipStream_t stream, const NvproPyramidPipelines* pipelines,
                                                const NvproPyramidImage* images, const uint32_t* mipLevels,
                                                uint32_t imageCount, const NvproPyramidBatchConfig* config)
{
  const char* who = "nvproPyramidBatchDispatchNoAlloc";
  NvproPyramidBatchConfig cfg;
  defaultBatchConfig(&cfg);
  cfg.mode = NVPRO_PYRAMID_BATCH_MODE_ROUNDS;
  if(config)
    cfg = *config;
  if(int e = checkBatchConfig(cfg, who))
    return e;
  if(cfg.mode != NVPRO_PYRAMID_BATCH_MODE_ROUNDS)
    return hostFail(hipErrorInvalidValue, "%s: mode must be NVPRO_PYRAMID_BATCH_MODE_ROUNDS", who);
  // Host-side planning state, reused from call to call on this thread.
  thread_local BatchPlan plan;
  thread_local std::vector<uint8_t> host;
  thread_local BatchRunner runner;
  if(int e = buildBatchPlan(pipelines, images, mipLevels, imageCount, cfg, plan, who))
    return e;
  const size_t bytes = BatchRunner::roundsWorkListBytes(plan);
  if(bytes > NVPRO_PYRAMID_NO_ALLOC_SCRATCH_BYTES)
    return hostFail(hipErrorInvalidValue, "%s: work list of %zu bytes (%zu items) exceeds the %u-byte scratch", who,
                    bytes, plan.items.size(), (unsigned)NVPRO_PYRAMID_NO_ALLOC_SCRATCH_BYTES);
  if(plan.items.empty())
  {
    hostClearError();
    return 0;
  }
  void* dev = nullptr;
  if(hipError_t e = scratchAddress(&dev, who))
    return e;
  hipError_t e = runner.stageRoundsInto(plan, *pipelines, cfg, stream, dev, host);
  if(e == hipSuccess)
    e = runner.launch(stream);
  if(e != hipSuccess)
    return hostFail(e, "%s: %s", who, hipGetErrorString(e));
  hostClearError();
  return 0;
}
