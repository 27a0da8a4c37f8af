// C-ABI host side: nvproCmdPyramidDispatch's level-splitting loop and default
// dispatchers (nvpro_pyramid/nvpro_pyramid_dispatch.hpp:109-304 of the
// reference), re-expressed over HIP streams.
//
// A dispatch call runs the dispatcher loop against an NvproPyramidCmd that only
// records steps; when the loop has finished without error the recorded steps are
// launched in order on the caller's stream.  Stream order stands in for the
// vkCmdPipelineBarrier the reference records between dispatches (:180-186).
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "pyramid_internal.hpp"

using namespace nvpro_pyramid;

namespace {

constexpr uint32_t kMaxSteps = 64;  // far above the <= 15 a 16-level chain can need

thread_local char gLastError[256] = "";
thread_local uint32_t gLastLaunches = 0;

int vfail(int code, const char* fmt, va_list ap)
{
  vsnprintf(gLastError, sizeof gLastError, fmt, ap);
  return code;
}

int fail(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vfail(code, fmt, ap);
  va_end(ap);
  return code;
}

// Arrival counters of fused launches (executePlan): a block per stream out of
// the kFusedSlots blocks of a static device array (fusedCounterBlock), so the
// dispatch path allocates nothing.  A stream runs its launches one after
// another, so a block is never in use by two launches at once, and every fused
// launch leaves it zeroed.  Streams are told apart by hipStreamGetId, which a
// later stream does not reuse; a slot stays with its stream.  There is no
// block (no fusion) while the stream is being captured: a graph may be
// launched on any stream, beside eager launches on the capturing one, and
// would share their block; nor for a stream of another device than the
// current one, nor while every slot is held, nor when NVPRO_PYRAMID_FUSE=0 is
// set in the environment.  nvproPyramidStreamRelease hands a stream's block
// back (after the stream's work) for later streams.
std::mutex gSlotLock;
std::map<std::pair<int, unsigned long long>, uint32_t> gSlots;  // by device, stream id
std::map<int, uint32_t> gSlotsUsed;                              // slots ever taken, by device
std::map<int, std::vector<uint32_t>> gSlotsFree;                 // released slots, by device

uint32_t* fusedCounters(hipStream_t stream)
{
  std::mutex& lock = gSlotLock;
  auto& slots      = gSlots;
  auto& used       = gSlotsUsed;
  const char* env = getenv("NVPRO_PYRAMID_FUSE");
  if(env && env[0] == '0')
    return nullptr;
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  unsigned long long id          = 0;
  int device                     = 0;
  if(hipStreamIsCapturing(stream, &capture) != hipSuccess || capture != hipStreamCaptureStatusNone ||
     hipStreamGetId(stream, &id) != hipSuccess || hipGetDevice(&device) != hipSuccess)
  {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> guard(lock);
  const auto key = std::make_pair(device, id);
  const auto it  = slots.find(key);
  if(it != slots.end())
    return fusedCounterBlock(it->second);
  int streamDevice = 0;
  std::vector<uint32_t>& freed = gSlotsFree[device];
  if((freed.empty() && used[device] >= kFusedSlots) || hipStreamGetDevice(stream, &streamDevice) != hipSuccess ||
     device != streamDevice)
  {
    (void)hipGetLastError();
    return nullptr;
  }
  const uint32_t slot = freed.empty() ? used[device] : freed.back();
  uint32_t* block     = fusedCounterBlock(slot);
  if(!block)
    return nullptr;
  if(freed.empty())
    used[device]++;
  else
    freed.pop_back();
  slots[key] = slot;
  return block;
}

}  // namespace

int nvpro_pyramid::hostFail(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vfail(code, fmt, ap);
  va_end(ap);
  return code;
}

void nvpro_pyramid::hostClearError() { gLastError[0] = 0; }

struct NvproPyramidCmd
{
  NvproPyramidPipelines pipelines;
  uint32_t width, height, mipLevels;
  NvproPyramidStep steps[kMaxSteps];
  uint32_t stepCount;
  int error;
};

static void levelDims(uint32_t w, uint32_t h, uint32_t level, uint32_t* lw, uint32_t* lh)
{
  for(uint32_t i = 0; i < level; ++i)
  {
    w = (w >> 1) | (w == 1u);
    h = (h >> 1) | (h == 1u);
  }
  *lw = w;
  *lh = h;
}

extern "C" {

uint32_t nvproPyramidFullLevelCount(uint32_t width, uint32_t height)
{
  uint32_t n = 0;
  while(width != 0 || height != 0)
  {
    width >>= 1;
    height >>= 1;
    ++n;
  }
  return n;
}

uint32_t nvproPyramidBytesPerTexel(uint32_t format)
{
  switch(format)
  {
    case NVPRO_PYRAMID_FORMAT_SRGBA8:
      return 4u;
    case NVPRO_PYRAMID_FORMAT_RGBA16F:
      return 8u;
    case NVPRO_PYRAMID_FORMAT_RGBA32F:
      return 16u;
    default:
      return 0u;
  }
}

const char* nvproPyramidGetLastError(void) { return gLastError; }

int nvproPyramidPackedLayout(uint32_t width, uint32_t height, uint32_t levelCount, uint32_t bytesPerTexel,
                             uint64_t* offsets, uint64_t* totalBytes)
{
  if(width == 0 || height == 0)
    return fail(hipErrorInvalidValue, "nvproPyramidPackedLayout: zero image size %ux%u", width, height);
  if(bytesPerTexel == 0)
    return fail(hipErrorInvalidValue, "nvproPyramidPackedLayout: bytesPerTexel is 0");
  const uint32_t full = nvproPyramidFullLevelCount(width, height);
  if(levelCount == 0)
    levelCount = full;
  if(levelCount > full)
    return fail(hipErrorInvalidValue, "nvproPyramidPackedLayout: %u levels exceed the %u of a %ux%u image",
                levelCount, full, width, height);
  uint64_t off = 0;
  for(uint32_t l = 0; l < levelCount; ++l)
  {
    uint32_t lw, lh;
    levelDims(width, height, l, &lw, &lh);
    if(offsets)
      offsets[l] = off;
    off += (uint64_t)lw * lh * bytesPerTexel;
  }
  if(totalBytes)
    *totalBytes = off;
  return 0;
}

int nvproPyramidCmdDispatch(NvproPyramidCmd* cmd, uint32_t pipeline, uint32_t pushConstant, uint32_t groupCountX)
{
  if(!cmd)
    return fail(hipErrorInvalidValue, "nvproPyramidCmdDispatch: null command");
  if(cmd->error)
    return cmd->error;
  const uint32_t level = pushConstant >> NVPRO_PYRAMID_INPUT_LEVEL_SHIFT;
  const uint32_t count = pushConstant & NVPRO_PYRAMID_LEVEL_COUNT_MASK;
  int err              = 0;
  if(pipeline != NVPRO_PYRAMID_PIPELINE_GENERAL && pipeline != NVPRO_PYRAMID_PIPELINE_FAST)
    err = fail(hipErrorInvalidValue, "nvproPyramidCmdDispatch: unknown pipeline %u", pipeline);
  else if(pipeline == NVPRO_PYRAMID_PIPELINE_FAST && !cmd->pipelines.fastPipeline)
    err = fail(hipErrorInvalidValue, "nvproPyramidCmdDispatch: fast pipeline is not available");
  else if(count == 0 || level + count >= cmd->mipLevels)
    err = fail(hipErrorInvalidValue, "nvproPyramidCmdDispatch: levels %u+1..%u outside a %u-level image", level,
               level + count, cmd->mipLevels);
  else if(pipeline == NVPRO_PYRAMID_PIPELINE_GENERAL && count > 2u)
    err = fail(hipErrorInvalidValue, "nvproPyramidCmdDispatch: general pipeline fills 1 or 2 levels, not %u", count);
  else if(pipeline == NVPRO_PYRAMID_PIPELINE_FAST)
  {
    uint32_t w, h;
    levelDims(cmd->width, cmd->height, level, &w, &h);
    const uint32_t m = (1u << (count > 6u ? 6u : count)) - 1u;
    if(count > 6u)
      err = fail(hipErrorInvalidValue, "nvproPyramidCmdDispatch: fast pipeline fills at most 6 levels, not %u", count);
    else if((w & m) != 0 || (h & m) != 0)
      err = fail(hipErrorInvalidValue,
                 "nvproPyramidCmdDispatch: fast pipeline needs level %u (%ux%u) divisible by %u", level, w, h, m + 1u);
  }
  if(!err && cmd->stepCount >= kMaxSteps)
    err = fail(hipErrorInvalidValue, "nvproPyramidCmdDispatch: more than %u dispatches", kMaxSteps);
  if(err)
  {
    cmd->error = err;
    return err;
  }
  NvproPyramidStep& s = cmd->steps[cmd->stepCount++];
  s.pipeline          = pipeline;
  s.inputLevel        = level;
  s.levelCount        = count;
  s.pushConstant      = pushConstant;
  s.groupCountX       = groupCountX;
  levelDims(cmd->width, cmd->height, level, &s.inputWidth, &s.inputHeight);
  s.reserved = 0;
  return 0;
}

uint32_t nvproPyramidFastDispatcherEx(NvproPyramidCmd* cmd, const NvproPyramidState* state, uint32_t divisibility,
                                      uint32_t maxLevels)
{
  // nvpro_pyramid_dispatch.hpp:204-206, :218: even divisibility, at most 6 levels.
  if(!cmd || !state || divisibility == 0 || divisibility % 2u != 0 || maxLevels == 0 || maxLevels > 6u)
    return 0;
  if(state->currentX % divisibility != 0 || state->currentY % divisibility != 0)
    return 0;
  uint32_t x = state->currentX, y = state->currentY, levels = 0;
  while(x % 2u == 0 && y % 2u == 0 && levels < state->remainingLevels && levels < maxLevels)
  {
    x /= 2u;
    y /= 2u;
    levels++;
  }
  const uint32_t pc      = state->currentLevel << NVPRO_PYRAMID_INPUT_LEVEL_SHIFT | levels;
  const uint32_t shift   = levels > 5u ? 12u : 10u;
  const uint32_t mask    = levels > 5u ? 4095u : 1023u;
  const uint64_t samples = (uint64_t)state->currentX * state->currentY;
  if(nvproPyramidCmdDispatch(cmd, NVPRO_PYRAMID_PIPELINE_FAST, pc, (uint32_t)((samples + mask) >> shift)))
    return 0;
  return levels;
}

uint32_t nvproPyramidDefaultFastDispatcher(NvproPyramidCmd* cmd, const NvproPyramidState* state)
{
  return nvproPyramidFastDispatcherEx(cmd, state, 4u, 6u);
}

uint32_t nvproPyramidDefaultGeneralDispatcher(NvproPyramidCmd* cmd, const NvproPyramidState* state)
{
  // py2_4_8_8 parameters (nvpro_pyramid_dispatch.hpp:254-258).
  constexpr uint32_t kMaxLevels = 2, kWarps = 4, kTileWidth = 8, kTileHeight = 8;
  if(!cmd || !state)
    return 0;
  const uint32_t levels = state->remainingLevels >= kMaxLevels ? kMaxLevels : state->remainingLevels;
  const uint32_t pc     = state->currentLevel << NVPRO_PYRAMID_INPUT_LEVEL_SHIFT | levels;
  uint32_t dstWidth     = state->currentX >> levels;
  dstWidth              = dstWidth ? dstWidth : 1u;
  uint32_t dstHeight    = state->currentY >> levels;
  dstHeight             = dstHeight ? dstHeight : 1u;
  uint32_t groups;
  if(levels == 1u)
  {
    const uint64_t samples = (uint64_t)dstWidth * dstHeight;
    const uint32_t threads = kWarps * 32u;
    groups                 = (uint32_t)((samples + (threads - 1u)) / threads);
  }
  else
  {
    groups = ((dstWidth + (kTileWidth - 1u)) / kTileWidth) * ((dstHeight + (kTileHeight - 1u)) / kTileHeight);
  }
  if(nvproPyramidCmdDispatch(cmd, NVPRO_PYRAMID_PIPELINE_GENERAL, pc, groups))
    return 0;
  return levels;
}

}  // extern "C"

// The dispatcher loop of nvpro_pyramid_dispatch.hpp:133-187 against a recording
// command.  Returns 0 or an error; on error nothing must be launched.
static int runDispatchLoop(NvproPyramidCmd* cmd, nvpro_pyramid_dispatcher_t general, nvpro_pyramid_dispatcher_t fast)
{
  if(!general)
    general = nvproPyramidDefaultGeneralDispatcher;
  if(!fast)
    fast = nvproPyramidDefaultFastDispatcher;
  NvproPyramidState state{0u, cmd->mipLevels - 1u, cmd->width, cmd->height};
  while(state.remainingLevels != 0u)
  {
    uint32_t levelsDone = 0;
    if(cmd->pipelines.fastPipeline)
      levelsDone = fast(cmd, &state);
    if(cmd->error)
      return cmd->error;
    if(levelsDone == 0)
    {
      levelsDone = general(cmd, &state);
      if(cmd->error)
        return cmd->error;
      if(levelsDone == 0)
        return fail(hipErrorInvalidValue, "general dispatcher filled no level at level %u", state.currentLevel);
    }
    if(levelsDone > state.remainingLevels)
      return fail(hipErrorInvalidValue, "dispatcher claimed %u levels with %u remaining", levelsDone,
                  state.remainingLevels);
    state.currentLevel += levelsDone;
    state.remainingLevels -= levelsDone;
    state.currentX >>= levelsDone;
    state.currentX = state.currentX ? state.currentX : 1u;
    state.currentY >>= levelsDone;
    state.currentY = state.currentY ? state.currentY : 1u;
  }
  return 0;
}

static int checkPipelines(const NvproPyramidPipelines* p, const char* who)
{
  if(!p)
    return fail(hipErrorInvalidValue, "%s: null pipelines", who);
  if(nvproPyramidBytesPerTexel(p->format) == 0)
    return fail(hipErrorInvalidValue, "%s: unknown format %u", who, p->format);
  if(p->arithmetic > NVPRO_PYRAMID_ARITH_PACKED_FP16 ||
     (p->arithmetic != NVPRO_PYRAMID_ARITH_FP32 && p->format != NVPRO_PYRAMID_FORMAT_RGBA16F))
    return fail(hipErrorInvalidValue, "%s: arithmetic %u not available for format %u", who, p->arithmetic, p->format);
  if(p->reserved != 0)
    return fail(hipErrorInvalidValue, "%s: reserved field must be 0", who);
  return 0;
}

static int initCmd(NvproPyramidCmd* cmd, const NvproPyramidPipelines* p, uint32_t w, uint32_t h, uint32_t mipLevels,
                   uint32_t available, const char* who)
{
  memset(cmd, 0, sizeof *cmd);
  if(w == 0 || h == 0)
    return fail(hipErrorInvalidValue, "%s: zero image size %ux%u", who, w, h);
  const uint32_t full = nvproPyramidFullLevelCount(w, h);
  if(full > NVPRO_PYRAMID_MAX_LEVELS)
    return fail(hipErrorInvalidValue, "%s: %ux%u needs %u levels, more than %u", who, w, h, full,
                NVPRO_PYRAMID_MAX_LEVELS);
  if(mipLevels == 0)
    mipLevels = full;
  if(mipLevels > full)
    return fail(hipErrorInvalidValue, "%s: mipLevels %u exceeds the %u of a %ux%u image", who, mipLevels, full, w, h);
  if(mipLevels > available)
    return fail(hipErrorInvalidValue, "%s: mipLevels %u exceeds the image's %u levels", who, mipLevels, available);
  cmd->pipelines = *p;
  cmd->width     = w;
  cmd->height    = h;
  cmd->mipLevels = mipLevels;
  return 0;
}

extern "C" int nvproPyramidPlan(const NvproPyramidPipelines* pipelines, uint32_t width, uint32_t height,
                                uint32_t mipLevels, nvpro_pyramid_dispatcher_t generalDispatcher,
                                nvpro_pyramid_dispatcher_t fastDispatcher, NvproPyramidStep* steps, uint32_t capacity,
                                uint32_t* stepCount)
{
  if(int e = checkPipelines(pipelines, "nvproPyramidPlan"))
    return e;
  NvproPyramidCmd cmd;
  if(int e = initCmd(&cmd, pipelines, width, height, mipLevels, NVPRO_PYRAMID_MAX_LEVELS, "nvproPyramidPlan"))
    return e;
  if(int e = runDispatchLoop(&cmd, generalDispatcher, fastDispatcher))
    return e;
  for(uint32_t i = 0; i < cmd.stepCount && i < capacity && steps; ++i)
    steps[i] = cmd.steps[i];
  if(stepCount)
    *stepCount = cmd.stepCount;
  return 0;
}

static int bindImage(const NvproPyramidImage* image, uint32_t format, uint32_t arithmetic, ImageDesc* desc,
                     const char* who)
{
  if(!image)
    return fail(hipErrorInvalidValue, "%s: null image", who);
  if(!image->data)
    return fail(hipErrorInvalidValue, "%s: null image data", who);
  if(image->reserved != 0)
    return fail(hipErrorInvalidValue, "%s: reserved field must be 0", who);
  const uint32_t bpt = nvproPyramidBytesPerTexel(format);
  if(((uintptr_t)image->data & 3u) != 0)
    return fail(hipErrorInvalidValue, "%s: image data must be 4-byte aligned", who);
  if(image->width == 0 || image->height == 0)
    return fail(hipErrorInvalidValue, "%s: zero image size %ux%u", who, image->width, image->height);
  const uint32_t full = nvproPyramidFullLevelCount(image->width, image->height);
  const uint32_t n    = image->levelCount ? image->levelCount : full;
  if(n > full || n > NVPRO_PYRAMID_MAX_LEVELS)
    return fail(hipErrorInvalidValue, "%s: image levelCount %u invalid for %ux%u (max %u)", who, n, image->width,
                image->height, full < NVPRO_PYRAMID_MAX_LEVELS ? full : NVPRO_PYRAMID_MAX_LEVELS);
  memset(desc, 0, sizeof *desc);
  desc->base       = (uint8_t*)image->data;
  desc->levelCount = n;
  desc->format     = format;
  desc->arithmetic = arithmetic;
  uint64_t packed  = 0;
  for(uint32_t l = 0; l < n; ++l)
  {
    uint32_t lw, lh;
    levelDims(image->width, image->height, l, &lw, &lh);
    const uint64_t tight = (uint64_t)lw * bpt;
    const uint64_t pitch = image->rowPitches ? image->rowPitches[l] : tight;
    const uint64_t off   = image->levelOffsets ? image->levelOffsets[l] : packed;
    if(pitch < tight || pitch % bpt != 0 || pitch > 0xffffffffull)
      return fail(hipErrorInvalidValue, "%s: level %u row pitch %llu invalid (width %u, %u bytes/texel)", who, l,
                  (unsigned long long)pitch, lw, bpt);
    if(off % bpt != 0)
      return fail(hipErrorInvalidValue, "%s: level %u offset %llu not a multiple of %u", who, l,
                  (unsigned long long)off, bpt);
    desc->lv[l].offset = off;
    desc->lv[l].width  = lw;
    desc->lv[l].height = lh;
    desc->lv[l].pitch  = (uint32_t)pitch;
    packed += tight * lh;
  }
  return 0;
}

int nvpro_pyramid::prepareImage(const NvproPyramidPipelines* pipelines, const NvproPyramidImage* image,
                                uint32_t mipLevels, nvpro_pyramid_dispatcher_t generalDispatcher,
                                nvpro_pyramid_dispatcher_t fastDispatcher, ImageDesc* desc, NvproPyramidStep* steps,
                                uint32_t* stepCount, const char* who)
{
  *stepCount = 0;
  if(int e = checkPipelines(pipelines, who))
    return e;
  if(int e = bindImage(image, pipelines->format, pipelines->arithmetic, desc, who))
    return e;
  NvproPyramidCmd cmd;
  if(int e = initCmd(&cmd, pipelines, image->width, image->height, mipLevels, desc->levelCount, who))
    return e;
  if(cmd.mipLevels <= 1u)
    return 0;
  if(int e = runDispatchLoop(&cmd, generalDispatcher, fastDispatcher))
    return e;
  if(cmd.stepCount > kMaxPlanSteps)
    return fail(hipErrorInvalidValue, "%s: plan of %u steps", who, cmd.stepCount);
  for(uint32_t i = 0; i < cmd.stepCount; ++i)
    steps[i] = cmd.steps[i];
  *stepCount = cmd.stepCount;
  return 0;
}

extern "C" int nvproPyramidDispatchCustom(hipStream_t stream, const NvproPyramidPipelines* pipelines,
                                          const NvproPyramidImage* image, uint32_t mipLevels,
                                          nvpro_pyramid_dispatcher_t generalDispatcher,
                                          nvpro_pyramid_dispatcher_t fastDispatcher)
{
  const char* who = "nvproPyramidDispatch";
  ImageDesc desc;
  NvproPyramidStep steps[kMaxPlanSteps];
  uint32_t count = 0;
  if(int e = prepareImage(pipelines, image, mipLevels, generalDispatcher, fastDispatcher, &desc, steps, &count, who))
    return e;
  if(count != 0)
  {
    uint32_t launches = 0;
    hipError_t e      = executePlan(stream, desc, steps, count, count >= 2u ? fusedCounters(stream) : nullptr, &launches);
    if(e != hipSuccess)
      return fail(e, "%s: kernel launch failed: %s", who, hipGetErrorString(e));
    gLastLaunches = launches;
  }
  else
    gLastLaunches = 0;
  gLastError[0] = 0;
  return 0;
}

extern "C" int nvproPyramidDispatch(hipStream_t stream, const NvproPyramidPipelines* pipelines,
                                    const NvproPyramidImage* image, uint32_t mipLevels)
{
  return nvproPyramidDispatchCustom(stream, pipelines, image, mipLevels, nvproPyramidDefaultGeneralDispatcher,
                                    nvproPyramidDefaultFastDispatcher);
}

extern "C" uint32_t nvproPyramidLastLaunchCount(void) { return gLastLaunches; }

extern "C" int nvproPyramidStreamRelease(hipStream_t stream)
{
  const char* who = "nvproPyramidStreamRelease";
  unsigned long long id = 0;
  int device            = 0;
  if(hipError_t e = hipStreamGetId(stream, &id))
    return fail(e, "%s: hipStreamGetId: %s", who, hipGetErrorString(e));
  if(hipError_t e = hipStreamGetDevice(stream, &device))
    return fail(e, "%s: hipStreamGetDevice: %s", who, hipGetErrorString(e));
  // The stream's last fused launch must be over before another stream may
  // take its block (every completed fused launch leaves the block zeroed).
  // The wait happens outside the lock, which every dispatch takes: other
  // threads keep dispatching meanwhile.  The caller enqueues no more work on
  // the stream once it releases it.
  if(hipError_t e = hipStreamSynchronize(stream))
    return fail(e, "%s: hipStreamSynchronize: %s", who, hipGetErrorString(e));
  std::lock_guard<std::mutex> guard(gSlotLock);
  const auto it = gSlots.find(std::make_pair(device, id));
  if(it != gSlots.end())  // else the stream never fused a launch
  {
    gSlotsFree[device].push_back(it->second);
    gSlots.erase(it);
  }
  gLastError[0] = 0;
  return 0;
}
