// C-ABI of batched chains (nvproPyramidBatch*): planning of the work list,
// device upload, launches (direct or through a hipGraph) and the autotuner.
//
// Every image is planned exactly as nvproPyramidDispatch plans it (default
// dispatchers, nvpro_pyramid_dispatch.hpp:109-304 of the reference).  Its
// steps become segments; a tail segment replaces the small end of the plan the
// way the per-image executor does.  Segments are cut into items of about
// itemBytes input bytes.  Round r holds step r of every image; the work list is
// round-major, largest segments first, so an item never depends on a later one.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "batch_internal.hpp"
#include "groups.hpp"

using namespace nvpro_pyramid;

namespace {

constexpr uint32_t kDefaultItemBytes = 65536;
constexpr uint32_t kSpinLimit        = 1u << 20;  // polls of >= 50 ns: far beyond any legitimate wait

// Input bytes one group of a segment reads (for item sizing).
uint64_t groupInputBytes(uint32_t kind, uint32_t bpt)
{
  if(kind == 1u)
    return (uint64_t)kGroupBS * 4u * bpt;  // 2x2 per lane
  if(kind <= 6u)
    return (uint64_t)kGroupBS * 16u * bpt;  // 4x4 per lane
  if(kind == kSegGeneral1)
    return (uint64_t)kGroupBS * 4u * bpt;
  if(kind == kSegGeneral2)  // the input under a tile's first level
    return bpt <= 8u ? (2ull * Gen2Geom<8>::MW + 1u) * (2ull * Gen2Geom<8>::MH + 1u) * bpt
                     : (2ull * Gen2Geom<16>::MW + 1u) * (2ull * Gen2Geom<16>::MH + 1u) * bpt;
  return 1u;
}

size_t align256(size_t v) { return (v + 255u) & ~size_t(255u); }

}  // namespace

int nvpro_pyramid::checkBatchConfig(const NvproPyramidBatchConfig& c, const char* who)
{
  if(c.mode > NVPRO_PYRAMID_BATCH_MODE_QUEUE)
    return hostFail(hipErrorInvalidValue, "%s: unknown mode %u", who, c.mode);
  if(c.tailTexels > kBatchTailTexels)
    return hostFail(hipErrorInvalidValue, "%s: tailTexels %u above %u", who, c.tailTexels, kBatchTailTexels);
  if(c.reserved[0] || c.reserved[1])
    return hostFail(hipErrorInvalidValue, "%s: reserved config fields must be 0", who);
  return 0;
}

int nvpro_pyramid::buildBatchPlan(const NvproPyramidPipelines* pipelines, const NvproPyramidImage* images,
                                  const uint32_t* mipLevels, uint32_t count, const NvproPyramidBatchConfig& cfg,
                                  BatchPlan& plan, const char* who)
{
  if(!images || count == 0)
    return hostFail(hipErrorInvalidValue, "%s: no images", who);
  const uint32_t bpt       = pipelines ? nvproPyramidBytesPerTexel(pipelines->format) : 0u;
  const uint32_t itemBytes = cfg.itemBytes ? cfg.itemBytes : kDefaultItemBytes;
  // Cleared, not replaced: a serving loop plans into the same BatchPlan batch
  // after batch and keeps its vectors' capacity.
  plan.imgs.clear();
  plan.segs.clear();
  plan.items.clear();
  plan.tails.clear();
  plan.roundEnd.clear();
  plan.numSlots         = 0;
  plan.algorithmicBytes = 0;
  plan.imgs.resize(count);
  // Per image: its steps as segments (before item cutting) in step order.
  struct Pending
  {
    QueueSeg seg;
    uint32_t groups, step;
  };
  std::vector<std::vector<Pending>> perImage(count);
  uint32_t rounds = 0;
  for(uint32_t i = 0; i < count; ++i)
  {
    NvproPyramidStep steps[kMaxPlanSteps];
    uint32_t n = 0;
    if(int e = prepareImage(pipelines, &images[i], mipLevels ? mipLevels[i] : 0u, nullptr, nullptr, &plan.imgs[i],
                            steps, &n, who))
      return e;
    const ImageDesc& d = plan.imgs[i];
    if(n == 0)
      continue;
    const uint32_t lastLevel = steps[n - 1].inputLevel + steps[n - 1].levelCount;
    plan.algorithmicBytes += (uint64_t)d.lv[0].width * d.lv[0].height * bpt;
    for(uint32_t l = 1; l <= lastLevel; ++l)
      plan.algorithmicBytes += (uint64_t)d.lv[l].width * d.lv[l].height * bpt;
    for(uint32_t s = 0; s < n; ++s)
    {
      const NvproPyramidStep& st = steps[s];
      QueueSeg q{};
      q.image = i;
      q.depSlot = q.doneSlot = kNoSlot;
      if(cfg.tailTexels && firstOutputTexels(d, st.inputLevel) <= cfg.tailTexels)
      {
        TailPlan tp{};
        tp.image = i;
        for(uint32_t k = s; k < n; ++k)
          tp.steps[tp.count++] = TailStep{steps[k].pipeline, steps[k].inputLevel, steps[k].levelCount, 0u};
        q.kind       = kSegTail;
        q.inputLevel = (uint32_t)plan.tails.size();
        plan.tails.push_back(tp);
        perImage[i].push_back(Pending{q, 1u, s});
        break;
      }
      q.inputLevel = st.inputLevel;
      uint32_t groups;
      if(st.pipeline == NVPRO_PYRAMID_PIPELINE_FAST)
      {
        const LevelDesc& in = d.lv[st.inputLevel];
        q.kind              = st.levelCount;
        q.tilesX            = in.width >> st.levelCount;
        q.numTiles          = q.tilesX * (in.height >> st.levelCount);
        groups              = fastGroupCount(in.width, in.height, st.levelCount);
      }
      else
      {
        const LevelDesc& out = d.lv[st.inputLevel + st.levelCount];
        q.kind               = st.levelCount == 1u ? kSegGeneral1 : kSegGeneral2;
        groups               = generalGroupCount(out.width, out.height, st.levelCount, bpt, &q.tilesX);
        q.numTiles           = groups;
      }
      perImage[i].push_back(Pending{q, groups, s});
    }
    rounds = std::max(rounds, (uint32_t)perImage[i].size());
  }
  // Completion slots: one per segment that has a successor.  Segment ids are
  // assigned round by round below; slots per (image, step position).
  std::vector<std::vector<uint32_t>> slotOf(count), itemsOf(count);
  for(uint32_t i = 0; i < count; ++i)
  {
    slotOf[i].assign(perImage[i].size(), kNoSlot);
    itemsOf[i].assign(perImage[i].size(), 0u);
    for(size_t k = 0; k + 1 < perImage[i].size(); ++k)
      slotOf[i][k] = plan.numSlots++;
  }
  for(uint32_t r = 0; r < rounds; ++r)
  {
    std::vector<std::pair<uint32_t, uint32_t>> order;  // (groups, image)
    for(uint32_t i = 0; i < count; ++i)
      if(r < perImage[i].size())
        order.push_back({perImage[i][r].groups, i});
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
                       return a.first > b.first;
                     });
    for(const auto& o : order)
    {
      const uint32_t i = o.second;
      Pending& p       = perImage[i][r];
      QueueSeg q       = p.seg;
      if(r > 0)
      {
        q.depSlot  = slotOf[i][r - 1];
        q.depCount = itemsOf[i][r - 1];
      }
      q.doneSlot             = slotOf[i][r];
      const uint32_t segId   = (uint32_t)plan.segs.size();
      const uint64_t perItem = std::max<uint64_t>(1u, itemBytes / groupInputBytes(q.kind, bpt));
      uint32_t nItems        = 0;
      for(uint64_t g = 0; g < p.groups; g += perItem, ++nItems)
        plan.items.push_back(
            QueueItem{segId, (uint32_t)g, (uint32_t)std::min<uint64_t>(perItem, p.groups - g), 0u});
      itemsOf[i][r] = nItems;
      plan.segs.push_back(q);
    }
    plan.roundEnd.push_back((uint32_t)plan.items.size());
  }
  return 0;
}

void nvpro_pyramid::defaultBatchConfig(NvproPyramidBatchConfig* c)
{
  memset(c, 0, sizeof *c);
  c->mode            = NVPRO_PYRAMID_BATCH_MODE_QUEUE;
  c->workgroupsPerCU = 0;
  c->itemBytes       = kDefaultItemBytes;
  c->tailTexels      = kBatchTailTexels;
  c->useGraph        = 1;
}

// ---------------------------------------------------------------------------
// BatchRunner
// ---------------------------------------------------------------------------

BatchRunner::~BatchRunner()
{
  if(dev_)
    (void)hipFree(dev_);
  if(pinned_)
    (void)hipHostFree(pinned_);
}

namespace {

// Device layout of a work list: images | segments | items | tails | extra |
// counters (zeroed; queue mode only).
struct WorkListLayout
{
  size_t oImg, oSeg, oItm, oTl, oExt, oCnt, size;
};

WorkListLayout layoutOf(const BatchPlan& p, size_t extraBytes, bool counters)
{
  WorkListLayout l;
  l.oImg = 0;
  l.oSeg = align256(l.oImg + p.imgs.size() * sizeof(ImageDesc));
  l.oItm = align256(l.oSeg + p.segs.size() * sizeof(QueueSeg));
  l.oTl  = align256(l.oItm + p.items.size() * sizeof(QueueItem));
  l.oExt = align256(l.oTl + p.tails.size() * sizeof(TailPlan));
  l.oCnt = align256(l.oExt + extraBytes);
  l.size = align256(l.oCnt + (counters ? batchCounterWords(p.numSlots) * sizeof(uint32_t) : 0u));
  return l;
}

// The host image of everything before the counters.
void writeWorkList(uint8_t* h, const BatchPlan& p, const WorkListLayout& l, const void* extra, size_t extraBytes)
{
  memset(h, 0, l.oCnt);
  memcpy(h + l.oImg, p.imgs.data(), p.imgs.size() * sizeof(ImageDesc));
  if(!p.segs.empty())
    memcpy(h + l.oSeg, p.segs.data(), p.segs.size() * sizeof(QueueSeg));
  if(!p.items.empty())
    memcpy(h + l.oItm, p.items.data(), p.items.size() * sizeof(QueueItem));
  if(!p.tails.empty())
    memcpy(h + l.oTl, p.tails.data(), p.tails.size() * sizeof(TailPlan));
  if(extra && extraBytes)
    memcpy(h + l.oExt, extra, extraBytes);
}

}  // namespace

void BatchRunner::bind(const BatchPlan& p, const NvproPyramidPipelines& pipes, const NvproPyramidBatchConfig& cfg,
                       uint8_t* base, size_t oImg, size_t oSeg, size_t oItm, size_t oTl, size_t oCnt, size_t size)
{
  args_.imgs      = (const ImageDesc*)(base + oImg);
  args_.segs      = (const QueueSeg*)(base + oSeg);
  args_.items     = (const QueueItem*)(base + oItm);
  args_.tails     = (const TailPlan*)(base + oTl);
  args_.counters  = (uint32_t*)(base + oCnt);
  args_.numSlots  = p.numSlots;
  args_.spinLimit = cfg.spinLimit ? cfg.spinLimit : kSpinLimit;
  format_         = pipes.format;
  arithmetic_     = pipes.arithmetic;
  queue_          = cfg.mode == NVPRO_PYRAMID_BATCH_MODE_QUEUE;
  roundEnd_       = p.roundEnd;
  items_          = (uint32_t)p.items.size();
  const uint32_t fit   = batchMaxWorkgroupsPerCU(format_, arithmetic_, queue_);
  const uint32_t perCU = cfg.workgroupsPerCU ? std::min(cfg.workgroupsPerCU, fit) : fit;
  blocks               = (uint32_t)std::max<uint64_t>(1u, std::min<uint64_t>((uint64_t)deviceComputeUnits() * perCU,
                                                                                items_));
  launches             = 0;
  uint32_t begin       = 0;
  for(uint32_t end : roundEnd_)
  {
    launches += end > begin ? 1u : 0u;
    begin = end;
  }
  if(queue_ && items_ != 0)
    launches = 1u;
  devBytes = size;
}

hipError_t BatchRunner::stage(const BatchPlan& p, const NvproPyramidPipelines& pipes,
                              const NvproPyramidBatchConfig& cfg, hipStream_t stream, const void* extra,
                              size_t extraBytes)
{
  const WorkListLayout l = layoutOf(p, extra ? extraBytes : 0u, true);
  // Grow geometrically (the previous buffers are idle: the caller staged
  // nothing in flight).  A serving loop stages a new work list per batch;
  // growing to the exact size would reallocate often, and hipFree waits for
  // the whole device.
  if(l.size > devCap_ || l.size > pinnedCap_)
    if(hipError_t e = reserve(std::max(l.size, 2u * std::min(devCap_, pinnedCap_))))
      return e;
  writeWorkList((uint8_t*)pinned_, p, l, extra, extraBytes);
  // The work list is copied, the counters after it are zeroed on the device.
  if(hipError_t e = launchStageCopy(stream, dev_, pinned_, l.oCnt, l.size))
    return e;
  uint8_t* base = (uint8_t*)dev_;
  extra_        = extra ? base + l.oExt : nullptr;
  bind(p, pipes, cfg, base, l.oImg, l.oSeg, l.oItm, l.oTl, l.oCnt, l.size);
  return hipSuccess;
}

size_t BatchRunner::roundsWorkListBytes(const BatchPlan& p) { return layoutOf(p, 0u, false).size; }

hipError_t BatchRunner::stageRoundsInto(const BatchPlan& p, const NvproPyramidPipelines& pipes,
                                        const NvproPyramidBatchConfig& cfg, hipStream_t stream, void* dev,
                                        std::vector<uint8_t>& host)
{
  if(cfg.mode != NVPRO_PYRAMID_BATCH_MODE_ROUNDS)
    return hipErrorInvalidValue;
  const WorkListLayout l = layoutOf(p, 0u, false);
  if(host.size() < l.size)
    host.resize(l.size);
  writeWorkList(host.data(), p, l, nullptr, 0u);
  // A copy from ordinary (pageable) host memory: the runtime has taken the
  // bytes when the call returns, so `host` may be reused at once.
  if(hipError_t e = hipMemcpyAsync(dev, host.data(), l.oCnt, hipMemcpyHostToDevice, stream))
    return e;
  extra_ = nullptr;
  bind(p, pipes, cfg, (uint8_t*)dev, l.oImg, l.oSeg, l.oItm, l.oTl, l.oCnt, l.size);
  return hipSuccess;
}

hipError_t BatchRunner::launch(hipStream_t stream) const
{
  if(items_ == 0)
    return hipSuccess;
  if(queue_)
  {
    BatchArgs a  = args_;
    a.itemBegin = 0;
    a.itemEnd   = items_;
    return launchBatch(stream, format_, arithmetic_, true, a, blocks);
  }
  uint32_t begin = 0;
  for(uint32_t end : roundEnd_)
  {
    if(end > begin)
    {
      BatchArgs a  = args_;
      a.itemBegin = begin;
      a.itemEnd   = end;
      if(hipError_t e = launchBatch(stream, format_, arithmetic_, false, a, std::min(end - begin, blocks)))
        return e;
    }
    begin = end;
  }
  return hipSuccess;
}

hipError_t BatchRunner::reserve(size_t bytes)
{
  bytes = align256(bytes);
  if(bytes > devCap_)
  {
    if(dev_)
      (void)hipFree(dev_);
    dev_    = nullptr;
    devCap_ = 0;
    if(hipError_t e = hipMalloc(&dev_, bytes))
      return e;
    devCap_ = bytes;
  }
  if(bytes > pinnedCap_)
  {
    if(pinned_)
      (void)hipHostFree(pinned_);
    pinned_    = nullptr;
    pinnedCap_ = 0;
    if(hipError_t e = hipHostMalloc(&pinned_, bytes, hipHostMallocDefault))
      return e;
    pinnedCap_ = bytes;
  }
  return hipSuccess;
}

uint32_t* BatchRunner::errorWord() const { return args_.counters ? args_.counters + kCounterError : nullptr; }

struct NvproPyramidBatch
{
  NvproPyramidBatchConfig cfg{};
  BatchPlan plan;
  BatchRunner runner;
  hipStream_t priv    = nullptr;
  hipGraph_t graph    = nullptr;
  hipGraphExec_t exec = nullptr;
};

static void releaseBatch(NvproPyramidBatch* b)
{
  if(b->exec)
    (void)hipGraphExecDestroy(b->exec);
  if(b->graph)
    (void)hipGraphDestroy(b->graph);
  if(b->priv)
  {
    (void)hipStreamSynchronize(b->priv);
    (void)hipStreamDestroy(b->priv);
  }
  delete b;
}

extern "C" {

int nvproPyramidBatchDefaultConfig(NvproPyramidBatchConfig* config)
{
  if(!config)
    return hostFail(hipErrorInvalidValue, "nvproPyramidBatchDefaultConfig: null config");
  defaultBatchConfig(config);
  return 0;
}

int nvproPyramidBatchCreate(const NvproPyramidPipelines* pipelines, const NvproPyramidImage* images,
                            const uint32_t* mipLevels, uint32_t imageCount, const NvproPyramidBatchConfig* config,
                            NvproPyramidBatch** batch)
{
  const char* who = "nvproPyramidBatchCreate";
  if(!batch)
    return hostFail(hipErrorInvalidValue, "%s: null output", who);
  *batch = nullptr;
  NvproPyramidBatchConfig cfg;
  defaultBatchConfig(&cfg);
  if(config)
    cfg = *config;
  if(int e = checkBatchConfig(cfg, who))
    return e;
  NvproPyramidBatch* b = new NvproPyramidBatch;
  b->cfg               = cfg;
  if(int e = buildBatchPlan(pipelines, images, mipLevels, imageCount, cfg, b->plan, who))
  {
    delete b;
    return e;
  }
  hipError_t e = hipStreamCreateWithFlags(&b->priv, hipStreamNonBlocking);
  if(e == hipSuccess)
    e = b->runner.stage(b->plan, *pipelines, cfg, b->priv);
  if(e == hipSuccess)
    e = hipStreamSynchronize(b->priv);
  if(e != hipSuccess)
  {
    releaseBatch(b);
    return hostFail(e, "%s: device work list: %s", who, hipGetErrorString(e));
  }
  if(cfg.useGraph && !b->plan.items.empty())
  {
    e = hipStreamBeginCapture(b->priv, hipStreamCaptureModeThreadLocal);
    if(e == hipSuccess)
    {
      hipError_t le = b->runner.launch(b->priv);
      e             = hipStreamEndCapture(b->priv, &b->graph);
      if(le != hipSuccess)
        e = le;
    }
    if(e == hipSuccess)
      e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if(e != hipSuccess)
    {
      releaseBatch(b);
      return hostFail(e, "%s: graph: %s", who, hipGetErrorString(e));
    }
  }
  *batch = b;
  hostClearError();
  return 0;
}

int nvproPyramidBatchLaunch(NvproPyramidBatch* batch, hipStream_t stream)
{
  if(!batch)
    return hostFail(hipErrorInvalidValue, "nvproPyramidBatchLaunch: null batch");
  hipError_t e = batch->exec ? hipGraphLaunch(batch->exec, stream) : batch->runner.launch(stream);
  if(e != hipSuccess)
    return hostFail(e, "nvproPyramidBatchLaunch: %s", hipGetErrorString(e));
  return 0;
}

int nvproPyramidBatchGetInfo(const NvproPyramidBatch* batch, NvproPyramidBatchInfo* info)
{
  if(!batch || !info)
    return hostFail(hipErrorInvalidValue, "nvproPyramidBatchGetInfo: null argument");
  memset(info, 0, sizeof *info);
  info->imageCount       = (uint32_t)batch->plan.imgs.size();
  info->segmentCount     = (uint32_t)batch->plan.segs.size();
  info->itemCount        = (uint32_t)batch->plan.items.size();
  info->roundCount       = (uint32_t)batch->plan.roundEnd.size();
  info->launchCount      = batch->runner.launches;
  info->workgroups       = batch->plan.items.empty() ? 0u : batch->runner.blocks;
  info->algorithmicBytes = batch->plan.algorithmicBytes;
  info->deviceBytes      = batch->runner.devBytes;
  info->config           = batch->cfg;
  return 0;
}

int nvproPyramidBatchReadErrors(NvproPyramidBatch* batch, uint32_t* flags)
{
  if(!batch || !flags)
    return hostFail(hipErrorInvalidValue, "nvproPyramidBatchReadErrors: null argument");
  hipError_t e = hipDeviceSynchronize();
  uint32_t v   = 0;
  uint32_t* p  = batch->runner.errorWord();
  if(e == hipSuccess)
    e = hipMemcpyAsync(&v, p, sizeof v, hipMemcpyDeviceToHost, batch->priv);
  const uint32_t zero = 0;
  if(e == hipSuccess)
    e = hipMemcpyAsync(p, &zero, sizeof zero, hipMemcpyHostToDevice, batch->priv);
  if(e == hipSuccess)
    e = hipStreamSynchronize(batch->priv);
  if(e != hipSuccess)
    return hostFail(e, "nvproPyramidBatchReadErrors: %s", hipGetErrorString(e));
  *flags = v;
  return 0;
}

int nvproPyramidBatchDestroy(NvproPyramidBatch* batch)
{
  if(batch)
    releaseBatch(batch);
  return 0;
}

int nvproPyramidBatchAutotune(const NvproPyramidPipelines* pipelines, const NvproPyramidImage* images,
                              const uint32_t* mipLevels, uint32_t imageCount, hipStream_t stream,
                              uint32_t repetitions, NvproPyramidBatchConfig* best, float* bestMicroseconds)
{
  const char* who = "nvproPyramidBatchAutotune";
  if(!best)
    return hostFail(hipErrorInvalidValue, "%s: null output", who);
  if(!pipelines || !images || imageCount == 0)
    return hostFail(hipErrorInvalidValue, "%s: no images", who);
  // Cache key: device, format, arithmetic and the exact list of image sizes and
  // level counts (two batches with the same count and texels but different
  // shapes can want different configurations).
  int dev = 0;
  (void)hipGetDevice(&dev);
  using Key = std::vector<uint64_t>;
  Key key{(uint64_t)dev, pipelines->format, pipelines->arithmetic, imageCount};
  key.reserve(4u + 2u * imageCount);
  for(uint32_t i = 0; i < imageCount; ++i)
  {
    key.push_back((uint64_t)images[i].width << 32 | images[i].height);
    key.push_back(mipLevels ? mipLevels[i] : 0u);
  }
  static std::mutex mtx;
  static std::map<Key, std::pair<NvproPyramidBatchConfig, float>> cache;
  {
    std::lock_guard<std::mutex> lock(mtx);
    auto it = cache.find(key);
    if(it != cache.end())
    {
      *best = it->second.first;
      if(bestMicroseconds)
        *bestMicroseconds = it->second.second;
      return 0;
    }
  }
  const uint32_t reps = repetitions ? repetitions : 5u;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t he = hipEventCreate(&e0);
  if(he == hipSuccess)
    he = hipEventCreate(&e1);
  if(he != hipSuccess)
    return hostFail(he, "%s: events: %s", who, hipGetErrorString(he));
  float bestUs = INFINITY;
  NvproPyramidBatchConfig winner;
  nvproPyramidBatchDefaultConfig(&winner);
  int status = 0;
  for(uint32_t mode : {NVPRO_PYRAMID_BATCH_MODE_QUEUE, NVPRO_PYRAMID_BATCH_MODE_ROUNDS})
    for(uint32_t perCU : {0u, 4u})
      for(uint32_t itemBytes : {16384u, 65536u, 262144u})
        for(uint32_t tail : {0u, kBatchTailTexels})
        {
          NvproPyramidBatchConfig c;
          nvproPyramidBatchDefaultConfig(&c);
          c.mode            = mode;
          c.workgroupsPerCU = perCU;
          c.itemBytes       = itemBytes;
          c.tailTexels      = tail;
          NvproPyramidBatch* b = nullptr;
          if((status = nvproPyramidBatchCreate(pipelines, images, mipLevels, imageCount, &c, &b)) != 0)
            goto done;
          status = nvproPyramidBatchLaunch(b, stream);  // warm-up
          if(status == 0 && (he = hipEventRecord(e0, stream)) != hipSuccess)
            status = hostFail(he, "%s: %s", who, hipGetErrorString(he));
          for(uint32_t r = 0; status == 0 && r < reps; ++r)
            status = nvproPyramidBatchLaunch(b, stream);
          if(status == 0 && (he = hipEventRecord(e1, stream)) != hipSuccess)
            status = hostFail(he, "%s: %s", who, hipGetErrorString(he));
          if(status == 0 && (he = hipEventSynchronize(e1)) != hipSuccess)
            status = hostFail(he, "%s: %s", who, hipGetErrorString(he));
          float ms = 0.0f;
          if(status == 0 && (he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess)
            status = hostFail(he, "%s: %s", who, hipGetErrorString(he));
          uint32_t flags = 0;
          if(status == 0)
            status = nvproPyramidBatchReadErrors(b, &flags);
          nvproPyramidBatchDestroy(b);
          if(status != 0)
            goto done;
          const float us = ms * 1000.0f / (float)reps;
          if(flags == 0 && us < bestUs)
          {
            bestUs = us;
            winner = c;
          }
        }
done:
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if(status != 0)
    return status;
  if(!(bestUs < INFINITY))
    return hostFail(hipErrorUnknown, "%s: every candidate reported errors", who);
  {
    std::lock_guard<std::mutex> lock(mtx);
    cache[key] = {winner, bestUs};
  }
  *best = winner;
  if(bestMicroseconds)
    *bestMicroseconds = bestUs;
  hostClearError();
  return 0;
}

}  // extern "C"
