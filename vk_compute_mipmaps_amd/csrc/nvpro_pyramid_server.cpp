// C-ABI of the serving loop (nvproPyramidServer*).  No reference counterpart:
// the reference regenerates one image's chain per frame (demo_app/
// mipmaps_app.cpp:362-369); this is a request loop around batched chains.
//
// A server owns one device arena of `budgetBytes` (a single hipMalloc at
// creation) from which every admitted request's chain is placed (first fit,
// 256-byte aligned) and where it stays resident until the client releases it.
// Requests wait in a FIFO until their chain fits.  Each poll:
//   1. retires in-flight batches whose completion event has fired,
//   2. admits waiting requests in arrival order while the arena has room,
//   3. groups the admitted requests by format into batches (at most
//      maxBatchImages each) and launches each batch on a free slot: its own
//      stream, event and reusable work list.  Level 0 is written first (copied
//      from the client's host texels, or the seeded synthetic image), then the
//      batched chain kernels run.
// Up to maxInFlight batches run at once on separate streams, so requests that
// arrive while others are on the device are admitted and overlap with them.
// Every chain is planned exactly as nvproPyramidDispatch plans it, so a result
// does not depend on what it was batched with.
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "batch_internal.hpp"

using namespace nvpro_pyramid;

namespace {

using Clock = std::chrono::steady_clock;

double usSince(Clock::time_point a, Clock::time_point b)
{
  return std::chrono::duration<double, std::micro>(b - a).count();
}

// Best-fit allocator over [0, size) with coalescing frees.  Free ranges are
// indexed by offset (to coalesce with neighbours) and by size (to find the
// smallest that fits), so an allocation or a free costs O(log n) in the number
// of free ranges.  (A first-fit scan of an offset-ordered list cost the serving
// loop tens of milliseconds per poll once thousands of resident chains had
// fragmented an 80 GB arena.)
class Arena
{
public:
  void reset(uint64_t size)
  {
    byOffset_.clear();
    bySize_.clear();
    if(size)
      insert(0, size);
    used_ = peak_ = 0;
  }
  bool alloc(uint64_t bytes, uint64_t* off)
  {
    bytes   = (bytes + 255u) & ~uint64_t(255u);
    auto it = bySize_.lower_bound(bytes);
    if(it == bySize_.end())
      return false;
    const uint64_t size = it->first, start = it->second;
    bySize_.erase(it);
    byOffset_.erase(start);
    if(size > bytes)
      insert(start + bytes, size - bytes);
    *off = start;
    used_ += bytes;
    peak_ = used_ > peak_ ? used_ : peak_;
    return true;
  }
  void release(uint64_t off, uint64_t bytes)
  {
    bytes = (bytes + 255u) & ~uint64_t(255u);
    used_ -= bytes;
    auto next = byOffset_.lower_bound(off);
    if(next != byOffset_.end() && off + bytes == next->first)
    {
      bytes += next->second;
      eraseSized(next->first, next->second);
      next = byOffset_.erase(next);
    }
    if(next != byOffset_.begin())
    {
      auto prev = std::prev(next);
      if(prev->first + prev->second == off)
      {
        off = prev->first;
        bytes += prev->second;
        eraseSized(prev->first, prev->second);
        byOffset_.erase(prev);
      }
    }
    insert(off, bytes);
  }
  uint64_t used() const { return used_; }
  uint64_t peak() const { return peak_; }
  size_t freeRanges() const { return byOffset_.size(); }

private:
  void insert(uint64_t off, uint64_t size)
  {
    byOffset_[off] = size;
    bySize_.emplace(size, off);
  }
  void eraseSized(uint64_t off, uint64_t size)
  {
    auto range = bySize_.equal_range(size);
    for(auto it = range.first; it != range.second; ++it)
      if(it->second == off)
      {
        bySize_.erase(it);
        return;
      }
  }
  std::map<uint64_t, uint64_t> byOffset_;     // offset -> size
  std::multimap<uint64_t, uint64_t> bySize_;  // size -> offset
  uint64_t used_ = 0, peak_ = 0;
};

enum class State
{
  Waiting,
  Admitted,
  Running,
  Done,
};

struct Request
{
  NvproPyramidRequest req;
  uint64_t ticket;
  State state;
  uint64_t offset = 0, bytes = 0;
  Clock::time_point submitted, admitted;
  int status = 0;
};

struct Slot
{
  hipStream_t stream = nullptr;
  hipEvent_t done    = nullptr;
  uint32_t* errHost  = nullptr;  // pinned copy of the batch's error flags (queue mode writes it)
  BatchRunner runner;           // work list, and the level-0 generator's jobs and chunks
  std::vector<uint64_t> tickets;
  bool busy      = false;
  int failStatus = 0;  // nonzero: the batch could not be launched; its requests complete with this
};

}  // namespace

struct NvproPyramidServer
{
  NvproPyramidServerConfig cfg{};
  uint8_t* arenaBase = nullptr;
  Arena arena;
  std::vector<Slot*> slots;
  std::unordered_map<uint64_t, Request> requests;
  std::deque<uint64_t> waiting;      // FIFO of tickets not yet admitted
  std::vector<uint64_t> admitted;    // admitted, not yet launched
  std::deque<NvproPyramidCompletion> completions;
  uint64_t nextTicket = 1;
  NvproPyramidServerStats stats{};
  std::mutex mtx;
  // Per-batch host scratch, reused from batch to batch (a batch of a thousand
  // images fills megabytes; new vectors each time cost page faults).
  std::vector<uint64_t> take, keep;
  std::vector<NvproPyramidImage> images;
  std::vector<uint32_t> mips;
  std::vector<SynthJob> jobs;
  std::vector<SynthChunk> chunks;
  std::vector<uint8_t> table;
  BatchPlan plan;
  // Test hook: the next this-many batch launches fail after their staging copy
  // is enqueued (NVPRO_PYRAMID_SERVER_TEST_FAIL_LAUNCHES at creation).
  uint32_t testFailLaunches = 0;
};

static uint64_t chainBytes(const NvproPyramidRequest& r)
{
  uint64_t total = 0;
  const uint32_t bpt = nvproPyramidBytesPerTexel(r.format);
  if(nvproPyramidPackedLayout(r.width, r.height, 0, bpt, nullptr, &total) != 0)
    return 0;
  return total;
}

static NvproPyramidImage imageOf(const NvproPyramidServer* s, const Request& r)
{
  return NvproPyramidImage{s->arenaBase + r.offset, r.req.width, r.req.height, 0, 0, nullptr, nullptr};
}

// Retire finished slots (step 1).  Caller holds the lock.
static int retire(NvproPyramidServer* s, Clock::time_point now)
{
  for(Slot* slot : s->slots)
  {
    if(!slot->busy)
      continue;
    const hipError_t q = hipEventQuery(slot->done);
    if(q == hipErrorNotReady)
      continue;
    const int status = q != hipSuccess       ? (int)q
                       : slot->failStatus    ? slot->failStatus
                       : *slot->errHost      ? (int)hipErrorUnknown
                                             : 0;
    if(slot->failStatus || *slot->errHost)
      s->stats.failedBatches++;
    slot->failStatus = 0;
    for(uint64_t t : slot->tickets)
    {
      Request& r = s->requests[t];
      r.state    = State::Done;
      r.status   = status;
      NvproPyramidCompletion c{};
      c.ticket     = t;
      c.status     = status;
      c.image      = imageOf(s, r);
      c.bytes      = r.bytes;
      c.format     = r.req.format;
      c.latencyUs  = usSince(r.submitted, now);
      c.waitUs     = usSince(r.submitted, r.admitted);
      s->completions.push_back(c);
      s->stats.completed++;
    }
    slot->tickets.clear();
    slot->busy = false;
    if(q != hipSuccess)
      return hostFail(q, "nvproPyramidServerPoll: batch failed: %s", hipGetErrorString(q));
  }
  return 0;
}

// Admit waiting requests in order while their chains fit (step 2).
static void admit(NvproPyramidServer* s, Clock::time_point now)
{
  while(!s->waiting.empty())
  {
    Request& r = s->requests[s->waiting.front()];
    uint64_t off;
    if(!s->arena.alloc(r.bytes, &off))
      break;
    r.offset   = off;
    r.state    = State::Admitted;
    r.admitted = now;
    s->admitted.push_back(r.ticket);
    s->waiting.pop_front();
    s->stats.admitted++;
  }
  s->stats.residentBytes     = s->arena.used();
  s->stats.peakResidentBytes = s->arena.peak();
}

// Launch admitted requests on free slots, one format per batch (step 3).
static int launch(NvproPyramidServer* s)
{
  const char* who = "nvproPyramidServerPoll";
  while(!s->admitted.empty())
  {
    Slot* slot = nullptr;
    for(Slot* c : s->slots)
      if(!c->busy)
      {
        slot = c;
        break;
      }
    if(!slot)
      return 0;
    const uint32_t format = s->requests[s->admitted.front()].req.format;
    std::vector<uint64_t>& take = s->take;
    std::vector<uint64_t>& keep = s->keep;
    take.clear();
    keep.clear();
    for(uint64_t t : s->admitted)
    {
      if(s->requests[t].req.format == format && take.size() < s->cfg.maxBatchImages)
        take.push_back(t);
      else
        keep.push_back(t);
    }
    s->admitted.swap(keep);
    // From here on a failure must not lose the taken requests: they complete
    // with the error (their chains stay resident until released) once the work
    // already enqueued on the slot's stream, which may still read the slot's
    // staging, has finished; retire() does that when the slot's event fires.
    auto abandon = [&](int code, const char* what) -> int {
      slot->tickets    = take;
      slot->failStatus = code ? code : (int)hipErrorUnknown;
      if(hipEventRecord(slot->done, slot->stream) != hipSuccess)
      {
        (void)hipStreamSynchronize(slot->stream);  // then the event's last record has fired
        (void)hipGetLastError();
      }
      slot->busy = true;
      s->stats.batches++;
      s->stats.batchedImages += take.size();
      return hostFail(code, "%s: %s: %s", who, what, hipGetErrorString((hipError_t)code));
    };
    std::vector<NvproPyramidImage>& images = s->images;
    std::vector<uint32_t>& mips            = s->mips;
    std::vector<SynthJob>& jobs            = s->jobs;
    std::vector<SynthChunk>& chunks        = s->chunks;
    images.clear();
    mips.clear();
    jobs.clear();
    chunks.clear();
    const uint32_t bpt   = nvproPyramidBytesPerTexel(format);
    uint64_t synthTexels = 0;
    for(uint64_t t : take)
    {
      Request& r = s->requests[t];
      images.push_back(imageOf(s, r));
      mips.push_back(r.req.mipLevels);
      if(r.req.level0)
      {
        const uint64_t n = (uint64_t)r.req.width * r.req.height * bpt;
        if(hipError_t he = hipMemcpyAsync(images.back().data, r.req.level0, n, hipMemcpyHostToDevice, slot->stream))
        {
          for(uint64_t u : take)
            s->requests[u].state = State::Running;
          return abandon(he, "level 0 upload");
        }
      }
      else
      {
        jobs.push_back(SynthJob{(uint8_t*)images.back().data, r.req.seed, r.req.width, r.req.height,
                                r.req.width * bpt, format});
        synthTexels += (uint64_t)r.req.width * r.req.height;
      }
      r.state = State::Running;
    }
    // Chunks of the level-0 generator: kSynthChunk texels, or larger when the
    // batch alone gives the launch enough of them (about 8k), so the table
    // stays small (a thousand mixed images in 4096-texel chunks are 200k
    // entries, 5 MB, to write and copy per batch).  Chunks of up to 65,536
    // texels keep the generator's lattice cache usable at any width.
    uint64_t chunk = kSynthChunk;
    while(chunk < 65536u && synthTexels / chunk > 8192u)
      chunk *= 2u;
    for(uint32_t j = 0; j < (uint32_t)jobs.size(); ++j)
      appendSynthChunks(chunks, j, jobs[j].width, jobs[j].height, chunk);
    // The level-0 generator's table (jobs, then chunks) travels with the work
    // list in one staging copy; one launch then writes every synthetic level 0
    // of the batch.
    std::vector<uint8_t>& table = s->table;
    table.resize(jobs.size() * sizeof(SynthJob) + chunks.size() * sizeof(SynthChunk));
    if(!chunks.empty())
    {
      memcpy(table.data(), jobs.data(), jobs.size() * sizeof(SynthJob));
      memcpy(table.data() + jobs.size() * sizeof(SynthJob), chunks.data(), chunks.size() * sizeof(SynthChunk));
    }
    const NvproPyramidPipelines pipes{format, 1u,
                                      format == NVPRO_PYRAMID_FORMAT_RGBA16F ? s->cfg.arithmetic16f
                                                                             : NVPRO_PYRAMID_ARITH_FP32,
                                      0u};
    BatchPlan& plan = s->plan;
    if(int e = buildBatchPlan(&pipes, images.data(), mips.data(), (uint32_t)images.size(), s->cfg.batch, plan, who))
      return abandon(e, "batch plan");
    *slot->errHost = 0;  // the slot is idle: no launch writes it now
    hipError_t e   = slot->runner.stage(plan, pipes, s->cfg.batch, slot->stream, chunks.empty() ? nullptr : table.data(),
                                        table.size());
    if(e == hipSuccess && s->testFailLaunches)
    {
      --s->testFailLaunches;
      e = hipErrorLaunchFailure;
    }
    if(e == hipSuccess && !chunks.empty())
    {
      const uint8_t* dev = (const uint8_t*)slot->runner.extraDevice();
      e = launchSynthBatch(slot->stream, (const SynthJob*)dev, (const SynthChunk*)(dev + jobs.size() * sizeof(SynthJob)),
                           (uint32_t)chunks.size());
    }
    if(e == hipSuccess)
      e = slot->runner.launch(slot->stream);
    if(e != hipSuccess)
      return abandon(e, "batch launch");
    if(hipError_t re = hipEventRecord(slot->done, slot->stream))
      return abandon(re, "completion event");
    slot->tickets = take;
    slot->busy    = true;
    s->stats.batches++;
    s->stats.batchedImages += take.size();
  }
  return 0;
}

// Before the first request: one small synthetic chain per format through every
// slot's generator and batch path, then a synchronize.  The first launch of a
// kernel in a process pays for loading its code (milliseconds); without this
// the first few hundred requests of a stream of arrivals would.  Uses the
// (still empty) arena's first bytes.
static hipError_t warmUp(NvproPyramidServer* s)
{
  const char* who = "nvproPyramidServerCreate";
  for(Slot* slot : s->slots)
  {
    for(uint32_t format : {NVPRO_PYRAMID_FORMAT_SRGBA8, NVPRO_PYRAMID_FORMAT_RGBA16F})
    {
      const uint32_t w = 67, h = 35, bpt = nvproPyramidBytesPerTexel(format);  // odd: general steps and a tail
      uint64_t bytes   = 0;
      if(nvproPyramidPackedLayout(w, h, 0, bpt, nullptr, &bytes) != 0 || bytes > s->cfg.budgetBytes)
        continue;  // an arena too small to hold it: the first request loads the kernels instead
      const NvproPyramidImage img{s->arenaBase, w, h, 0, 0, nullptr, nullptr};
      struct
      {
        SynthJob job;
        SynthChunk chunk;
      } table{SynthJob{(uint8_t*)s->arenaBase, 1u, w, h, w * bpt, format}, SynthChunk{0u, 0u, 0u, h}};
      const NvproPyramidPipelines pipes{
          format, 1u, format == NVPRO_PYRAMID_FORMAT_RGBA16F ? s->cfg.arithmetic16f : NVPRO_PYRAMID_ARITH_FP32, 0u};
      const uint32_t mips = 0;
      BatchPlan plan;
      if(int pe = buildBatchPlan(&pipes, &img, &mips, 1u, s->cfg.batch, plan, who))
        return (hipError_t)pe;
      hipError_t e = slot->runner.stage(plan, pipes, s->cfg.batch, slot->stream, &table, sizeof table);
      if(e == hipSuccess)
      {
        const uint8_t* dev = (const uint8_t*)slot->runner.extraDevice();
        e = launchSynthBatch(slot->stream, (const SynthJob*)dev,
                             (const SynthChunk*)(dev + offsetof(decltype(table), chunk)), 1u);
      }
      if(e == hipSuccess)
        e = slot->runner.launch(slot->stream);
      if(e == hipSuccess)
        e = hipStreamSynchronize(slot->stream);
      if(e != hipSuccess)
        return e;
    }
  }
  return hipSuccess;
}

static void destroyServer(NvproPyramidServer* s)
{
  for(Slot* slot : s->slots)
  {
    if(slot->stream)
      (void)hipStreamSynchronize(slot->stream);
    if(slot->done)
      (void)hipEventDestroy(slot->done);
    if(slot->errHost)
      (void)hipHostFree(slot->errHost);
    if(slot->stream)
      (void)hipStreamDestroy(slot->stream);
    delete slot;
  }
  if(s->arenaBase)
    (void)hipFree(s->arenaBase);
  delete s;
}

extern "C" {

int nvproPyramidServerDefaultConfig(NvproPyramidServerConfig* config)
{
  if(!config)
    return hostFail(hipErrorInvalidValue, "nvproPyramidServerDefaultConfig: null config");
  memset(config, 0, sizeof *config);
  config->budgetBytes    = 80ull * 1000ull * 1000ull * 1000ull;
  config->maxInFlight    = 4;
  config->maxBatchImages = 1024;
  config->arithmetic16f  = NVPRO_PYRAMID_ARITH_FP32;
  defaultBatchConfig(&config->batch);
  config->batch.mode     = NVPRO_PYRAMID_BATCH_MODE_ROUNDS;  // the faster mode on mixed-size batches
  config->batch.useGraph = 0;                                // a new work list every batch: direct launches
  return 0;
}

int nvproPyramidServerCreate(const NvproPyramidServerConfig* config, NvproPyramidServer** server)
{
  const char* who = "nvproPyramidServerCreate";
  if(!server)
    return hostFail(hipErrorInvalidValue, "%s: null output", who);
  *server = nullptr;
  NvproPyramidServerConfig cfg;
  nvproPyramidServerDefaultConfig(&cfg);
  if(config)
    cfg = *config;
  if(cfg.budgetBytes == 0 || cfg.maxInFlight == 0 || cfg.maxInFlight > 16 || cfg.maxBatchImages == 0)
    return hostFail(hipErrorInvalidValue, "%s: budget, maxInFlight (1..16) and maxBatchImages must be set", who);
  if(cfg.arithmetic16f > NVPRO_PYRAMID_ARITH_PACKED_FP16)
    return hostFail(hipErrorInvalidValue, "%s: unknown arithmetic %u", who, cfg.arithmetic16f);
  if(cfg.reserved[0] || cfg.reserved[1])
    return hostFail(hipErrorInvalidValue, "%s: reserved fields must be 0", who);
  if(cfg.batch.useGraph)
    return hostFail(hipErrorInvalidValue, "%s: batches are launched directly (batch.useGraph must be 0)", who);
  if(int e = checkBatchConfig(cfg.batch, who))
    return e;
  NvproPyramidServer* s = new NvproPyramidServer;
  s->cfg                = cfg;
  hipError_t e          = hipMalloc(&s->arenaBase, cfg.budgetBytes);
  // A work list with the level-0 generator's table takes up to about 4 KB per
  // image of a mixed-size batch (sizes to 4096, 64 KiB items; up to 6 KB with
  // RGBA16F's larger items count); sized now for a full batch of such images,
  // so that the staging of ordinary batches never grows (growth frees, and
  // hipFree waits for every batch in flight: 8-16 ms stalls of the poll loop
  // were measured with the staging sized at 4 KB per image).
  const size_t stageReserve =
      std::min<size_t>(32u << 20, std::max<size_t>(256u << 10, (size_t)cfg.maxBatchImages * 8192u));
  for(uint32_t i = 0; e == hipSuccess && i < cfg.maxInFlight; ++i)
  {
    Slot* slot = new Slot;
    s->slots.push_back(slot);
    e = hipStreamCreateWithFlags(&slot->stream, hipStreamNonBlocking);
    if(e == hipSuccess)
      e = hipEventCreateWithFlags(&slot->done, hipEventDisableTiming);
    if(e == hipSuccess)
      e = hipHostMalloc((void**)&slot->errHost, sizeof(uint32_t), hipHostMallocDefault);
    if(e == hipSuccess)
      *slot->errHost = 0;
    if(e == hipSuccess)
      e = slot->runner.reserve(stageReserve);
    if(cfg.batch.mode == NVPRO_PYRAMID_BATCH_MODE_QUEUE)
      slot->runner.setErrorHost(slot->errHost);  // rounds mode has no waits, so no error flags
  }
  if(e == hipSuccess)
    e = warmUp(s);
  if(const char* t = getenv("NVPRO_PYRAMID_SERVER_TEST_FAIL_LAUNCHES"))
    s->testFailLaunches = (uint32_t)strtoul(t, nullptr, 10);
  if(e != hipSuccess)
  {
    destroyServer(s);
    return hostFail(e, "%s: %s", who, hipGetErrorString(e));
  }
  s->arena.reset(cfg.budgetBytes);
  s->stats.budgetBytes = cfg.budgetBytes;
  *server              = s;
  hostClearError();
  return 0;
}

int nvproPyramidServerSubmit(NvproPyramidServer* server, const NvproPyramidRequest* request, uint64_t* ticket)
{
  const char* who = "nvproPyramidServerSubmit";
  if(!server || !request || !ticket)
    return hostFail(hipErrorInvalidValue, "%s: null argument", who);
  const NvproPyramidRequest& q = *request;
  if(q.format != NVPRO_PYRAMID_FORMAT_SRGBA8 && q.format != NVPRO_PYRAMID_FORMAT_RGBA16F)
    return hostFail(hipErrorInvalidValue, "%s: format %u not served", who, q.format);
  if(q.width == 0 || q.height == 0)
    return hostFail(hipErrorInvalidValue, "%s: zero image size %ux%u", who, q.width, q.height);
  const uint32_t full = nvproPyramidFullLevelCount(q.width, q.height);
  if(full > NVPRO_PYRAMID_MAX_LEVELS || q.mipLevels > full)
    return hostFail(hipErrorInvalidValue, "%s: %ux%u with %u levels not supported", who, q.width, q.height,
                    q.mipLevels);
  const uint64_t bytes = chainBytes(q);
  std::lock_guard<std::mutex> lock(server->mtx);
  if(((bytes + 255u) & ~uint64_t(255u)) > server->cfg.budgetBytes)
    return hostFail(hipErrorOutOfMemory, "%s: chain of %llu bytes exceeds the %llu-byte budget", who,
                    (unsigned long long)bytes, (unsigned long long)server->cfg.budgetBytes);
  Request r{};
  r.req       = q;
  r.ticket    = server->nextTicket++;
  r.state     = State::Waiting;
  r.bytes     = bytes;
  r.submitted = Clock::now();
  server->requests.emplace(r.ticket, r);
  server->waiting.push_back(r.ticket);
  server->stats.submitted++;
  *ticket = r.ticket;
  return 0;
}

int nvproPyramidServerPoll(NvproPyramidServer* server, NvproPyramidCompletion* completions, uint32_t capacity,
                           uint32_t* count)
{
  if(!server || !count || (capacity && !completions))
    return hostFail(hipErrorInvalidValue, "nvproPyramidServerPoll: null argument");
  std::lock_guard<std::mutex> lock(server->mtx);
  *count                           = 0;
  const Clock::time_point now      = Clock::now();
  if(int e = retire(server, now))
    return e;
  admit(server, now);
  if(int e = launch(server))
    return e;
  while(*count < capacity && !server->completions.empty())
  {
    completions[(*count)++] = server->completions.front();
    server->completions.pop_front();
  }
  uint32_t running = 0;
  for(Slot* slot : server->slots)
    running += slot->busy ? 1u : 0u;
  server->stats.inFlightBatches = running;
  server->stats.waiting         = (uint32_t)server->waiting.size();
  return 0;
}

int nvproPyramidServerRelease(NvproPyramidServer* server, uint64_t ticket)
{
  if(!server)
    return hostFail(hipErrorInvalidValue, "nvproPyramidServerRelease: null server");
  std::lock_guard<std::mutex> lock(server->mtx);
  auto it = server->requests.find(ticket);
  if(it == server->requests.end() || it->second.state != State::Done)
    return hostFail(hipErrorInvalidValue, "nvproPyramidServerRelease: ticket %llu is not a completed request",
                    (unsigned long long)ticket);
  server->arena.release(it->second.offset, it->second.bytes);
  server->requests.erase(it);
  server->stats.released++;
  server->stats.residentBytes = server->arena.used();
  return 0;
}

int nvproPyramidServerGetStats(NvproPyramidServer* server, NvproPyramidServerStats* stats)
{
  if(!server || !stats)
    return hostFail(hipErrorInvalidValue, "nvproPyramidServerGetStats: null argument");
  std::lock_guard<std::mutex> lock(server->mtx);
  *stats = server->stats;
  return 0;
}

int nvproPyramidServerDestroy(NvproPyramidServer* server)
{
  if(server)
    destroyServer(server);
  return 0;
}

}  // extern "C"
