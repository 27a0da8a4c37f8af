// C++ host of the C-ABI: the counterpart of the reference's minimal_app
// (minimal_app/minimal_mipmaps.cpp:59-241) and of the per-image benchmark loop
// of demo_app/mipmaps_app.cpp:672-722, over HIP streams.
//
//   nvpro_pyramid_demo minimal [W H [seed]]
//       Synthesises a W x H sRGBA8 level 0 on the device, fills the chain with
//       nvproPyramidDispatch, downloads it and prints each level's size and an
//       FNV-1a checksum.
//   nvpro_pyramid_demo batch [--count N] [--seed S] [--format srgba8|rgba16f]
//                            [--packed] [--reps R] [--autotune] [--no-graph]
//       N textures of the mixed-size workload (sides log-uniform on [64, 4096],
//       40% odd; the same sizes as vk_compute_mipmaps_amd/synth.py mixed_sizes).
//       Times a loop of plain per-texture dispatches (eager, and captured into
//       one hipGraph like a recorded command buffer) against one batched launch,
//       checks that both write identical bytes, prints one JSON object.
//   nvpro_pyramid_demo serve [--rate R] [--count N] [--warmup W] [--seed S] [--hold-ms H]
//                            [--latency-csv PATH] [--budget-gb B] [--in-flight K] [--max-batch M]
//                            [--packed] [--queue | --rounds]
//       Serving loop: W + N requests arrive as a Poisson stream of R per second;
//       the first W warm the GPU's clocks and the server up and are left out
//       of every figure but the failure count;
//       sizes from the mixed-size workload, 70% sRGBA8 / 30% RGBA16F; level 0
//       is the seeded synthetic image, written on the device at launch.  The
//       server admits against a B GB budget and batches what arrived; each
//       completed chain stays resident for H ms before the client releases it.
//       Prints chains/s, p50/p99 latency and batching statistics as JSON;
//       --latency-csv also writes every request's arrival, size and latency.
//   nvpro_pyramid_demo sizes N S
//       Prints the workload's N texture sizes for seed S (no device use).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <unordered_map>
#include <vector>

#include "nvpro_pyramid_hip.h"

#define CHECK_HIP(x)                                                                                                 \
  do                                                                                                                 \
  {                                                                                                                  \
    hipError_t e_ = (x);                                                                                             \
    if(e_ != hipSuccess)                                                                                             \
    {                                                                                                                \
      fprintf(stderr, "%s:%d: %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_));                            \
      exit(2);                                                                                                       \
    }                                                                                                                \
  } while(0)
#define CHECK_PYR(x)                                                                                                 \
  do                                                                                                                 \
  {                                                                                                                  \
    if(int e_ = (x))                                                                                                 \
    {                                                                                                                \
      fprintf(stderr, "%s:%d: %s: error %d: %s\n", __FILE__, __LINE__, #x, e_, nvproPyramidGetLastError());       \
      exit(3);                                                                                                       \
    }                                                                                                                \
  } while(0)

static uint64_t fnv1a(const uint8_t* p, size_t n)
{
  uint64_t h = 1469598103934665603ull;
  for(size_t i = 0; i < n; ++i)
    h = (h ^ p[i]) * 1099511628211ull;
  return h;
}

// synth.py's counter-based hash and mixed_sizes (log-uniform sides, odd fraction).
static uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static uint64_t hash4(uint64_t seed, uint64_t s, uint64_t i, uint64_t j)
{
  return mix64(mix64(mix64(seed ^ (s << 56)) ^ i) ^ j);
}
static std::vector<std::pair<uint32_t, uint32_t>> mixedSizes(uint32_t count, uint64_t seed)
{
  const double lo = 64.0, hi = 4096.0, odd = 0.4;
  std::vector<std::pair<uint32_t, uint32_t>> out;
  for(uint32_t k = 0; k < count; ++k)
  {
    const double u1 = (double)(hash4(seed, 0, k, k) >> 11) * 0x1p-53;
    const double u2 = (double)(hash4(seed, 1, k, k) >> 11) * 0x1p-53;
    const double u3 = (double)(hash4(seed, 2, k, k) >> 11) * 0x1p-53;
    const double w  = exp(log(lo) + u1 * (log(hi) - log(lo)));
    const double h  = exp(log(lo) + u2 * (log(hi) - log(lo)));
    long a = lround(w), b = lround(h);  // Python round(): half to even; ties do not occur here
    a = std::max(64l, std::min(4096l, a));
    b = std::max(64l, std::min(4096l, b));
    if(u3 < odd)
    {
      a = std::min(a | 1l, 4095l);
      b = std::min(b | 1l, 4095l);
    }
    else
    {
      a &= ~1l;
      b &= ~1l;
    }
    out.push_back({(uint32_t)a, (uint32_t)b});
  }
  return out;
}

static int minimal(int argc, char** argv)
{
  const uint32_t w    = argc > 2 ? (uint32_t)atoi(argv[2]) : 1024u;
  const uint32_t h    = argc > 3 ? (uint32_t)atoi(argv[3]) : 1024u;
  const uint64_t seed = argc > 4 ? (uint64_t)atoll(argv[4]) : 1u;
  const uint32_t levels = nvproPyramidFullLevelCount(w, h);
  std::vector<uint64_t> offsets(levels);
  uint64_t bytes = 0;
  CHECK_PYR(nvproPyramidPackedLayout(w, h, 0, 4, offsets.data(), &bytes));
  void* d = nullptr;
  hipStream_t s;
  CHECK_HIP(hipStreamCreate(&s));
  CHECK_HIP(hipMalloc(&d, bytes));
  NvproPyramidPipelines pipes{NVPRO_PYRAMID_FORMAT_SRGBA8, 1u, NVPRO_PYRAMID_ARITH_FP32, 0u};
  NvproPyramidImage image{d, w, h, 0, 0, nullptr, nullptr};
  CHECK_PYR(nvproPyramidSynthesize(s, NVPRO_PYRAMID_FORMAT_SRGBA8, &image, seed));
  CHECK_PYR(nvproPyramidDispatch(s, &pipes, &image, 0));
  std::vector<uint8_t> chain(bytes);
  CHECK_HIP(hipMemcpyAsync(chain.data(), d, bytes, hipMemcpyDeviceToHost, s));
  CHECK_HIP(hipStreamSynchronize(s));
  for(uint32_t l = 0; l < levels; ++l)
  {
    const uint64_t end = l + 1 < levels ? offsets[l + 1] : bytes;
    printf("level %2u  %5u x %-5u  fnv1a %016llx\n", l, std::max(1u, w >> l), std::max(1u, h >> l),
           (unsigned long long)fnv1a(chain.data() + offsets[l], end - offsets[l]));
  }
  CHECK_HIP(hipFree(d));
  CHECK_HIP(hipStreamDestroy(s));
  return 0;
}

struct Args
{
  uint32_t count = 512, reps = 20, format = NVPRO_PYRAMID_FORMAT_SRGBA8, arith = NVPRO_PYRAMID_ARITH_FP32;
  uint64_t seed = 8;
  bool autotune = false, graph = true;
};

static float elapsedMs(hipEvent_t a, hipEvent_t b)
{
  float ms = 0;
  CHECK_HIP(hipEventElapsedTime(&ms, a, b));
  return ms;
}

static double percentile(std::vector<double> v, double p)
{
  std::sort(v.begin(), v.end());
  const double idx = p * (double)(v.size() - 1);
  const size_t i   = (size_t)floor(idx);
  const size_t j   = std::min(v.size() - 1, i + 1);
  return v[i] + (v[j] - v[i]) * (idx - (double)i);
}

static int batch(const Args& a)
{
  const uint32_t bpt = nvproPyramidBytesPerTexel(a.format);
  const auto sizes   = mixedSizes(a.count, a.seed);
  std::vector<NvproPyramidImage> images(a.count);
  std::vector<uint64_t> bytes(a.count);
  uint64_t algorithmic = 0, padded = 0;  // base read once + every level written once = the packed chain sizes
  for(uint32_t k = 0; k < a.count; ++k)
  {
    CHECK_PYR(nvproPyramidPackedLayout(sizes[k].first, sizes[k].second, 0, bpt, nullptr, &bytes[k]));
    algorithmic += bytes[k];
    padded += (bytes[k] + 255u) & ~uint64_t(255u);
  }
  uint8_t* pool = nullptr;
  CHECK_HIP(hipMalloc(&pool, padded));
  hipStream_t s;
  CHECK_HIP(hipStreamCreate(&s));
  uint64_t off = 0;
  for(uint32_t k = 0; k < a.count; ++k)
  {
    images[k] = NvproPyramidImage{pool + off, sizes[k].first, sizes[k].second, 0, 0, nullptr, nullptr};
    off += (bytes[k] + 255u) & ~uint64_t(255u);
    CHECK_PYR(nvproPyramidSynthesize(s, a.format, &images[k], a.seed * 1000u + k));
  }
  NvproPyramidPipelines pipes{a.format, 1u, a.arith, 0u};
  hipEvent_t e0, e1;
  CHECK_HIP(hipEventCreate(&e0));
  CHECK_HIP(hipEventCreate(&e1));
  auto loopOnce = [&]() {
    for(uint32_t k = 0; k < a.count; ++k)
      CHECK_PYR(nvproPyramidDispatch(s, &pipes, &images[k], 0));
  };
  // Reference bytes: the loop of plain dispatches.
  loopOnce();
  CHECK_HIP(hipStreamSynchronize(s));
  std::vector<uint8_t> want(padded), got(padded);
  CHECK_HIP(hipMemcpy(want.data(), pool, padded, hipMemcpyDeviceToHost));

  // Loop, eager.
  std::vector<double> loopMs;
  for(uint32_t r = 0; r < a.reps; ++r)
  {
    CHECK_HIP(hipEventRecord(e0, s));
    loopOnce();
    CHECK_HIP(hipEventRecord(e1, s));
    CHECK_HIP(hipEventSynchronize(e1));
    loopMs.push_back(elapsedMs(e0, e1));
  }
  // Loop captured into one graph (every launch recorded ahead, as a command buffer).
  hipGraph_t g          = nullptr;
  hipGraphExec_t ge     = nullptr;
  std::vector<double> graphMs;
  if(a.graph)
  {
    CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    loopOnce();
    CHECK_HIP(hipStreamEndCapture(s, &g));
    CHECK_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    CHECK_HIP(hipGraphLaunch(ge, s));
    for(uint32_t r = 0; r < a.reps; ++r)
    {
      CHECK_HIP(hipEventRecord(e0, s));
      CHECK_HIP(hipGraphLaunch(ge, s));
      CHECK_HIP(hipEventRecord(e1, s));
      CHECK_HIP(hipEventSynchronize(e1));
      graphMs.push_back(elapsedMs(e0, e1));
    }
  }
  else
    graphMs = loopMs;  // not measured without graphs

  // Batched.
  NvproPyramidBatchConfig cfg;
  CHECK_PYR(nvproPyramidBatchDefaultConfig(&cfg));
  float tunedUs = 0;
  if(a.autotune)
    CHECK_PYR(nvproPyramidBatchAutotune(&pipes, images.data(), nullptr, a.count, s, 3, &cfg, &tunedUs));
  cfg.useGraph = a.graph ? 1u : 0u;
  NvproPyramidBatch* b = nullptr;
  CHECK_PYR(nvproPyramidBatchCreate(&pipes, images.data(), nullptr, a.count, &cfg, &b));
  CHECK_HIP(hipMemsetAsync(pool, 0, padded, s));  // outputs and inputs cleared ...
  for(uint32_t k = 0; k < a.count; ++k)           // ... and the inputs written again
    CHECK_PYR(nvproPyramidSynthesize(s, a.format, &images[k], a.seed * 1000u + k));
  CHECK_PYR(nvproPyramidBatchLaunch(b, s));
  CHECK_HIP(hipStreamSynchronize(s));
  CHECK_HIP(hipMemcpy(got.data(), pool, padded, hipMemcpyDeviceToHost));
  const bool identical = memcmp(got.data(), want.data(), padded) == 0;
  uint32_t flags = 0;
  CHECK_PYR(nvproPyramidBatchReadErrors(b, &flags));
  std::vector<double> batchMs;
  for(uint32_t r = 0; r < a.reps; ++r)
  {
    CHECK_HIP(hipEventRecord(e0, s));
    CHECK_PYR(nvproPyramidBatchLaunch(b, s));
    CHECK_HIP(hipEventRecord(e1, s));
    CHECK_HIP(hipEventSynchronize(e1));
    batchMs.push_back(elapsedMs(e0, e1));
  }
  NvproPyramidBatchInfo info;
  CHECK_PYR(nvproPyramidBatchGetInfo(b, &info));

  const double loopMed = percentile(loopMs, 0.5), graphMed = percentile(graphMs, 0.5);
  const double bMed = percentile(batchMs, 0.5), bP99 = percentile(batchMs, 0.99), bMin = percentile(batchMs, 0.0);
  const double bestLoop = std::min(loopMed, graphMed);
  printf("{\"images\": %u, \"format\": \"%s\", \"arithmetic\": \"%s\", \"seed\": %llu, \"algorithmic_bytes\": %llu, "
         "\"reps\": %u,\n",
         a.count, a.format == NVPRO_PYRAMID_FORMAT_SRGBA8 ? "srgba8" : "rgba16f",
         a.arith == NVPRO_PYRAMID_ARITH_PACKED_FP16 ? "packed_fp16" : "fp32", (unsigned long long)a.seed,
         (unsigned long long)algorithmic, a.reps);
  printf(" \"loop_eager\": {\"ms_p50\": %.4f, \"ms_p99\": %.4f, \"chains_per_s\": %.1f},\n", loopMed,
         percentile(loopMs, 0.99), a.count / (loopMed * 1e-3));
  printf(" \"loop_graph\": {\"ms_p50\": %.4f, \"ms_p99\": %.4f, \"chains_per_s\": %.1f},\n", graphMed,
         percentile(graphMs, 0.99), a.count / (graphMed * 1e-3));
  printf(" \"batch\": {\"mode\": \"%s\", \"workgroups_per_cu\": %u, \"item_bytes\": %u, \"tail_texels\": %u, "
         "\"graph\": %s, \"autotuned\": %s, \"launches\": %u, \"items\": %u, \"rounds\": %u,\n",
         info.config.mode == NVPRO_PYRAMID_BATCH_MODE_QUEUE ? "queue" : "rounds", info.config.workgroupsPerCU,
         info.config.itemBytes, info.config.tailTexels, info.config.useGraph ? "true" : "false",
         a.autotune ? "true" : "false", info.launchCount, info.itemCount, info.roundCount);
  printf("   \"ms_min\": %.4f, \"ms_p50\": %.4f, \"ms_p99\": %.4f, \"chains_per_s\": %.1f, \"gb_per_s\": %.1f},\n", bMin,
         bMed, bP99, a.count / (bMed * 1e-3), algorithmic / (bMed * 1e-3) / 1e9);
  printf(" \"speedup_vs_best_loop\": %.3f, \"identical_to_loop\": %s, \"error_flags\": %u}\n", bestLoop / bMed,
         identical ? "true" : "false", flags);
  CHECK_PYR(nvproPyramidBatchDestroy(b));
  if(ge)
    CHECK_HIP(hipGraphExecDestroy(ge));
  if(g)
    CHECK_HIP(hipGraphDestroy(g));
  CHECK_HIP(hipFree(pool));
  CHECK_HIP(hipStreamDestroy(s));
  return identical && flags == 0 ? 0 : 1;
}

struct ServeArgs
{
  double rate = 20000.0, holdMs = 20.0, budgetGB = 80.0;
  uint32_t count = 20000, inFlight = 4, maxBatch = 0;
  uint64_t seed = 9;
  bool packed = false;
  int mode    = -1;  // batch mode; -1: the server's default
  uint32_t warmup        = 0;        // requests before the measured ones: served, not counted
  const char* latencyCsv = nullptr;  // per request: index, arrival s, width, height, format, latency us
};

static int serve(const ServeArgs& a)
{
  using Clock        = std::chrono::steady_clock;
  const uint32_t total = a.warmup + a.count;  // requests 0 .. warmup-1 are the warm-up
  const auto sizes     = mixedSizes(total, a.seed);
  std::vector<double> arrival(total);
  std::vector<uint32_t> format(total);
  double t = 0;
  for(uint32_t k = 0; k < total; ++k)
  {
    const double u = ((double)(hash4(a.seed, 3, k, k) >> 11) + 0.5) * 0x1p-53;
    t += -log(u) / a.rate;  // exponential inter-arrival times
    arrival[k] = t;
    const double f = (double)(hash4(a.seed, 4, k, k) >> 11) * 0x1p-53;
    format[k]      = f < 0.7 ? NVPRO_PYRAMID_FORMAT_SRGBA8 : NVPRO_PYRAMID_FORMAT_RGBA16F;
  }
  NvproPyramidServerConfig cfg;
  CHECK_PYR(nvproPyramidServerDefaultConfig(&cfg));
  cfg.budgetBytes   = (uint64_t)(a.budgetGB * 1e9);
  cfg.maxInFlight   = a.inFlight;
  if(a.maxBatch)
    cfg.maxBatchImages = a.maxBatch;
  cfg.arithmetic16f = a.packed ? NVPRO_PYRAMID_ARITH_PACKED_FP16 : NVPRO_PYRAMID_ARITH_FP32;
  if(a.mode >= 0)
    cfg.batch.mode = (uint32_t)a.mode;
  NvproPyramidServer* srv = nullptr;
  CHECK_PYR(nvproPyramidServerCreate(&cfg, &srv));
  std::vector<NvproPyramidCompletion> done(4096);
  std::vector<double> latency, wait;
  std::vector<std::pair<double, uint64_t>> holding;  // (release time, ticket)
  std::unordered_map<uint64_t, uint32_t> indexOf;      // ticket -> request index
  std::vector<std::pair<uint32_t, double>> order;      // (request index, latency us)
  uint64_t bytes = 0;
  uint32_t next = 0, finished = 0, measured = 0, failed = 0;
  const Clock::time_point t0 = Clock::now();
  const double start         = arrival[a.warmup];  // the measured window opens with the first measured arrival
  double lastDone            = start;
  double lastIter = 0, gapMax = 0, pollMax = 0;  // host loop in the window: longest gap between iterations, longest poll
  uint32_t slowPolls = 0;                        // polls over 1 ms
  double pollSum = 0, submitSum = 0;             // host time in the window inside Poll and Submit, seconds
  NvproPyramidServerStats atStart{};             // server counters when the window opens
  bool opened = a.warmup == 0;
  while(finished < total)
  {
    const double now = std::chrono::duration<double>(Clock::now() - t0).count();
    while(next < total && arrival[next] <= now)
    {
      if(next == a.warmup && a.warmup != 0)
      {
        CHECK_PYR(nvproPyramidServerGetStats(srv, &atStart));
        opened = true;
      }
      NvproPyramidRequest r{format[next], sizes[next].first, sizes[next].second, 0, a.seed * 100000u + next,
                            nullptr};
      uint64_t ticket;
      const Clock::time_point s0 = Clock::now();
      CHECK_PYR(nvproPyramidServerSubmit(srv, &r, &ticket));
      if(opened)
        submitSum += std::chrono::duration<double>(Clock::now() - s0).count();
      indexOf[ticket] = next;
      ++next;
    }
    uint32_t n = 0;
    const Clock::time_point p0 = Clock::now();
    CHECK_PYR(nvproPyramidServerPoll(srv, done.data(), (uint32_t)done.size(), &n));
    const double pollUs = std::chrono::duration<double, std::micro>(Clock::now() - p0).count();
    if(opened)
    {
      pollMax = std::max(pollMax, pollUs);
      slowPolls += pollUs > 1000.0 ? 1u : 0u;
      pollSum += pollUs * 1e-6;
      if(lastIter >= start)
        gapMax = std::max(gapMax, (now - lastIter) * 1e6);
    }
    lastIter = now;
    for(uint32_t i = 0; i < n; ++i)
    {
      const auto it     = indexOf.find(done[i].ticket);
      const uint32_t ix = it->second;
      indexOf.erase(it);
      failed += done[i].status != 0 ? 1u : 0u;
      holding.push_back({now + a.holdMs * 1e-3, done[i].ticket});
      ++finished;
      if(ix < a.warmup)
        continue;
      latency.push_back(done[i].latencyUs);
      if(a.latencyCsv)
        order.push_back({ix, done[i].latencyUs});
      wait.push_back(done[i].waitUs);
      bytes += done[i].bytes;
      ++measured;
      lastDone = now;
    }
    size_t keep = 0;
    for(size_t i = 0; i < holding.size(); ++i)
    {
      if(holding[i].first <= now)
        CHECK_PYR(nvproPyramidServerRelease(srv, holding[i].second));
      else
        holding[keep++] = holding[i];
    }
    holding.resize(keep);
  }
  NvproPyramidServerStats st;
  CHECK_PYR(nvproPyramidServerGetStats(srv, &st));
  for(const auto& h : holding)
    CHECK_PYR(nvproPyramidServerRelease(srv, h.second));
  CHECK_PYR(nvproPyramidServerDestroy(srv));
  const double span         = lastDone - start;
  const uint64_t batches    = st.batches - atStart.batches;
  const uint64_t batchedImg = st.batchedImages - atStart.batchedImages;
  if(a.latencyCsv)
  {
    FILE* f = fopen(a.latencyCsv, "w");
    if(!f)
      return 4;
    fprintf(f, "index,arrival_s,width,height,format,latency_us\n");
    std::sort(order.begin(), order.end());
    for(const auto& o : order)
      fprintf(f, "%u,%.6f,%u,%u,%u,%.1f\n", o.first, arrival[o.first], sizes[o.first].first, sizes[o.first].second,
              format[o.first], o.second);
    fclose(f);
  }
  printf("{\"requests\": %u, \"warmup_requests\": %u, \"rate_per_s\": %.1f, "
         "\"mix\": \"70%% sRGBA8 / 30%% RGBA16F (%s)\", "
         "\"budget_bytes\": %llu, \"in_flight\": %u, \"hold_ms\": %.1f, \"batch_mode\": \"%s\",\n",
         a.count, a.warmup, a.rate, a.packed ? "packed_fp16" : "fp32", (unsigned long long)cfg.budgetBytes, a.inFlight,
         a.holdMs, cfg.batch.mode == NVPRO_PYRAMID_BATCH_MODE_QUEUE ? "queue" : "rounds");
  printf(" \"seconds\": %.4f, \"chains_per_s\": %.1f, \"GBps\": %.1f, \"failed\": %u,\n", span,
         measured / span, bytes / span / 1e9, failed);
  printf(" \"latency_us\": {\"p50\": %.1f, \"p90\": %.1f, \"p99\": %.1f, \"max\": %.1f}, "
         "\"admission_wait_us\": {\"p50\": %.1f, \"p99\": %.1f},\n",
         percentile(latency, 0.5), percentile(latency, 0.9), percentile(latency, 0.99), percentile(latency, 1.0),
         percentile(wait, 0.5), percentile(wait, 0.99));
  // The second half of the measured completions alone.
  const std::vector<double> late(latency.begin() + latency.size() / 2, latency.end());
  printf(" \"latency_us_second_half\": {\"p50\": %.1f, \"p99\": %.1f, \"max\": %.1f},\n", percentile(late, 0.5),
         percentile(late, 0.99), percentile(late, 1.0));
  printf(" \"host_loop_us\": {\"max_gap\": %.1f, \"max_poll\": %.1f, \"polls_over_1ms\": %u, "
         "\"poll_busy_frac\": %.3f, \"submit_busy_frac\": %.3f},\n",
         gapMax, pollMax, slowPolls, pollSum / span, submitSum / span);
  printf(" \"batches\": %llu, \"mean_batch_images\": %.2f, \"batch_image_cap\": %u, "
         "\"peak_resident_bytes\": %llu}\n",
         (unsigned long long)batches, batches ? (double)batchedImg / batches : 0.0,
         cfg.maxBatchImages, (unsigned long long)st.peakResidentBytes);
  return failed == 0 ? 0 : 1;
}

int main(int argc, char** argv)
{
  if(argc > 1 && !strcmp(argv[1], "serve"))
  {
    ServeArgs a;
    for(int i = 2; i < argc; ++i)
    {
      const std::string o = argv[i];
      if(o == "--rate" && i + 1 < argc)
        a.rate = atof(argv[++i]);
      else if(o == "--count" && i + 1 < argc)
        a.count = (uint32_t)atoi(argv[++i]);
      else if(o == "--warmup" && i + 1 < argc)
        a.warmup = (uint32_t)atoi(argv[++i]);
      else if(o == "--seed" && i + 1 < argc)
        a.seed = (uint64_t)atoll(argv[++i]);
      else if(o == "--hold-ms" && i + 1 < argc)
        a.holdMs = atof(argv[++i]);
      else if(o == "--budget-gb" && i + 1 < argc)
        a.budgetGB = atof(argv[++i]);
      else if(o == "--in-flight" && i + 1 < argc)
        a.inFlight = (uint32_t)atoi(argv[++i]);
      else if(o == "--max-batch" && i + 1 < argc)
        a.maxBatch = (uint32_t)atoi(argv[++i]);
      else if(o == "--packed")
        a.packed = true;
      else if(o == "--queue")
        a.mode = NVPRO_PYRAMID_BATCH_MODE_QUEUE;
      else if(o == "--rounds")
        a.mode = NVPRO_PYRAMID_BATCH_MODE_ROUNDS;
      else if(o == "--latency-csv" && i + 1 < argc)
        a.latencyCsv = argv[++i];
      else
      {
        fprintf(stderr, "unknown option %s\n", o.c_str());
        return 4;
      }
    }
    if(a.count == 0 || !(a.rate > 0) || a.warmup > (1u << 24) || a.count > (1u << 24))
      return 4;
    return serve(a);
  }
  if(argc > 3 && !strcmp(argv[1], "sizes"))  // host only: the workload's texture sizes
  {
    for(const auto& wh : mixedSizes((uint32_t)atoi(argv[2]), (uint64_t)atoll(argv[3])))
      printf("%u %u\n", wh.first, wh.second);
    return 0;
  }
  if(argc > 1 && !strcmp(argv[1], "minimal"))
    return minimal(argc, argv);
  if(argc > 1 && !strcmp(argv[1], "batch"))
  {
    Args a;
    for(int i = 2; i < argc; ++i)
    {
      const std::string o = argv[i];
      if(o == "--count" && i + 1 < argc)
        a.count = (uint32_t)atoi(argv[++i]);
      else if(o == "--seed" && i + 1 < argc)
        a.seed = (uint64_t)atoll(argv[++i]);
      else if(o == "--reps" && i + 1 < argc)
        a.reps = (uint32_t)atoi(argv[++i]);
      else if(o == "--format" && i + 1 < argc)
        a.format = !strcmp(argv[++i], "rgba16f") ? NVPRO_PYRAMID_FORMAT_RGBA16F : NVPRO_PYRAMID_FORMAT_SRGBA8;
      else if(o == "--packed")
        a.arith = NVPRO_PYRAMID_ARITH_PACKED_FP16;
      else if(o == "--autotune")
        a.autotune = true;
      else if(o == "--no-graph")
        a.graph = false;
      else
      {
        fprintf(stderr, "unknown option %s\n", o.c_str());
        return 4;
      }
    }
    if(a.count == 0 || a.reps == 0)
      return 4;
    return batch(a);
  }
  fprintf(stderr, "usage: %s minimal [W H [seed]] | batch [--count N] [--seed S] [--format srgba8|rgba16f] "
                  "[--packed] [--reps R] [--autotune]\n",
          argv[0]);
  return 4;
}
