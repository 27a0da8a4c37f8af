// Host-side checks of the library and the oracle, built with AddressSanitizer
// and UndefinedBehaviorSanitizer (Makefile beside this file) and run without a
// GPU: every call here either is host-only (layout, planner, batch work-list
// planning, TGA files, the oracle) or fails its argument checks before any
// device call.  Any out-of-bounds access, use after free, leak or undefined
// operation aborts the run with the sanitizer's report; a failed expectation
// prints its line and exits nonzero.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <string>
#include <vector>

#include "batch_internal.hpp"
#include "nvpro_pyramid_hip.h"
#include "orc.h"

using namespace nvpro_pyramid;

static int gFailures = 0;
static uint64_t gChecks = 0;
#define EXPECT(cond)                                                      \
  do                                                                      \
  {                                                                       \
    ++gChecks;                                                            \
    if(!(cond))                                                           \
    {                                                                     \
      fprintf(stderr, "%s:%d: expected %s\n", __FILE__, __LINE__, #cond); \
      if(++gFailures > 20)                                                \
        exit(1);                                                          \
    }                                                                     \
  } while(0)

static bool lastErrorHas(const char* what)
{
  return strstr(nvproPyramidGetLastError(), what) != nullptr;
}

// ---------------------------------------------------------------------------
// Layout and level counts against the oracle (mipmap_storage.hpp:53-76).
static void checkLayout()
{
  std::vector<std::pair<uint32_t, uint32_t>> sizes = {{1, 1},     {2, 1},     {1, 65535}, {4095, 1},
                                                      {4096, 4096}, {16384, 16384}, {65535, 65535}, {3, 65535}};
  for(uint32_t w = 1; w <= 70; ++w)
    for(uint32_t h = 1; h <= 70; h += 3)
      sizes.push_back({w, h});
  for(auto [w, h] : sizes)
  {
    const uint32_t full = nvproPyramidFullLevelCount(w, h);
    EXPECT(full == orc_full_levels(w, h));
    uint64_t texelOffsets[NVPRO_PYRAMID_MAX_LEVELS];
    const uint64_t texels = orc_packed_offsets(w, h, full, texelOffsets);
    for(uint32_t bpt : {4u, 8u, 16u})
    {
      // Exactly `full` entries on the heap, so a write past the end is caught.
      std::vector<uint64_t> offsets(full);
      uint64_t total = 0;
      EXPECT(nvproPyramidPackedLayout(w, h, 0, bpt, offsets.data(), &total) == 0);
      EXPECT(total == texels * bpt);
      for(uint32_t l = 0; l < full; ++l)
        EXPECT(offsets[l] == texelOffsets[l] * bpt);
      EXPECT(nvproPyramidPackedLayout(w, h, full, bpt, nullptr, nullptr) == 0);
      EXPECT(nvproPyramidPackedLayout(w, h, full + 1u, bpt, nullptr, &total) != 0);
    }
  }
  EXPECT(nvproPyramidBytesPerTexel(NVPRO_PYRAMID_FORMAT_SRGBA8) == 4u);
  EXPECT(nvproPyramidBytesPerTexel(NVPRO_PYRAMID_FORMAT_RGBA16F) == 8u);
  EXPECT(nvproPyramidBytesPerTexel(NVPRO_PYRAMID_FORMAT_RGBA32F) == 16u);
  EXPECT(nvproPyramidBytesPerTexel(7u) == 0u);
}

// ---------------------------------------------------------------------------
// Planner against the oracle's restatement (nvpro_pyramid_dispatch.hpp:109-292).
static uint32_t fast26(NvproPyramidCmd* cmd, const NvproPyramidState* st)
{
  return nvproPyramidFastDispatcherEx(cmd, st, 2, 6);
}
static uint32_t fast43(NvproPyramidCmd* cmd, const NvproPyramidState* st)
{
  return nvproPyramidFastDispatcherEx(cmd, st, 4, 3);
}

static void comparePlan(uint32_t w, uint32_t h, uint32_t mip, bool fast, uint32_t div, uint32_t maxLevels,
                        nvpro_pyramid_dispatcher_t fastDispatcher)
{
  const NvproPyramidPipelines pipes{NVPRO_PYRAMID_FORMAT_SRGBA8, fast ? 1u : 0u, 0u, 0u};
  orc_step want[64];
  const int n = orc_plan(w, h, mip, fast ? 1 : 0, div, maxLevels, want, 64);
  EXPECT(n >= 0);
  uint32_t count = 0;
  // Capacity 1 first: only one step may be written, the count is still exact.
  NvproPyramidStep* one = new NvproPyramidStep[1];
  EXPECT(nvproPyramidPlan(&pipes, w, h, mip, nullptr, fastDispatcher, one, 1, &count) == 0);
  EXPECT((int)count == n);
  delete[] one;
  std::vector<NvproPyramidStep> got(count ? count : 1u);
  EXPECT(nvproPyramidPlan(&pipes, w, h, mip, nullptr, fastDispatcher, got.data(), count, &count) == 0);
  EXPECT((int)count == n);
  for(uint32_t s = 0; s < count && (int)s < n; ++s)
  {
    EXPECT(got[s].pipeline == want[s].pipeline);
    EXPECT(got[s].inputLevel == want[s].inputLevel);
    EXPECT(got[s].levelCount == want[s].levelCount);
    EXPECT(got[s].pushConstant == want[s].pushConstant);
    EXPECT(got[s].groupCountX == want[s].groupCountX);
    EXPECT(got[s].inputWidth == want[s].inputWidth);
    EXPECT(got[s].inputHeight == want[s].inputHeight);
  }
}

static void checkPlanner()
{
  for(uint32_t w = 1; w <= 130; ++w)
    for(uint32_t h = 1; h <= 130; ++h)
      for(bool fast : {false, true})
        comparePlan(w, h, 0, fast, 4, 6, nullptr);
  for(auto [w, h] : std::vector<std::pair<uint32_t, uint32_t>>{
          {4096, 4096}, {4094, 4094}, {1920, 1080}, {4095, 2047}, {2047, 4097}, {16384, 16384}, {65535, 3}, {1, 65535}})
  {
    for(uint32_t mip = 0; mip <= nvproPyramidFullLevelCount(w, h); ++mip)
      comparePlan(w, h, mip, true, 4, 6, nullptr);
    comparePlan(w, h, 0, true, 2, 6, fast26);
    comparePlan(w, h, 0, true, 4, 3, fast43);
  }
}

// Dispatchers that break the reference's contract (its asserts,
// nvpro_pyramid_dispatch.hpp:169, :172) make the plan fail, not overrun.
static uint32_t generalNothing(NvproPyramidCmd*, const NvproPyramidState*) { return 0; }
static uint32_t generalGreedy(NvproPyramidCmd*, const NvproPyramidState* st) { return st->remainingLevels + 1u; }
static uint32_t generalPastEnd(NvproPyramidCmd* cmd, const NvproPyramidState* st)
{
  nvproPyramidCmdDispatch(cmd, NVPRO_PYRAMID_PIPELINE_GENERAL, (st->currentLevel + 5u) << 5 | 2u, 1);
  return 2;
}
static uint32_t fastTooDeep(NvproPyramidCmd* cmd, const NvproPyramidState* st)
{
  nvproPyramidCmdDispatch(cmd, NVPRO_PYRAMID_PIPELINE_FAST, st->currentLevel << 5 | 3u, 1);
  return 3;
}
static uint32_t generalNull(NvproPyramidCmd*, const NvproPyramidState* st)
{
  return std::min(2u, st->remainingLevels);
}

static void checkDispatcherContract()
{
  const NvproPyramidPipelines slow{NVPRO_PYRAMID_FORMAT_SRGBA8, 0u, 0u, 0u}, fast{NVPRO_PYRAMID_FORMAT_SRGBA8, 1u, 0u, 0u};
  NvproPyramidStep steps[16];
  uint32_t n = 0;
  EXPECT(nvproPyramidPlan(&slow, 100, 60, 0, generalNothing, nullptr, steps, 16, &n) != 0);
  EXPECT(lastErrorHas("filled no level"));
  EXPECT(nvproPyramidPlan(&slow, 100, 60, 0, generalGreedy, nullptr, steps, 16, &n) != 0);
  EXPECT(lastErrorHas("claimed"));
  EXPECT(nvproPyramidPlan(&slow, 64, 64, 0, generalPastEnd, nullptr, steps, 16, &n) != 0);
  EXPECT(lastErrorHas("outside"));
  EXPECT(nvproPyramidPlan(&fast, 12, 12, 0, nullptr, fastTooDeep, steps, 16, &n) != 0);
  EXPECT(lastErrorHas("divisible by 8"));
  EXPECT(nvproPyramidPlan(&slow, 256, 256, 0, generalNull, nullptr, steps, 16, &n) == 0);
  EXPECT(n == 0);
  EXPECT(nvproPyramidPlan(&slow, 0, 5, 0, nullptr, nullptr, steps, 16, &n) != 0);
  EXPECT(nvproPyramidPlan(&slow, 65536, 4, 0, nullptr, nullptr, steps, 16, &n) != 0);
}

// ---------------------------------------------------------------------------
// Dispatch, batch and generator argument checks (no device call is reached).
static void checkArgumentErrors()
{
  const NvproPyramidPipelines pipes{NVPRO_PYRAMID_FORMAT_SRGBA8, 1u, 0u, 0u};
  void* fake = (void*)(uintptr_t)0x10000;
  NvproPyramidImage img{nullptr, 64, 64, 0, 0, nullptr, nullptr};
  EXPECT(nvproPyramidDispatch(nullptr, &pipes, &img, 0) != 0);
  img.data = (void*)(uintptr_t)0x10002;
  EXPECT(nvproPyramidDispatch(nullptr, &pipes, &img, 0) != 0);
  img.data = fake;
  img.levelCount = 9;
  EXPECT(nvproPyramidDispatch(nullptr, &pipes, &img, 0) != 0);
  img.levelCount = 3;
  EXPECT(nvproPyramidDispatch(nullptr, &pipes, &img, 5) != 0);
  img.levelCount = 7;
  const uint32_t pitches[7] = {64 * 4 - 4, 0, 0, 0, 0, 0, 0};
  img.rowPitches = pitches;
  EXPECT(nvproPyramidDispatch(nullptr, &pipes, &img, 0) != 0);
  img.rowPitches = nullptr;
  const uint64_t offsets[7] = {0, 6, 100, 200, 300, 400, 500};
  img.levelOffsets = offsets;
  EXPECT(nvproPyramidDispatch(nullptr, &pipes, &img, 0) != 0);
  img.levelOffsets = nullptr;
  EXPECT(nvproPyramidDispatch(nullptr, &pipes, &img, 1) == 0);  // mipLevels 1 fills nothing
  img.width = img.height = 1;
  img.levelCount = 0;
  EXPECT(nvproPyramidDispatch(nullptr, &pipes, &img, 0) == 0);  // a 1x1 image has nothing to fill

  NvproPyramidImage two[2] = {{fake, 64, 64, 0, 0, nullptr, nullptr}, {(void*)(uintptr_t)0x20000, 33, 7, 0, 0, nullptr, nullptr}};
  NvproPyramidBatchConfig cfg;
  EXPECT(nvproPyramidBatchDefaultConfig(&cfg) == 0);
  EXPECT(cfg.mode == NVPRO_PYRAMID_BATCH_MODE_QUEUE && cfg.itemBytes == 65536u && cfg.tailTexels == 128u);
  NvproPyramidBatch* batch = nullptr;
  cfg.tailTexels = 129;
  EXPECT(nvproPyramidBatchCreate(&pipes, two, nullptr, 2, &cfg, &batch) != 0 && lastErrorHas("tailTexels"));
  nvproPyramidBatchDefaultConfig(&cfg);
  cfg.mode = 7;
  EXPECT(nvproPyramidBatchCreate(&pipes, two, nullptr, 2, &cfg, &batch) != 0 && lastErrorHas("unknown mode"));
  EXPECT(nvproPyramidBatchCreate(&pipes, two, nullptr, 0, nullptr, &batch) != 0 && lastErrorHas("no images"));
  const uint32_t mips[2] = {0, 7};
  EXPECT(nvproPyramidBatchCreate(&pipes, two, mips, 2, nullptr, &batch) != 0 && lastErrorHas("exceeds"));
  EXPECT(batch == nullptr);
  nvproPyramidBatchDefaultConfig(&cfg);  // queue mode: the no-alloc entry point takes rounds only
  EXPECT(nvproPyramidBatchDispatchNoAlloc(nullptr, &pipes, two, nullptr, 2, &cfg) != 0 && lastErrorHas("ROUNDS"));
  NvproPyramidImage odd{(void*)(uintptr_t)0x10004, 8, 8, 0, 0, nullptr, nullptr};
  EXPECT(nvproPyramidSynthesize(nullptr, NVPRO_PYRAMID_FORMAT_RGBA32F, &two[0], 1) != 0);
  EXPECT(nvproPyramidSynthesize(nullptr, NVPRO_PYRAMID_FORMAT_RGBA16F, &odd, 1) != 0);
}

// ---------------------------------------------------------------------------
// Batch work lists: dependency order, item coverage and the no-alloc size.
static void checkBatchPlans()
{
  std::mt19937 rng(7);
  BatchPlan plan;  // reused from plan to plan, as the serving loop does
  for(int trial = 0; trial < 60; ++trial)
  {
    const uint32_t count = 1u + rng() % 96u;
    std::vector<NvproPyramidImage> images(count);
    std::vector<uint32_t> mips(count);
    for(uint32_t i = 0; i < count; ++i)
    {
      const uint32_t w = 1u + rng() % (trial % 3 ? 700u : 4100u), h = 1u + rng() % 700u;
      images[i]        = NvproPyramidImage{(void*)(uintptr_t)(0x100000u * (i + 1u)), w, h, 0, 0, nullptr, nullptr};
      mips[i]          = rng() % 4u == 0u ? 1u + rng() % nvproPyramidFullLevelCount(w, h) : 0u;
    }
    NvproPyramidBatchConfig cfg;
    defaultBatchConfig(&cfg);
    cfg.mode       = trial % 2 ? NVPRO_PYRAMID_BATCH_MODE_ROUNDS : NVPRO_PYRAMID_BATCH_MODE_QUEUE;
    cfg.itemBytes  = trial % 4 == 3 ? 4096u : 65536u;
    cfg.tailTexels = trial % 5 == 4 ? 0u : 128u;
    const uint32_t format = trial % 3 == 0 ? NVPRO_PYRAMID_FORMAT_RGBA16F : NVPRO_PYRAMID_FORMAT_SRGBA8;
    const NvproPyramidPipelines pipes{format, trial % 7 ? 1u : 0u, 0u, 0u};
    EXPECT(buildBatchPlan(&pipes, images.data(), mips.data(), count, cfg, plan, "host checks") == 0);
    EXPECT(plan.imgs.size() == count);
    EXPECT(!plan.roundEnd.empty() || plan.items.empty());
    EXPECT(plan.roundEnd.empty() || plan.roundEnd.back() == plan.items.size());
    // Items of each segment cover groups [0, n) in order; remember where each
    // segment's items are.
    std::vector<uint32_t> nextGroup(plan.segs.size(), 0u), firstItem(plan.segs.size(), UINT32_MAX),
        lastItem(plan.segs.size(), 0u), itemCount(plan.segs.size(), 0u);
    std::vector<uint32_t> roundOf(plan.items.size(), 0u);
    for(uint32_t r = 0, i = 0; r < plan.roundEnd.size(); ++r)
      for(; i < plan.roundEnd[r]; ++i)
        roundOf[i] = r;
    for(uint32_t i = 0; i < plan.items.size(); ++i)
    {
      const QueueItem& it = plan.items[i];
      EXPECT(it.seg < plan.segs.size());
      if(it.seg >= plan.segs.size())
        continue;
      EXPECT(it.firstGroup == nextGroup[it.seg] && it.groupCount > 0u);
      nextGroup[it.seg] += it.groupCount;
      firstItem[it.seg] = std::min(firstItem[it.seg], i);
      lastItem[it.seg]  = i;
      itemCount[it.seg]++;
    }
    // Every wait is on a slot that earlier items (in earlier rounds) complete,
    // exactly depCount times: persistent workgroups taking items in list order
    // can then never wait on work that no one has started.
    std::vector<int64_t> producer(plan.numSlots, -1);
    for(uint32_t s = 0; s < plan.segs.size(); ++s)
    {
      const QueueSeg& q = plan.segs[s];
      EXPECT(q.image < count);
      if(q.doneSlot != 0xffffffffu)
      {
        EXPECT(q.doneSlot < plan.numSlots && producer[q.doneSlot] < 0);
        producer[q.doneSlot] = s;
      }
      if(q.kind == 9u)  // a tail: inputLevel indexes the tail plans
      {
        EXPECT(q.inputLevel < plan.tails.size());
        if(q.inputLevel < plan.tails.size())
        {
          EXPECT(plan.tails[q.inputLevel].image == q.image);
          EXPECT(plan.tails[q.inputLevel].count >= 1u && plan.tails[q.inputLevel].count <= NVPRO_PYRAMID_MAX_LEVELS);
        }
      }
    }
    for(uint32_t s = 0; s < plan.segs.size(); ++s)
    {
      const QueueSeg& q = plan.segs[s];
      if(q.depSlot == 0xffffffffu)
        continue;
      EXPECT(q.depSlot < plan.numSlots);
      if(q.depSlot >= plan.numSlots || producer[q.depSlot] < 0)
        continue;
      const uint32_t p = (uint32_t)producer[q.depSlot];
      EXPECT(plan.segs[p].image == q.image);
      EXPECT(q.depCount == itemCount[p]);
      EXPECT(lastItem[p] < firstItem[s]);
      EXPECT(roundOf[lastItem[p]] < roundOf[firstItem[s]]);
    }
    // Algorithmic bytes: each image's level 0 read and its filled levels written once.
    const uint32_t bpt = nvproPyramidBytesPerTexel(format);
    uint64_t want      = 0;
    for(uint32_t i = 0; i < count; ++i)
    {
      const uint32_t full = nvproPyramidFullLevelCount(images[i].width, images[i].height);
      const uint32_t last = mips[i] ? mips[i] : full;
      if(last <= 1u)
        continue;
      for(uint32_t l = 0; l < last; ++l)
      {
        uint32_t lw, lh;
        orc_level_dims(images[i].width, images[i].height, l, &lw, &lh);
        want += (uint64_t)lw * lh * bpt;
      }
    }
    EXPECT(plan.algorithmicBytes == want);
    const size_t bytes = BatchRunner::roundsWorkListBytes(plan);
    EXPECT(bytes >= plan.imgs.size() * sizeof(ImageDesc) + plan.items.size() * sizeof(QueueItem));
  }
}

// ---------------------------------------------------------------------------
// TGA files (mipmap_storage.hpp:441-479): round trips, hand-made run-length
// and 24-bit files, then truncated and mutated files that must be rejected or
// read without any access outside the file or the caller's buffer.
static std::string tmpDir()
{
  char tmpl[] = "/tmp/nvpro_pyramid_hostXXXXXX";
  const char* d = mkdtemp(tmpl);
  EXPECT(d != nullptr);
  return d ? d : "/tmp";
}

static void writeFile(const std::string& name, const std::vector<uint8_t>& bytes)
{
  FILE* f = fopen(name.c_str(), "wb");
  EXPECT(f != nullptr);
  if(!f)
    return;
  EXPECT(fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size());
  fclose(f);
}

static std::vector<uint8_t> tgaHeader(uint8_t type, uint32_t w, uint32_t h, uint8_t bpp, uint8_t desc)
{
  std::vector<uint8_t> hdr(18, 0);
  hdr[2]  = type;
  hdr[12] = (uint8_t)w;
  hdr[13] = (uint8_t)(w >> 8);
  hdr[14] = (uint8_t)h;
  hdr[15] = (uint8_t)(h >> 8);
  hdr[16] = bpp;
  hdr[17] = desc;
  return hdr;
}

static void checkTga()
{
  const std::string dir = tmpDir();
  std::mt19937 rng(3);
  for(auto [w, h] : std::vector<std::pair<uint32_t, uint32_t>>{{1, 1}, {3, 5}, {64, 1}, {300, 17}, {129, 130}})
  {
    uint64_t total = 0;
    nvproPyramidPackedLayout(w, h, 0, 4, nullptr, &total);
    std::vector<uint8_t> chain(total);
    for(auto& b : chain)
      b = (uint8_t)rng();
    const std::string base = dir + "/chain.tga";
    EXPECT(nvproPyramidWriteTga(chain.data(), w, h, 0, base.c_str()) == 0);
    const uint32_t full = nvproPyramidFullLevelCount(w, h);
    uint64_t offsets[NVPRO_PYRAMID_MAX_LEVELS];
    nvproPyramidPackedLayout(w, h, full, 4, offsets, nullptr);
    for(uint32_t l = 0; l < full; ++l)
    {
      const std::string name = l ? dir + "/chain." + std::to_string(l) + ".tga" : base;
      uint32_t lw = 0, lh = 0;
      EXPECT(nvproPyramidReadTga(name.c_str(), nullptr, 0, &lw, &lh) == 0);
      std::vector<uint8_t> back((size_t)lw * lh * 4u);
      EXPECT(nvproPyramidReadTga(name.c_str(), back.data(), back.size(), &lw, &lh) == 0);
      EXPECT(memcmp(back.data(), chain.data() + offsets[l], back.size()) == 0);
      if(back.size() > 4u)
        EXPECT(nvproPyramidReadTga(name.c_str(), back.data(), back.size() - 4u, &lw, &lh) != 0);
    }
  }
  // Run-length encoded 32-bit, bottom-up, 5x3: a run packet, a raw packet, a run.
  std::vector<uint8_t> rle = tgaHeader(10, 5, 3, 32, 0x08);
  const uint8_t a[4] = {10, 20, 30, 40}, b[4] = {50, 60, 70, 80};
  rle.push_back(0x80 | 6);  // 7 texels of a
  rle.insert(rle.end(), a, a + 4);
  rle.push_back(1);  // 2 raw texels: b, a
  rle.insert(rle.end(), b, b + 4);
  rle.insert(rle.end(), a, a + 4);
  rle.push_back(0x80 | 5);  // 6 texels of b
  rle.insert(rle.end(), b, b + 4);
  writeFile(dir + "/rle.tga", rle);
  std::vector<uint8_t> got(5 * 3 * 4);
  uint32_t w = 0, h = 0;
  EXPECT(nvproPyramidReadTga((dir + "/rle.tga").c_str(), got.data(), got.size(), &w, &h) == 0);
  EXPECT(w == 5 && h == 3);
  // File order: texels 0..6 a, 7 b, 8 a, 9..14 b, rows from the bottom.
  for(uint32_t i = 0; i < 15; ++i)
  {
    const uint32_t y = 2u - i / 5u, x = i % 5u;
    const uint8_t* want = (i < 7 || i == 8) ? a : b;
    const uint8_t* t    = got.data() + (y * 5u + x) * 4u;
    EXPECT(t[0] == want[2] && t[1] == want[1] && t[2] == want[0] && t[3] == want[3]);
  }
  // 24-bit uncompressed, top-down: alpha reads as 255.
  std::vector<uint8_t> rgb = tgaHeader(2, 2, 2, 24, 0x20);
  for(int i = 0; i < 12; ++i)
    rgb.push_back((uint8_t)(i + 1));
  writeFile(dir + "/rgb.tga", rgb);
  std::vector<uint8_t> rgba(16);
  EXPECT(nvproPyramidReadTga((dir + "/rgb.tga").c_str(), rgba.data(), rgba.size(), &w, &h) == 0);
  EXPECT(rgba[0] == 3 && rgba[1] == 2 && rgba[2] == 1 && rgba[3] == 255 && rgba[15] == 255);
  // Errors.
  EXPECT(nvproPyramidReadTga((dir + "/missing.tga").c_str(), rgba.data(), rgba.size(), &w, &h) != 0);
  std::vector<uint8_t> paletted = tgaHeader(1, 2, 2, 8, 0);
  paletted[1] = 1;
  writeFile(dir + "/pal.tga", paletted);
  EXPECT(nvproPyramidReadTga((dir + "/pal.tga").c_str(), rgba.data(), rgba.size(), &w, &h) != 0);
  EXPECT(nvproPyramidWriteTga(rgba.data(), 0, 4, 0, (dir + "/x.tga").c_str()) != 0);
  EXPECT(nvproPyramidWriteTga(rgba.data(), 4, 4, 4, (dir + "/x.tga").c_str()) != 0);
  // Every truncation of both files, then random byte mutations (the reader
  // gets a buffer of exactly the size it asks for).
  for(const std::vector<uint8_t>* file : {&rle, &rgb})
  {
    for(size_t len = 0; len < file->size(); ++len)
    {
      writeFile(dir + "/cut.tga", std::vector<uint8_t>(file->begin(), file->begin() + (long)len));
      std::vector<uint8_t> out(64);
      EXPECT(nvproPyramidReadTga((dir + "/cut.tga").c_str(), out.data(), out.size(), &w, &h) != 0);
    }
    for(int trial = 0; trial < 1500; ++trial)
    {
      std::vector<uint8_t> m = *file;
      for(int k = 1 + (int)(rng() % 4u); k > 0; --k)
        m[18u + rng() % (m.size() - 18u)] = (uint8_t)rng();
      if(trial % 3 == 0)
        m[16] = (uint8_t)(rng() % 2u ? 24 : 32);
      if(trial % 5 == 0)
        m[0] = (uint8_t)(rng() % 8u);  // image id length
      if(trial % 7 == 0)
        m.resize(m.size() - rng() % 8u);
      writeFile(dir + "/mut.tga", m);
      uint32_t mw = 0, mh = 0;
      if(nvproPyramidReadTga((dir + "/mut.tga").c_str(), nullptr, 0, &mw, &mh) != 0)
        continue;
      std::vector<uint8_t> out((size_t)mw * mh * 4u);
      (void)nvproPyramidReadTga((dir + "/mut.tga").c_str(), out.data(), out.size(), &mw, &mh);
    }
  }
  const std::string rmCmd = "rm -rf '" + dir + "'";
  EXPECT(system(rmCmd.c_str()) == 0);
}

// ---------------------------------------------------------------------------
// The oracle's emulators on exact-size heap chains: odd and one-texel-wide
// sizes reach every 3-tap and edge branch.
static void checkOracle()
{
  std::mt19937 rng(11);
  for(auto [w, h] : std::vector<std::pair<uint32_t, uint32_t>>{
          {1, 1}, {1, 7}, {7, 1}, {5, 3}, {33, 17}, {64, 64}, {127, 65}, {3, 129}, {256, 5}})
  {
    const uint32_t full = orc_full_levels(w, h);
    uint64_t offsets[NVPRO_PYRAMID_MAX_LEVELS];
    const uint64_t texels = orc_packed_offsets(w, h, full, offsets);
    for(bool fast : {false, true})
    {
      orc_step steps[64];
      const int n = orc_plan(w, h, 0, fast ? 1 : 0, 4, 6, steps, 64);
      EXPECT(n >= 0);
      std::vector<uint8_t> c32(texels * 4u), c64, cpu;
      for(uint64_t i = 0; i < (uint64_t)w * h * 4u; ++i)
        c32[i] = (uint8_t)rng();
      orc_premultiply_srgba8(c32.data(), (uint64_t)w * h, 1);
      c64 = cpu = c32;
      EXPECT(orc_pyramid_srgba8_f32(c32.data(), w, h, full, steps, (uint32_t)n) == 0);
      EXPECT(orc_pyramid_srgba8_f64(c64.data(), w, h, full, steps, (uint32_t)n) == 0);
      orc_reference_cpu_srgba8_mt(cpu.data(), w, h, 3);
      uint32_t wx, wy, wl, wc;
      uint64_t diff = 0, total = 0;
      EXPECT(orc_compare_srgba8(c32.data(), c64.data(), w, h, full, &wx, &wy, &wl, &wc, &diff, &total) <= 1u);
      EXPECT(orc_compare_srgba8(c32.data(), cpu.data(), w, h, full, &wx, &wy, &wl, &wc, &diff, &total) <= 5u);
      std::vector<float> f32(texels * 4u);
      for(uint64_t i = 0; i < (uint64_t)w * h * 4u; ++i)
        f32[i] = (float)(rng() % 10000u) * 1e-3f;
      EXPECT(orc_pyramid_rgba32f_f32(f32.data(), w, h, full, steps, (uint32_t)n) == 0);
      std::vector<double> lin(texels * 4u);
      for(uint64_t i = 0; i < (uint64_t)w * h * 4u; ++i)
        lin[i] = f32[i];
      orc_pyramid_linear_f64(lin.data(), w, h, full);
      for(uint64_t i = (uint64_t)w * h * 4u; i < texels * 4u; ++i)
        EXPECT(fabs(lin[i] - f32[i]) <= 1e-4 * (1.0 + fabs(lin[i])));
      for(int packed : {0, 1})
      {
        std::vector<uint16_t> half(texels * 4u);
        for(uint64_t i = 0; i < (uint64_t)w * h * 4u; ++i)
          half[i] = orc_half_from_double((double)(rng() % 4096u) / 64.0);
        EXPECT(orc_pyramid_rgba16f(half.data(), w, h, full, steps, (uint32_t)n, packed) == 0);
      }
    }
  }
}

int main()
{
  checkLayout();
  checkPlanner();
  checkDispatcherContract();
  checkArgumentErrors();
  checkBatchPlans();
  checkTga();
  checkOracle();
  if(gFailures)
  {
    fprintf(stderr, "%d host check(s) failed\n", gFailures);
    return 1;
  }
  printf("all host checks passed (%llu expectations)\n", (unsigned long long)gChecks);
  return 0;
}
