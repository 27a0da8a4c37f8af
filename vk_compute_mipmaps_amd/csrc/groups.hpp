// Work-group bodies of the three pipelines, shared by the per-step kernels and
// the batched kernel.  Each body does one unit of a step ("group"): a tile group
// of a fast step, 256 outputs or a 16x8 (8x8) tile of a general step, or a
// whole tail plan.  They take the LDS decode/encode tables (`tab`, any table type the
// functor's load/store accept) and the LDS scratch they need as pointers, so a
// kernel can overlay the scratch of different bodies.  Every thread of the
// 256-thread workgroup calls a body (a tail: any workgroup size); bodies with
// barriers say so.
#pragma once

#include "kernel_common.hpp"

namespace nvpro_pyramid {

constexpr int kGroupBS = 256;  // threads per workgroup, every pipeline

// ---------------------------------------------------------------------------
// Fast pipeline (nvpro_pyramid.glsl:209-534).
//
// One step fills COUNT (1..6) levels from an input level whose edges are
// divisible by 2^COUNT.  The input is cut into 2^COUNT x 2^COUNT tiles.  Each
// lane owns a 4x4 input block (2x2 when COUNT is 1), loads it with one 16-byte
// load per row and reduces it in registers to levels 1 and 2, so a group of 256
// threads covers 4096 input texels: one 64x64 tile for COUNT 6, four 32x32
// tiles for COUNT 5, and so on.  Lanes are numbered in Morton order within a
// tile, so the level-2 values of a 2x2 are those of four consecutive threads.
// Levels 3..COUNT are reduced by one "duty" wave per group (rotating over the
// workgroup's waves with `iter`): the level-2 values go through a 4 KiB LDS
// tile, duty lane k reduces threads 4k..4k+3 to level 3 (all 64 lanes busy), and
// the levels above are cross-lane reductions inside the duty wave (DPP row
// shifts, then ds_bpermute).  Per level the 2x2 summation order is the
// reference's (fastOrder).
// ---------------------------------------------------------------------------
template <int COUNT>
struct FastGeom
{
  static constexpr int LB           = COUNT >= 2 ? 2 : 1;                // lane block edge, log2
  static constexpr int TEAM         = COUNT >= 2 ? 2 * (COUNT - 2) : 0;  // threads per tile, log2 (<= 8)
  static constexpr int TILES_PER_WG = kGroupBS >> TEAM;
  static constexpr bool UPPER       = COUNT >= 3;                        // levels 3.. exist
};

// Groups of a fast step over a width x height input level.
__host__ __device__ inline uint32_t fastGroupCount(uint32_t width, uint32_t height, uint32_t count)
{
  const uint32_t tiles = (width >> count) * (height >> count);
  const uint32_t team  = count >= 2u ? 2u * (count - 2u) : 0u;
  const uint32_t perWG = (uint32_t)kGroupBS >> team;
  return (tiles + perWG - 1u) / perWG;
}

// The input block of one lane for group g of a fast step: its 2^LB rows, all
// loads issued before any arithmetic.
template <class F, int COUNT>
struct FastWindow
{
  typename F::Raw raw[1 << FastGeom<COUNT>::LB][1 << FastGeom<COUNT>::LB];
};

template <class F, int COUNT>
__device__ __forceinline__ void fastLoad(const ImageDesc& img, uint32_t inLevel, uint32_t tilesX, uint32_t numTiles,
                                         uint32_t g, FastWindow<F, COUNT>& w)
{
  using G           = FastGeom<COUNT>;
  constexpr int LB  = G::LB;
  constexpr int BLK = 1 << LB;
  const uint32_t tile = g * G::TILES_PER_WG + (threadIdx.x >> G::TEAM);
  if(tile >= numTiles)
    return;
  const uint32_t inTile = threadIdx.x & ((1u << G::TEAM) - 1u);
  const uint32_t px = compactEven(inTile), py = compactEven(inTile >> 1);
  const uint32_t ty = tile / tilesX, tx = tile - ty * tilesX;
  const uint32_t x0 = (tx << COUNT) + (px << LB), y0 = (ty << COUNT) + (py << LB);
  const LevelDesc& in = img.lv[inLevel];
  const uint8_t* p    = img.base + in.offset + (uint64_t)y0 * in.pitch + (uint64_t)x0 * F::kBytes;
#pragma unroll
  for(int r = 0; r < BLK; ++r)
    loadRow<F, BLK>(p + (uint64_t)r * in.pitch, w.raw[r]);
}

// Group g of a fast step of `img` from level inLevel (tilesX x tilesY tiles in
// all, numTiles = their product), from the lane blocks fastLoad put in `w`.
// `level2`: kGroupBS values of LDS scratch (unused when COUNT < 3).  Contains
// barriers when COUNT >= 3.
template <class F, int COUNT, bool CO = false, class TT>
__device__ __forceinline__ void fastCompute(const TT& tab, typename F::Val* level2, const ImageDesc& img,
                                            uint32_t inLevel, uint32_t tilesX, uint32_t numTiles, uint32_t g,
                                            uint32_t iter, const FastWindow<F, COUNT>& w)
{
  using G            = FastGeom<COUNT>;
  using Raw          = typename F::Raw;
  using V            = typename F::Val;
  constexpr int LB   = G::LB;
  constexpr int BLK  = 1 << LB;
  constexpr bool PRE = F::kPrescale;

  const uint32_t lane   = threadIdx.x & 63u;
  const uint32_t wave   = threadIdx.x >> 6;
  const uint32_t inTile = threadIdx.x & ((1u << G::TEAM) - 1u);
  const uint32_t px     = compactEven(inTile);
  const uint32_t py     = compactEven(inTile >> 1);

  // Levels 1 and 2 of this thread's block.
  const uint32_t tile = g * G::TILES_PER_WG + (threadIdx.x >> G::TEAM);
  V v{};
  if(tile < numTiles)  // whole tiles
  {
    const uint32_t ty = tile / tilesX, tx = tile - ty * tilesX;
    const uint32_t x0 = (tx << COUNT) + (px << LB), y0 = (ty << COUNT) + (py << LB);
    const auto& raw = w.raw;

    // Level 1: LOAD_REDUCE4 as four loads (premultiplied by 0.25 when exact).
    V l1[BLK / 2][BLK / 2];
#pragma unroll
    for(int i = 0; i < BLK / 2; ++i)
    {
      Raw row[BLK / 2];
#pragma unroll
      for(int j = 0; j < BLK / 2; ++j)
      {
        l1[i][j] = box<F, PRE>(F::template load<PRE>(tab, raw[2 * i][2 * j]),
                               F::template load<PRE>(tab, raw[2 * i][2 * j + 1]),
                               F::template load<PRE>(tab, raw[2 * i + 1][2 * j]),
                               F::template load<PRE>(tab, raw[2 * i + 1][2 * j + 1]), fastOrder(COUNT, 1u));
        row[j] = F::store(tab, l1[i][j]);
      }
      if constexpr(CO && COUNT == 1)
        storeTexelCoherent<F>(texelPtr(img, inLevel + 1u, x0 >> 1, (y0 >> 1) + i, F::kBytes), row[0]);
      else
        storeRow<F, BLK / 2>(texelPtr(img, inLevel + 1u, x0 >> 1, (y0 >> 1) + i, F::kBytes), row);
    }
    v = l1[0][0];
    if constexpr(LB == 2)
    {
      v = box<F, false>(l1[0][0], l1[0][1], l1[1][0], l1[1][1], fastOrder(COUNT, 2u));
      storeTexelTo<F, CO && COUNT == 2>(texelPtr(img, inLevel + 2u, x0 >> 2, y0 >> 2, F::kBytes), F::store(tab, v));
    }
  }
  if constexpr(G::UPPER)
  {
    __syncthreads();  // the previous duty wave has read level2[]
    level2[threadIdx.x] = v;
    __syncthreads();
    if(wave == (iter & 3u))
    {
      // Duty lane k: level 3 from threads 4k..4k+3 (a Morton 2x2 of level 2).
      constexpr int TEAM3 = 2 * (COUNT - 3);  // duty lanes per tile, log2 (<= 6)
      const uint32_t t3   = g * G::TILES_PER_WG + (lane >> TEAM3);
      const uint32_t k3   = lane & ((1u << TEAM3) - 1u);
      const uint32_t qx   = compactEven(k3), qy = compactEven(k3 >> 1);
      const V* c          = &level2[4u * lane];
      V u                 = box<F, false>(c[0], c[1], c[2], c[3], fastOrder(COUNT, 3u));
      if(t3 < numTiles)
      {
        const uint32_t ty = t3 / tilesX, tx = t3 - ty * tilesX;
        storeTexelTo<F, CO && COUNT == 3>(
            texelPtr(img, inLevel + 3u, (tx << (COUNT - 3)) + qx, (ty << (COUNT - 3)) + qy, F::kBytes),
            F::store(tab, u));
#pragma unroll
        for(int k = 0; k < 3; ++k)
        {
          const int r = 4 + k;
          if(r > COUNT)
            break;
          if(k == 0)
            u = crossLaneBox<F, 1>(u, fastOrder(COUNT, (uint32_t)r));
          else if(k == 1)
            u = crossLaneBox<F, 4>(u, fastOrder(COUNT, (uint32_t)r));
          else
            u = crossLaneBox<F, 16>(u, fastOrder(COUNT, (uint32_t)r));
          if((k3 & ((4u << (2 * k)) - 1u)) == 0u)
          {
            const uint32_t sft = (uint32_t)(r - 3);
            uint8_t* const p   = texelPtr(img, inLevel + (uint32_t)r, (tx << (COUNT - r)) + (qx >> sft),
                                          (ty << (COUNT - r)) + (qy >> sft), F::kBytes);
            if(CO && r == COUNT)
              storeTexelCoherent<F>(p, F::store(tab, u));
            else
              storeTexel<F>(p, F::store(tab, u));
          }
        }
      }
    }
  }
}

// Group g, loads included.  Contains barriers when COUNT >= 3.  CO: the last
// level is stored with storeTexelCoherent (for a continuation in the same
// launch, see arriveLast).
template <class F, int COUNT, bool CO = false, class TT>
__device__ __forceinline__ void fastGroup(const TT& tab, typename F::Val* level2, const ImageDesc& img,
                                          uint32_t inLevel, uint32_t tilesX, uint32_t numTiles, uint32_t g,
                                          uint32_t iter)
{
  FastWindow<F, COUNT> w;
  fastLoad<F, COUNT>(img, inLevel, tilesX, numTiles, g, w);
  fastCompute<F, COUNT, CO>(tab, level2, img, inLevel, tilesX, numTiles, g, iter, w);
}

// ---------------------------------------------------------------------------
// General pipeline (nvpro_pyramid.glsl:536-884): 1 or 2 levels of any size with
// the 2x2 / 2x3 / 3x2 / 3x3 energy-conserving kernel.
// ---------------------------------------------------------------------------

// Geometry of a two-level general group: a TX x TY tile of the second level,
// whose first level (up to (2 TX + 1) x (2 TY + 1) texels, sharedLevel_ of
// nvpro_pyramid.glsl:555) is computed into LDS by threads that each own a
// vertical run of R first-level texels of one column.  Narrow texels take the
// wide tile; 16-byte texels a smaller one to bound registers.
template <uint32_t BYTES>
struct Gen2Geom
{
  static constexpr uint32_t TX = BYTES <= 8u ? 16u : 8u;
  static constexpr uint32_t TY = 8u;
  static constexpr uint32_t R  = BYTES <= 8u ? 3u : 2u;
  static constexpr uint32_t MW = 2u * TX + 1u;  // LDS row stride of the first level
  static constexpr uint32_t MH = 2u * TY + 1u;
  static_assert(MW * ((MH + R - 1u) / R) <= (uint32_t)kGroupBS, "one pass of the workgroup");
};

// Groups of a one-level general step (a thread per output) and of a two-level
// one (a Gen2Geom tile of the second level each; *tilesX receives the tile
// columns).
__host__ __device__ inline uint32_t generalGroupCount(uint32_t dstWidth, uint32_t dstHeight, uint32_t count,
                                                      uint32_t bytesPerTexel, uint32_t* tilesX)
{
  if(count == 1u)
  {
    *tilesX = 0;
    return (uint32_t)(((uint64_t)dstWidth * dstHeight + kGroupBS - 1) / kGroupBS);
  }
  const uint32_t tx = bytesPerTexel <= 8u ? Gen2Geom<8>::TX : Gen2Geom<16>::TX;
  const uint32_t ty = bytesPerTexel <= 8u ? Gen2Geom<8>::TY : Gen2Geom<16>::TY;
  *tilesX           = (dstWidth + tx - 1u) / tx;
  return *tilesX * ((dstHeight + ty - 1u) / ty);
}

// One level (glsl:828-842): outputs [256 g, 256 g + 256) of level inLevel+1, in
// row-major order, a thread each.  No barrier.
template <class F, class TT>
__device__ __forceinline__ void generalOneGroup(const TT& tab, const ImageDesc& img, uint32_t inLevel, uint32_t g)
{
  const LevelDesc& src = img.lv[inLevel];
  const LevelDesc& dst = img.lv[inLevel + 1u];
  const int kx = kernelSize(src.width), ky = kernelSize(src.height);
  const uint32_t i = g * kGroupBS + threadIdx.x;
  if(i >= dst.width * dst.height)
    return;
  const uint32_t dy = i / dst.width, dx = i - dy * dst.width;
  const uint8_t* p  = img.base + src.offset + (uint64_t)(2u * dy) * src.pitch + (uint64_t)(2u * dx) * F::kBytes;
  auto at           = [&](int j, int r) {
    return F::template load<false>(tab, loadTexel<F>(p + (uint64_t)r * src.pitch + (uint64_t)j * F::kBytes));
  };
  const typename F::Val v = generalSample<F>(at, kx, ky, axis3(dst.width), axis3(dst.height), dx, dy);
  storeTexel<F>(texelPtr(img, inLevel + 1u, dx, dy, F::kBytes), F::store(tab, v));
}

// LDS values of a two-level general group.
template <class F>
constexpr uint32_t gen2MidTexels()
{
  return Gen2Geom<F::kBytes>::MW * Gen2Geom<F::kBytes>::MH;
}

// Two levels: the Gen2Geom tile `tile` of level inLevel+2 and the first-level
// texels it needs.  First level (fillIntermediateTile_, glsl:732-781): thread t
// owns column t % MW of the first-level tile, rows R (t / MW) .. + R-1.
// gen2Load issues the loads of the thread's whole input window (up to 2R+1
// rows x 3 columns) into a Gen2Window; gen2Compute then decodes each input
// texel once and computes the run top to bottom (reduceStoreSample_'s order:
// vertical per column, then horizontal), storing each value and keeping it in
// `mid` ([y][x], stride MW).  Second level (fillLastTile_, glsl:791-820) from
// LDS, a thread per texel.  Neighbouring tiles recompute their shared
// first-level border texels and store identical values.  (Keeping a second
// window in flight across tiles costs two waves of occupancy per SIMD and
// measured slower.)
template <class F>
struct Gen2Window
{
  typename F::Raw raw[2u * Gen2Geom<F::kBytes>::R + 1u][3];
};

struct Gen2Run
{
  uint32_t gx, gy0, n;  // first-level column, first row, outputs of the run (0: idle thread)
};

template <class F>
__device__ __forceinline__ Gen2Run gen2Run(const ImageDesc& img, uint32_t inLevel, uint32_t tilesX, uint32_t tile)
{
  using G              = Gen2Geom<F::kBytes>;
  const LevelDesc& l1  = img.lv[inLevel + 1u];
  const int kx2 = kernelSize(l1.width), ky2 = kernelSize(l1.height);
  const uint32_t nx = 2u * G::TX + (kx2 == 3 ? 1u : 0u), ny = 2u * G::TY + (ky2 == 3 ? 1u : 0u);
  const uint32_t ty = tile / tilesX, tx = tile - ty * tilesX;
  const uint32_t lx = threadIdx.x % G::MW, ly0 = (threadIdx.x / G::MW) * G::R;
  Gen2Run run{2u * G::TX * tx + lx, 2u * G::TY * ty + ly0, 0u};
  if(lx < nx && ly0 < ny && run.gx < l1.width && run.gy0 < l1.height)
  {
    uint32_t n = G::R;
    n      = ny - ly0 < n ? ny - ly0 : n;
    run.n  = l1.height - run.gy0 < n ? l1.height - run.gy0 : n;
  }
  return run;
}

template <class F>
__device__ __forceinline__ void gen2Load(const ImageDesc& img, uint32_t inLevel, uint32_t tilesX, uint32_t tile,
                                         Gen2Window<F>& w)
{
  constexpr uint32_t NIN = 2u * Gen2Geom<F::kBytes>::R + 1u;
  const Gen2Run run      = gen2Run<F>(img, inLevel, tilesX, tile);
  if(run.n == 0u)
    return;
  const LevelDesc& src = img.lv[inLevel];
  const int kx1 = kernelSize(src.width), ky1 = kernelSize(src.height);
  const uint32_t rows = 2u * (run.n - 1u) + (uint32_t)ky1;  // input rows the run reads
  const uint8_t* p = img.base + src.offset + (uint64_t)(2u * run.gy0) * src.pitch + (uint64_t)(2u * run.gx) * F::kBytes;
#pragma unroll
  for(uint32_t r = 0; r < NIN; ++r)
#pragma unroll
    for(int j = 0; j < 3; ++j)
      if(r < rows && j < kx1)
        w.raw[r][j] = loadTexel<F>(p + (uint64_t)r * src.pitch + (uint64_t)j * F::kBytes);
}

// Contains barriers.
template <class F, class TT>
__device__ __forceinline__ void gen2Compute(const TT& tab, typename F::Val* mid, const ImageDesc& img,
                                            uint32_t inLevel, uint32_t tilesX, uint32_t tile, const Gen2Window<F>& w)
{
  using G              = Gen2Geom<F::kBytes>;
  using V              = typename F::Val;
  const LevelDesc& src = img.lv[inLevel];
  const LevelDesc& l1  = img.lv[inLevel + 1u];
  const LevelDesc& l2  = img.lv[inLevel + 2u];
  const int kx1 = kernelSize(src.width), ky1 = kernelSize(src.height);
  const int kx2 = kernelSize(l1.width), ky2 = kernelSize(l1.height);
  const Gen2Run run = gen2Run<F>(img, inLevel, tilesX, tile);
  if(run.n != 0u)
  {
    const Axis3 ax = axis3(l1.width), ay = axis3(l1.height);
    const uint32_t lx = threadIdx.x % G::MW, ly0 = (threadIdx.x / G::MW) * G::R;
    V top[3];  // decoded input row 2k of output k (the last row of output k-1 when 3-tall)
#pragma unroll
    for(int j = 0; j < 3; ++j)
      if(j < kx1)
        top[j] = F::template load<false>(tab, w.raw[0][j]);
#pragma unroll
    for(uint32_t k = 0; k < G::R; ++k)
    {
      if(k < run.n)
      {
        const Weights3 wy = weights3(ay, run.gy0 + k);
        V h[3];
#pragma unroll
        for(int j = 0; j < 3; ++j)
        {
          if(j < kx1)
          {
            const V a0 = top[j];
            const V a1 = ky1 >= 2 ? F::template load<false>(tab, w.raw[2u * k + 1u][j]) : a0;
            const V a2 = ky1 == 3 ? F::template load<false>(tab, w.raw[2u * k + 2u][j]) : a0;
            h[j]       = verticalTap<F>(wy, ky1, a0, a1, a2);
            if(k + 1u < run.n)  // the next output's first row: shared (3-tall) or the row after (2-tall)
              top[j] = ky1 == 3 ? a2 : F::template load<false>(tab, w.raw[2u * k + 2u][j]);
          }
        }
        const V v = horizontalTap<F>(ax, run.gx, kx1, h[0], h[1], h[2]);
        storeTexel<F>(texelPtr(img, inLevel + 1u, run.gx, run.gy0 + k, F::kBytes), F::store(tab, v));
        mid[(ly0 + k) * G::MW + lx] = v;
      }
    }
  }
  __syncthreads();
  if(threadIdx.x < G::TX * G::TY)
  {
    const uint32_t ty = tile / tilesX, tx = tile - ty * tilesX;
    const uint32_t ox = threadIdx.x % G::TX, oy = threadIdx.x / G::TX;
    const uint32_t x = G::TX * tx + ox, y = G::TY * ty + oy;
    if(x < l2.width && y < l2.height)
    {
      auto at = [&](int j, int r) { return mid[(2u * oy + r) * G::MW + 2u * ox + j]; };
      const V v = generalSample<F>(at, kx2, ky2, axis3(l2.width), axis3(l2.height), x, y);
      storeTexel<F>(texelPtr(img, inLevel + 2u, x, y, F::kBytes), F::store(tab, v));
    }
  }
  __syncthreads();
}

// One tile.  Contains barriers.
template <class F, class TT>
__device__ __forceinline__ void generalTwoGroup(const TT& tab, typename F::Val* mid, const ImageDesc& img,
                                                uint32_t inLevel, uint32_t tilesX, uint32_t tile)
{
  Gen2Window<F> w;
  gen2Load<F>(img, inLevel, tilesX, tile, w);
  gen2Compute<F>(tab, mid, img, inLevel, tilesX, tile, w);
}

// ---------------------------------------------------------------------------
// Tail: every remaining step of a plan in one workgroup.
//
// Once a plan reaches small levels, one kernel per step only buys launch
// latency.  A tail runs the remaining steps, fast and general alike: every level
// is computed by the whole workgroup, stored to memory, and kept in LDS for the
// next level; at a step boundary the value kept is the decode of the stored
// texel, exactly what the next dispatch would load.  The arithmetic per level is
// that of the per-step bodies (fast: box with the reference's order; general:
// reduceStoreSample_), so results do not depend on which kernel ran a step.
// `buf0`, `buf1`: LDS ping-pong of maxTexels values each, where maxTexels is at
// least the first output level's texel count.  BS threads (the workgroup).
// Contains barriers.
// ---------------------------------------------------------------------------
//
// The first level reads memory: its outputs go in chunks of PF passes of the
// workgroup, every load of a chunk issued before its arithmetic, so a chunk
// pays one memory latency instead of one per pass (9 PF texels of registers).
template <class F, int BS, int PF = 1, class TT>
__device__ __forceinline__ void tailRun(const TT& tab, typename F::Val* buf0, typename F::Val* buf1,
                                        const ImageDesc& img, const TailPlan& plan)
{
  using V      = typename F::Val;
  using Raw    = typename F::Raw;
  const V* cur = nullptr;  // the current source level in LDS, null: read it from memory
  V* dst           = buf0;
  for(uint32_t s = 0; s < plan.count; ++s)
  {
    const uint32_t L     = plan.steps[s].inputLevel;
    const uint32_t count = plan.steps[s].levelCount;
    const bool fast      = plan.steps[s].pipeline == NVPRO_PYRAMID_PIPELINE_FAST;
    for(uint32_t r = 1; r <= count; ++r)
    {
      const LevelDesc& sl = img.lv[L + r - 1u];
      const LevelDesc& dl = img.lv[L + r];
      const int kx = kernelSize(sl.width), ky = kernelSize(sl.height);
      const int order     = fastOrder(count, r);
      const uint32_t n    = dl.width * dl.height;
      const Axis3 ax = axis3(dl.width), ay = axis3(dl.height);
      auto finish         = [&](uint32_t i, uint32_t dx, uint32_t dy, V v) {
        const Raw q = F::store(tab, v);
        storeTexel<F>(img.base + dl.offset + (uint64_t)dy * dl.pitch + (uint64_t)dx * F::kBytes, q);
        if(r == count)
          v = F::template load<false>(tab, q);  // the next step reads the stored texel
        dst[i] = v;
      };
      if(!cur)
      {
        for(uint32_t base = 0; base < n; base += PF * BS)
        {
          Raw raw[PF][3][3];
#pragma unroll
          for(int p = 0; p < PF; ++p)
          {
            const uint32_t i = base + (uint32_t)p * BS + threadIdx.x;
            if(i < n)
            {
              const uint32_t dy = i / dl.width, dx = i - dy * dl.width;
              const uint8_t* q  = img.base + sl.offset + (uint64_t)(2u * dy) * sl.pitch + (uint64_t)(2u * dx) * F::kBytes;
#pragma unroll
              for(int k = 0; k < 3; ++k)
#pragma unroll
                for(int j = 0; j < 3; ++j)
                  if(j < kx && k < ky)
                    raw[p][k][j] = loadTexel<F>(q + (uint64_t)k * sl.pitch + (uint64_t)j * F::kBytes);
            }
          }
#pragma unroll
          for(int p = 0; p < PF; ++p)
          {
            const uint32_t i = base + (uint32_t)p * BS + threadIdx.x;
            if(i < n)
            {
              const uint32_t dy = i / dl.width, dx = i - dy * dl.width;
              auto at = [&](int j, int k) -> V { return F::template load<false>(tab, raw[p][k][j]); };
              finish(i, dx, dy, fast ? box<F, false>(at(0, 0), at(1, 0), at(0, 1), at(1, 1), order)
                                     : generalSample<F>(at, kx, ky, ax, ay, dx, dy));
            }
          }
        }
      }
      else
      {
        for(uint32_t i = threadIdx.x; i < n; i += BS)
        {
          const uint32_t dy = i / dl.width, dx = i - dy * dl.width;
          auto at = [&](int j, int k) -> V { return cur[(2u * dy + (uint32_t)k) * sl.width + 2u * dx + (uint32_t)j]; };
          finish(i, dx, dy, fast ? box<F, false>(at(0, 0), at(1, 0), at(0, 1), at(1, 1), order)
                                 : generalSample<F>(at, kx, ky, ax, ay, dx, dy));
        }
      }
      __syncthreads();
      cur = dst;
      dst = dst == buf0 ? buf1 : buf0;
    }
  }
}

// Texel count of level `inputLevel + 1` (the first output level of a step).
__host__ __device__ inline uint64_t firstOutputTexels(const ImageDesc& img, uint32_t inputLevel)
{
  return (uint64_t)img.lv[inputLevel + 1u].width * img.lv[inputLevel + 1u].height;
}

}  // namespace nvpro_pyramid
