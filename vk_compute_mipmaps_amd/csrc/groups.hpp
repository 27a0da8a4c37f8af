// Work-group bodies of the three pipelines, shared by the per-step kernels and
// the batched kernel.  Each body does one unit of a step ("group"): a tile group
// of a fast step, 256 outputs or a 31x7 (31x5) tile of a general step, or a
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
// tiles for COUNT 5, and so on.  The group's tiles are consecutive in tile
// order (side by side along a tile row, except where a row ends), and lanes are
// numbered row-major over the blocks of that strip (STRIP, the per-step
// kernels' COUNT 6 groups), so the lanes of one wave load four texel rows of
// 256 contiguous bytes per instruction instead of the eight 128-byte pieces of
// a Morton order within the tile (the other counts and the batched kernel:
// measured faster there).
// Levels 3..COUNT are reduced by one "duty" wave per group (rotating over the
// workgroup's waves with `iter`): the level-2 values go through a 4 KiB LDS
// strip, duty lane k reads the 2x2 of level 2 under its level-3 texel (all 64
// lanes busy; duty lanes are in Morton order within a tile), and the levels
// above are cross-lane reductions inside the duty wave (DPP row shifts, then
// ds_bpermute).  Per level the 2x2 summation order is the reference's
// (fastOrder).
// ---------------------------------------------------------------------------
template <int COUNT>
struct FastGeom
{
  static constexpr int LB           = COUNT >= 2 ? 2 : 1;                // lane block edge, log2
  static constexpr int TEAM         = COUNT >= 2 ? 2 * (COUNT - 2) : 0;  // threads per tile, log2 (<= 8)
  static constexpr int TILES_PER_WG = kGroupBS >> TEAM;
  static constexpr bool UPPER       = COUNT >= 3;                        // levels 3.. exist
  static constexpr int T            = COUNT >= 2 ? 1 << (COUNT - 2) : 1;  // tile edge in lane blocks
  static constexpr int RW           = TILES_PER_WG * T;                  // strip width in lane blocks
};

// Position of lane t in the group: its tile among the group's tiles and its
// block (px, py) in that tile.  STRIP: row-major over the strip of the group's
// tiles; otherwise Morton order within each tile, tiles one after the other.
// Both layouts compute every value from the same operands in the same order.
// Measured on MI355X (sRGBA8, per step): the strip takes F(0,6) of 16384^2
// from 299 to 291 us and of 8192^2 RGBA16F from 142.5 to 137.9 us, but F(0,4)
// of 3840x2160 from 10.2 to 10.7 us, F(0,5) of 2560x1440 from 6.6 to 6.8 us and
// the one-group F(6,4) of 1024^2 by 0.9 us (one small tile then spans all four
// waves); in the batched kernel its index arithmetic costs registers that the
// many inlined bodies cannot spare.
template <int COUNT, bool STRIP>
__device__ __forceinline__ void fastLanePos(uint32_t t, uint32_t& tileInGroup, uint32_t& px, uint32_t& py)
{
  using G = FastGeom<COUNT>;
  if constexpr(STRIP)
  {
    const uint32_t rx = t & (uint32_t)(G::RW - 1);
    tileInGroup       = rx / (uint32_t)G::T;
    px                = rx & (uint32_t)(G::T - 1);
    py                = t / (uint32_t)G::RW;
  }
  else
  {
    const uint32_t inTile = t & ((1u << G::TEAM) - 1u);
    tileInGroup           = t >> G::TEAM;
    px                    = compactEven(inTile);
    py                    = compactEven(inTile >> 1);
  }
}

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

// NT: load with the non-temporal hint (data read once).
template <class F, int COUNT, bool STRIP, bool NT = false>
__device__ __forceinline__ void fastLoad(const ImageDesc& img, uint32_t inLevel, uint32_t tilesX, uint32_t numTiles,
                                         uint32_t g, FastWindow<F, COUNT>& w)
{
  using G           = FastGeom<COUNT>;
  constexpr int LB  = G::LB;
  constexpr int BLK = 1 << LB;
  uint32_t tig, px, py;
  fastLanePos<COUNT, STRIP>(threadIdx.x, tig, px, py);
  const uint32_t tile = g * G::TILES_PER_WG + tig;
  if(tile >= numTiles)
    return;
  const uint32_t ty = tile / tilesX, tx = tile - ty * tilesX;
  const uint32_t x0 = (tx << COUNT) + (px << LB), y0 = (ty << COUNT) + (py << LB);
  const LevelDesc& in = img.lv[inLevel];
  const uint8_t* p    = img.base + in.offset + (uint64_t)y0 * in.pitch + (uint64_t)x0 * F::kBytes;
  if constexpr(NT)
  {
#pragma unroll
    for(int r = 0; r < BLK; ++r)
      loadRowStream<F, BLK>(p + (uint64_t)r * in.pitch, w.raw[r]);
  }
  else
  {
#pragma unroll
    for(int r = 0; r < BLK; ++r)
      loadRow<F, BLK>(p + (uint64_t)r * in.pitch, w.raw[r]);
  }
}

// Row addressing of a fast step's output levels inLevel+1..inLevel+COUNT.
// HOIST: read from the image descriptor once, before any arithmetic.  The
// descriptor is kernel-argument memory read through the scalar cache; read at
// each level's first store, every read that misses stalls the wave there,
// which is what a one-group successor step meets on levels no other workgroup
// has addressed.  The empty asm makes the values (scalar registers) ready at
// this point, so their loads overlap the input loads in flight.  Without HOIST
// (the batched kernel, whose many inlined bodies cannot spare the registers)
// each store reads its level's descriptor.
template <int COUNT, bool HOIST>
struct FastOut
{
  uint8_t* base[COUNT];
  uint32_t pitch[COUNT];

  __device__ __forceinline__ FastOut(const ImageDesc& img, uint32_t inLevel)
  {
#pragma unroll
    for(int r = 0; r < COUNT; ++r)
    {
      const LevelDesc& d = img.lv[inLevel + 1u + (uint32_t)r];
      base[r]            = img.base + d.offset;
      pitch[r]           = d.pitch;
      __asm__ volatile("" ::"s"(base[r]), "s"(pitch[r]));
    }
  }
  // Texel (x, y) of level inLevel + r, r in 1..COUNT.
  __device__ __forceinline__ uint8_t* at(int r, uint32_t x, uint32_t y, uint32_t bytes) const
  {
    return base[r - 1] + (uint64_t)y * pitch[r - 1] + (uint64_t)x * bytes;
  }
};

template <int COUNT>
struct FastOut<COUNT, false>
{
  const ImageDesc& img;
  uint32_t inLevel;

  __device__ __forceinline__ FastOut(const ImageDesc& i, uint32_t l) : img(i), inLevel(l) {}
  __device__ __forceinline__ uint8_t* at(int r, uint32_t x, uint32_t y, uint32_t bytes) const
  {
    return texelPtr(img, inLevel + (uint32_t)r, x, y, bytes);
  }
};

// Group g of a fast step of `img` from level inLevel (tilesX x tilesY tiles in
// all, numTiles = their product), from the lane blocks fastLoad put in `w`.
// `level2`: kGroupBS values of LDS scratch (unused when COUNT < 3).  Contains
// barriers when COUNT >= 3.
template <class F, int COUNT, bool CO, bool STRIP, bool HOIST, class TT>
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

  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t tig, px, py;
  fastLanePos<COUNT, STRIP>(threadIdx.x, tig, px, py);
  const FastOut<COUNT, HOIST> out(img, inLevel);

  // Levels 1 and 2 of this thread's block.
  const uint32_t tile = g * G::TILES_PER_WG + tig;
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
        storeTexelCoherent<F>(out.at(1, x0 >> 1, (y0 >> 1) + i, F::kBytes), row[0]);
      else
        storeRow<F, BLK / 2>(out.at(1, x0 >> 1, (y0 >> 1) + i, F::kBytes), row);
    }
    v = l1[0][0];
    if constexpr(LB == 2)
    {
      v = box<F, false>(l1[0][0], l1[0][1], l1[1][0], l1[1][1], fastOrder(COUNT, 2u));
      storeTexelTo<F, CO && COUNT == 2>(out.at(2, x0 >> 2, y0 >> 2, F::kBytes), F::store(tab, v));
    }
  }
  if constexpr(G::UPPER)
  {
    __syncthreads();  // the previous duty wave has read level2[]
    level2[threadIdx.x] = v;
    __syncthreads();
    if(wave == (iter & 3u))
    {
      // Duty lane k: level-3 texel (qx, qy) of tile lane >> TEAM3 (Morton
      // order within the tile), from the 2x2 of level 2 under it: threads
      // 4k..4k+3 in Morton order, or two rows of the strip.
      constexpr int TEAM3 = 2 * (COUNT - 3);  // duty lanes per tile, log2 (<= 6)
      const uint32_t t3   = g * G::TILES_PER_WG + (lane >> TEAM3);
      const uint32_t k3   = lane & ((1u << TEAM3) - 1u);
      const uint32_t qx   = compactEven(k3), qy = compactEven(k3 >> 1);
      const uint32_t b  = STRIP ? 2u * qy * (uint32_t)G::RW + (lane >> TEAM3) * (uint32_t)G::T + 2u * qx : 4u * lane;
      const uint32_t dy = STRIP ? (uint32_t)G::RW : 2u;
      V u = box<F, false>(level2[b], level2[b + 1u], level2[b + dy], level2[b + dy + 1u], fastOrder(COUNT, 3u));
      if(t3 < numTiles)
      {
        const uint32_t ty = t3 / tilesX, tx = t3 - ty * tilesX;
        storeTexelTo<F, CO && COUNT == 3>(out.at(3, (tx << (COUNT - 3)) + qx, (ty << (COUNT - 3)) + qy, F::kBytes),
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
            uint8_t* const p   = out.at(r, (tx << (COUNT - r)) + (qx >> sft), (ty << (COUNT - r)) + (qy >> sft),
                                        F::kBytes);
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
// launch, see arriveLast).  STRIP: the lane layout (fastLanePos).  NT: the
// input loads carry the non-temporal hint.  HOIST: output addressing read up
// front (FastOut).
template <class F, int COUNT, bool CO = false, bool STRIP = (COUNT == 6), bool NT = false, bool HOIST = true,
          class TT>
__device__ __forceinline__ void fastGroup(const TT& tab, typename F::Val* level2, const ImageDesc& img,
                                          uint32_t inLevel, uint32_t tilesX, uint32_t numTiles, uint32_t g,
                                          uint32_t iter)
{
  FastWindow<F, COUNT> w;
  fastLoad<F, COUNT, STRIP, NT>(img, inLevel, tilesX, numTiles, g, w);
  fastCompute<F, COUNT, CO, STRIP, HOIST>(tab, level2, img, inLevel, tilesX, numTiles, g, iter, w);
}

// ---------------------------------------------------------------------------
// General pipeline (nvpro_pyramid.glsl:536-884): 1 or 2 levels of any size with
// the 2x2 / 2x3 / 3x2 / 3x3 energy-conserving kernel.
// ---------------------------------------------------------------------------

// Geometry of a two-level general group: a TX x TY tile of the second level,
// whose first level (up to MW x MH texels, sharedLevel_ of
// nvpro_pyramid.glsl:555) is computed into LDS.  The layout follows the wave:
// lane c of every wave owns first-level column c of the tile (MW = 63, so a
// wave spans the tile's width and lane 63 is a helper), and wave w owns the
// first-level rows [w R, w R + R).  Narrow texels take 7 second-level rows;
// 16-byte texels 5, to bound the registers of a lane's input window.  THIN
// tiles (3 second-level rows, 2 first-level rows per wave) are for steps too
// small to fill the GPU: each lane then runs half the serial work, and a
// step's latency is one tile's.
template <uint32_t BYTES, bool THIN = false>
struct Gen2Geom
{
  static constexpr uint32_t TX = 31;
  static constexpr uint32_t TY = THIN ? 3u : (BYTES <= 8u ? 7u : 5u);
  static constexpr uint32_t R  = THIN ? 2u : (BYTES <= 8u ? 4u : 3u);
  static constexpr uint32_t MW = 2u * TX + 1u;  // LDS row stride of the first level
  static constexpr uint32_t MH = 2u * TY + 1u;
  static_assert(MW == 63u, "one lane per first-level column, one wave across the tile");
  static_assert(4u * R >= MH, "the workgroup's four waves cover the tile's first-level rows");
  static_assert(TX * TY <= (uint32_t)kGroupBS, "one pass of the workgroup for the second level");
};

// Groups of a one-level general step (a thread per output) and of a two-level
// one (a Gen2Geom tile of the second level each, thin or not; *tilesX receives
// the tile columns).
__host__ __device__ inline uint32_t generalGroupCount(uint32_t dstWidth, uint32_t dstHeight, uint32_t count,
                                                      uint32_t bytesPerTexel, uint32_t* tilesX, bool thin = false)
{
  if(count == 1u)
  {
    *tilesX = 0;
    return (uint32_t)(((uint64_t)dstWidth * dstHeight + kGroupBS - 1) / kGroupBS);
  }
  const uint32_t tx = Gen2Geom<8>::TX;
  static_assert(Gen2Geom<8>::TX == Gen2Geom<16>::TX && Gen2Geom<8, true>::TX == Gen2Geom<16, true>::TX, "one width");
  const uint32_t ty = thin ? Gen2Geom<8, true>::TY : bytesPerTexel <= 8u ? Gen2Geom<8>::TY : Gen2Geom<16>::TY;
  *tilesX           = (dstWidth + tx - 1u) / tx;
  return *tilesX * ((dstHeight + ty - 1u) / ty);
}

// One level (glsl:828-842): outputs [256 g, 256 g + 256) of level inLevel+1, in
// row-major order, a thread each.  No barrier.
template <class F, bool CO = false, class TT>
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
  storeTexelTo<F, CO>(texelPtr(img, inLevel + 1u, dx, dy, F::kBytes), F::store(tab, v));
}

// LDS values of a two-level general group.
template <class F, bool THIN = false>
constexpr uint32_t gen2MidTexels()
{
  return Gen2Geom<F::kBytes, THIN>::MW * Gen2Geom<F::kBytes, THIN>::MH;
}

// Two levels: the Gen2Geom tile `tile` of level inLevel+2 and the first-level
// texels it needs.
//
// First level (fillIntermediateTile_, glsl:732-781): a lane computes the
// texels of its column in its wave's rows, top to bottom.  Every first-level
// texel is the reference's reduceStoreSample_ (glsl:575-656): a vertical 3-tap
// (or 2-tap) reduction of each of its input columns 2c, 2c+1, 2c+2, then a
// horizontal one.  The vertical reduction of column 2c+2 is the same value as
// that of the next lane's column 2(c+1), so a lane computes only columns 2c
// and 2c+1 and takes the third from lane c+1 with a wave shift (DPP
// wave_shl:1) -- the same operands in the same order, so the same value.  A
// lane thus loads and decodes each input texel of its two columns once: the
// thread's input window (Gen2Window, up to 2R+1 rows x 2 columns) is loaded by
// gen2Load before any arithmetic.  The values go to memory and to `mid`
// ([y][x], stride MW).
//
// Second level (fillLastTile_, glsl:791-820): from LDS, a thread per texel.
// Neighbouring tiles recompute their shared first-level border texels and
// store identical values.
template <class F, bool THIN = false>
struct Gen2Window
{
  typename F::Raw raw[2u * Gen2Geom<F::kBytes, THIN>::R + 1u][2];
};

// Lane i receives lane i+1's value (DPP wave_shl:1); lane 63 receives 0, and
// no caller uses it (it is a helper lane).  With no old value to keep, the
// compiler folds the shift into the instruction that consumes it (a DPP
// operand of the FMA).
template <class V>
__device__ __forceinline__ V waveShiftDown1(V v)
{
  static_assert(sizeof(V) % 4 == 0, "32-bit words");
  uint32_t w[sizeof(V) / 4];
  __builtin_memcpy(w, &v, sizeof v);
#pragma unroll
  for(int k = 0; k < (int)(sizeof(V) / 4); ++k)
    w[k] = (uint32_t)__builtin_amdgcn_mov_dpp((int)w[k], 0x130, 0xf, 0xf, true);
  V r;
  __builtin_memcpy(&r, w, sizeof r);
  return r;
}

struct Gen2Lane
{
  uint32_t gx, gy0, n;  // first-level column, first row of the wave's rows, rows (the same in the whole wave)
  bool out;             // the lane computes first-level texels (else it only lends column 2 gx)
};

// The lane's place in tile `tile` (the same for every kernel size).
template <class F, bool THIN = false>
__device__ __forceinline__ Gen2Lane gen2Lane(const ImageDesc& img, uint32_t inLevel, uint32_t tilesX, uint32_t tile)
{
  using G             = Gen2Geom<F::kBytes, THIN>;
  const LevelDesc& l1 = img.lv[inLevel + 1u];
  const int kx2 = kernelSize(l1.width), ky2 = kernelSize(l1.height);
  const uint32_t lane = threadIdx.x & 63u, ly0 = (threadIdx.x >> 6) * G::R;
  const uint32_t ty = tile / tilesX, tx = tile - ty * tilesX;
  // First-level texels the tile needs (its second level's 2x / 3x windows).
  const uint32_t nx = 2u * G::TX + (kx2 == 3 ? 1u : 0u), ny = 2u * G::TY + (ky2 == 3 ? 1u : 0u);
  Gen2Lane r;
  r.gx  = 2u * G::TX * tx + lane;
  r.gy0 = 2u * G::TY * ty + ly0;
  r.n   = 0u;
  if(ly0 < ny && r.gy0 < l1.height)
  {
    uint32_t n = G::R;
    n          = ny - ly0 < n ? ny - ly0 : n;
    r.n        = l1.height - r.gy0 < n ? l1.height - r.gy0 : n;
  }
  r.out = lane < nx && r.gx < l1.width;
  return r;
}

// KX, KY: the first level's kernel sizes (kernelSize of the input width and
// height) when known at compile time, 0 to read them from the image.  Each
// specialisation computes exactly what the general one does.
template <class F, int KX = 0, int KY = 0, bool THIN = false>
__device__ __forceinline__ void gen2Load(const ImageDesc& img, uint32_t inLevel, uint32_t tilesX, uint32_t tile,
                                         Gen2Window<F, THIN>& w)
{
  constexpr uint32_t NIN = 2u * Gen2Geom<F::kBytes, THIN>::R + 1u;
  const Gen2Lane L       = gen2Lane<F, THIN>(img, inLevel, tilesX, tile);
  const LevelDesc& src   = img.lv[inLevel];
  // Input rows the wave reads (the same for all its lanes; in the level).
  const uint32_t rows = L.n == 0u ? 0u : 2u * (L.n - 1u) + (uint32_t)(KY ? KY : kernelSize(src.height));
  // Every lane loads both of its columns, clamped into the level, rather than
  // branching per load.  Only three kinds of values are used: a computing
  // lane's two columns (the second only for a kernel 2 or 3 wide), and column
  // 2 gx of a lane whose left neighbour computes with a 3-wide kernel, which is
  // then inside the level.  Whatever else a lane reads goes unused.
  const uint32_t last = src.width - 1u;
  const uint32_t x0 = 2u * L.gx < last ? 2u * L.gx : last, x1 = 2u * L.gx + 1u < last ? 2u * L.gx + 1u : last;
  const uint8_t* p  = img.base + src.offset + (uint64_t)(2u * L.gy0) * src.pitch;
#pragma unroll
  for(uint32_t r = 0; r < NIN; ++r)
  {
    if(r < rows)
    {
      w.raw[r][0] = loadTexel<F>(p + (uint64_t)r * src.pitch + (uint64_t)x0 * F::kBytes);
      w.raw[r][1] = loadTexel<F>(p + (uint64_t)r * src.pitch + (uint64_t)x1 * F::kBytes);
    }
    else
    {
      w.raw[r][0] = typename F::Raw{};
      w.raw[r][1] = typename F::Raw{};
    }
  }
}

// The second level of a two-level tile when its horizontal kernel is 3 wide
// (KY2: its vertical kernel size when known at compile time, else 0 and
// `ky2`).  Threads form rows of 32 lanes, TX outputs and a helper; each lane
// takes the vertical taps of mid columns 2 ox and 2 ox + 1 and receives column
// 2 ox + 2 from lane + 1 (its column 2 (ox + 1)); the helper (ox = TX) only
// lends its column.  Per output these are generalSample's operations on the
// same values, from six LDS reads instead of nine.
template <class F, bool CO, int KY2, bool THIN, class TT>
__device__ __forceinline__ void gen2SecondLevel3(const TT& tab, const typename F::Val* mid, const ImageDesc& img,
                                                 uint32_t inLevel, uint32_t tilesX, uint32_t tile, int ky2)
{
  using G = Gen2Geom<F::kBytes, THIN>;
  using V = typename F::Val;
  static_assert(G::TX + 1u == 32u, "a row of outputs and its helper fill half a wave");
  const int ky = KY2 ? KY2 : ky2;
  if(threadIdx.x < 32u * G::TY)  // whole half-waves: the lane shift sees every lane of a row
  {
    const LevelDesc& l2 = img.lv[inLevel + 2u];
    const uint32_t ty = tile / tilesX, tx = tile - ty * tilesX;
    const uint32_t ox = threadIdx.x & 31u, oy = threadIdx.x >> 5;
    const uint32_t x = G::TX * tx + ox, y = G::TY * ty + oy;
    const Weights3 wy = weights3(axis3(l2.height), y);
    const uint32_t c0 = 2u * ox, c1 = c0 + 1u < G::MW ? c0 + 1u : c0;  // the helper's second column is past the row
    V h[2];
#pragma unroll
    for(int j = 0; j < 2; ++j)
    {
      const V* col = mid + (2u * oy) * G::MW + (j ? c1 : c0);
      const V a0   = col[0];
      const V a1   = ky >= 2 ? col[G::MW] : a0;
      const V a2   = ky == 3 ? col[2u * G::MW] : a0;
      h[j]         = verticalTap<F>(wy, ky, a0, a1, a2);
    }
    const V h2 = waveShiftDown1(h[0]);
    if(ox < G::TX && x < l2.width && y < l2.height)
    {
      const V v = horizontalTap<F>(axis3(l2.width), x, 3, h[0], h[1], h2);
      storeTexelTo<F, CO>(texelPtr(img, inLevel + 2u, x, y, F::kBytes), F::store(tab, v));
    }
  }
}

// Contains barriers.  CO: the second level is stored with storeTexelCoherent.
template <class F, bool CO = false, int KX = 0, int KY = 0, bool THIN = false, class TT>
__device__ __forceinline__ void gen2Compute(const TT& tab, typename F::Val* mid, const ImageDesc& img,
                                            uint32_t inLevel, uint32_t tilesX, uint32_t tile,
                                            const Gen2Window<F, THIN>& w)
{
  using G              = Gen2Geom<F::kBytes, THIN>;
  using V              = typename F::Val;
  const LevelDesc& src = img.lv[inLevel];
  const LevelDesc& l1  = img.lv[inLevel + 1u];
  const LevelDesc& l2  = img.lv[inLevel + 2u];
  const int kx1 = KX ? KX : kernelSize(src.width), ky1 = KY ? KY : kernelSize(src.height);
  const int kx2 = kernelSize(l1.width), ky2 = kernelSize(l1.height);
  const Gen2Lane L = gen2Lane<F, THIN>(img, inLevel, tilesX, tile);
  if(L.n != 0u)  // the same in the whole wave: the wave shift below sees every lane
  {
    const Axis3 ax = axis3(l1.width), ay = axis3(l1.height);
    const uint32_t lane = threadIdx.x & 63u, ly0 = (threadIdx.x >> 6) * G::R;
    V top[2];  // decoded input row 2k of output k (the last row of output k-1 when 3-tall)
#pragma unroll
    for(int j = 0; j < 2; ++j)
      top[j] = F::template load<false>(tab, w.raw[0][j]);
#pragma unroll
    for(uint32_t k = 0; k < G::R; ++k)
    {
      if(k < L.n)
      {
        const Weights3 wy = weights3(ay, L.gy0 + k);
        V h[2];
#pragma unroll
        for(int j = 0; j < 2; ++j)
        {
          const V a0 = top[j];
          const V a1 = ky1 >= 2 ? F::template load<false>(tab, w.raw[2u * k + 1u][j]) : a0;
          const V a2 = ky1 == 3 ? F::template load<false>(tab, w.raw[2u * k + 2u][j]) : a0;
          h[j]       = verticalTap<F>(wy, ky1, a0, a1, a2);
          if(k + 1u < L.n)  // the next output's first row: shared (3-tall) or the row after (2-tall)
            top[j] = ky1 == 3 ? a2 : F::template load<false>(tab, w.raw[2u * k + 2u][j]);
        }
        const V h2 = waveShiftDown1(h[0]);  // column 2 gx + 2, from lane + 1
        if(L.out)
        {
          const V v = horizontalTap<F>(ax, L.gx, kx1, h[0], h[1], h2);
          storeTexel<F>(texelPtr(img, inLevel + 1u, L.gx, L.gy0 + k, F::kBytes), F::store(tab, v));
          mid[(ly0 + k) * G::MW + lane] = v;
        }
      }
    }
  }
  __syncthreads();
  if(kx2 == 3)  // the same for every tile of a step
  {
    if(ky2 == 3)
      gen2SecondLevel3<F, CO, 3, THIN>(tab, mid, img, inLevel, tilesX, tile, ky2);
    else if(ky2 == 2)
      gen2SecondLevel3<F, CO, 2, THIN>(tab, mid, img, inLevel, tilesX, tile, ky2);
    else
      gen2SecondLevel3<F, CO, 0, THIN>(tab, mid, img, inLevel, tilesX, tile, ky2);
  }
  else if(threadIdx.x < G::TX * G::TY)
  {
    const uint32_t ty = tile / tilesX, tx = tile - ty * tilesX;
    const uint32_t ox = threadIdx.x % G::TX, oy = threadIdx.x / G::TX;
    const uint32_t x = G::TX * tx + ox, y = G::TY * ty + oy;
    if(x < l2.width && y < l2.height)
    {
      auto at = [&](int j, int r) { return mid[(2u * oy + r) * G::MW + 2u * ox + j]; };
      const V v = generalSampleAny<F>(at, kx2, ky2, axis3(l2.width), axis3(l2.height), x, y);
      storeTexelTo<F, CO>(texelPtr(img, inLevel + 2u, x, y, F::kBytes), F::store(tab, v));
    }
  }
  __syncthreads();
}

// One tile.  Contains barriers.
template <class F, bool CO = false, int KX = 0, int KY = 0, bool THIN = false, class TT>
__device__ __forceinline__ void generalTwoGroup(const TT& tab, typename F::Val* mid, const ImageDesc& img,
                                                uint32_t inLevel, uint32_t tilesX, uint32_t tile)
{
  Gen2Window<F, THIN> w;
  gen2Load<F, KX, KY, THIN>(img, inLevel, tilesX, tile, w);
  gen2Compute<F, CO, KX, KY, THIN>(tab, mid, img, inLevel, tilesX, tile, w);
}

// One tile, through the specialisation for the first level's kernel sizes
// (2 or 3 each; a side of 1 takes the general code).  The choice is the same
// for every tile of a step.
template <class F, bool CO = false, class TT>
__device__ __forceinline__ void generalTwoGroupAny(const TT& tab, typename F::Val* mid, const ImageDesc& img,
                                                   uint32_t inLevel, uint32_t tilesX, uint32_t tile)
{
  const int kx = kernelSize(img.lv[inLevel].width), ky = kernelSize(img.lv[inLevel].height);
  if(kx == 3 && ky == 3)
    generalTwoGroup<F, CO, 3, 3>(tab, mid, img, inLevel, tilesX, tile);
  else if(kx == 2 && ky == 2)
    generalTwoGroup<F, CO, 2, 2>(tab, mid, img, inLevel, tilesX, tile);
  else if(kx == 3 && ky == 2)
    generalTwoGroup<F, CO, 3, 2>(tab, mid, img, inLevel, tilesX, tile);
  else if(kx == 2 && ky == 3)
    generalTwoGroup<F, CO, 2, 3>(tab, mid, img, inLevel, tilesX, tile);
  else
    generalTwoGroup<F, CO>(tab, mid, img, inLevel, tilesX, tile);
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

// The (up to) 3x3 input window of output i of level dl, from level sl.
template <class F>
struct TailWindow
{
  typename F::Raw raw[3][3];
};

template <class F>
__device__ __forceinline__ void tailLoad(const ImageDesc& img, const LevelDesc& sl, const LevelDesc& dl, uint32_t i,
                                         TailWindow<F>& w)
{
  const int kx = kernelSize(sl.width), ky = kernelSize(sl.height);
  const uint32_t dy = i / dl.width, dx = i - dy * dl.width;
  const uint8_t* q  = img.base + sl.offset + (uint64_t)(2u * dy) * sl.pitch + (uint64_t)(2u * dx) * F::kBytes;
#pragma unroll
  for(int k = 0; k < 3; ++k)
#pragma unroll
    for(int j = 0; j < 3; ++j)
      if(j < kx && k < ky)
        w.raw[k][j] = loadTexel<F>(q + (uint64_t)k * sl.pitch + (uint64_t)j * F::kBytes);
}

// The first level reads memory: its outputs go in chunks of PF passes of the
// workgroup, every load of a chunk issued before its arithmetic, so a chunk
// pays one memory latency instead of one per pass (9 PF texels of registers).
// SPEC: general levels through generalSampleAny (constant kernel sizes, more
// code); off where registers are tight.
template <class F, int BS, int PF = 1, bool SPEC = true, class TT>
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
          TailWindow<F> win[PF];
#pragma unroll
          for(int p = 0; p < PF; ++p)
          {
            const uint32_t i = base + (uint32_t)p * BS + threadIdx.x;
            if(i < n)
              tailLoad<F>(img, sl, dl, i, win[p]);
          }
#pragma unroll
          for(int p = 0; p < PF; ++p)
          {
            const uint32_t i = base + (uint32_t)p * BS + threadIdx.x;
            if(i < n)
            {
              const uint32_t dy = i / dl.width, dx = i - dy * dl.width;
              auto at = [&](int j, int k) -> V { return F::template load<false>(tab, win[p].raw[k][j]); };
              finish(i, dx, dy, fast ? box<F, false>(at(0, 0), at(1, 0), at(0, 1), at(1, 1), order)
                                     : (SPEC ? generalSampleAny<F>(at, kx, ky, ax, ay, dx, dy)
                                            : generalSample<F>(at, kx, ky, ax, ay, dx, dy)));
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
                                 : (SPEC ? generalSampleAny<F>(at, kx, ky, ax, ay, dx, dy)
                                            : generalSample<F>(at, kx, ky, ax, ay, dx, dy)));
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
