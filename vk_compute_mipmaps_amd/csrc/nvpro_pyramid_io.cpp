// Host-side file formats of the chain: per-level TGA files.
//
// writeMipmapsTga (include/mipmap_storage.hpp:441-479 of the reference) writes
// every level of a packed sRGBA8 chain to its own TGA file: level 0 to the base
// name, level l to the base name with "l" inserted before the extension
// ("out.tga" -> "out.1.tga").  This writes the same set of files, uncompressed
// 32-bit, top-left origin, BGRA texel order.  The reader takes uncompressed or
// run-length encoded 24/32-bit true-colour TGA, either origin (the reference
// writes through stb_image_write, which emits run-length encoded files).
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "pyramid_internal.hpp"

using namespace nvpro_pyramid;

namespace {

std::string levelFilename(const char* base, uint32_t level)
{
  if(level == 0)
    return base;
  const char* dot = strrchr(base, '.');
  const std::string prefix = dot ? std::string(base, (size_t)(dot - base) + 1u) : std::string(base);
  const std::string suffix = dot ? std::string(dot) : std::string();
  return prefix + std::to_string(level) + suffix;
}

int writeTga(const std::string& name, const uint8_t* rgba, uint32_t w, uint32_t h)
{
  FILE* f = fopen(name.c_str(), "wb");
  if(!f)
    return hostFail(hipErrorInvalidValue, "nvproPyramidWriteTga: cannot open %s", name.c_str());
  uint8_t hdr[18] = {};
  hdr[2]          = 2;  // uncompressed true colour
  hdr[12]         = (uint8_t)(w & 0xffu);
  hdr[13]         = (uint8_t)(w >> 8);
  hdr[14]         = (uint8_t)(h & 0xffu);
  hdr[15]         = (uint8_t)(h >> 8);
  hdr[16]         = 32;
  hdr[17]         = 0x28;  // 8 alpha bits, top-left origin
  bool ok         = fwrite(hdr, 1, sizeof hdr, f) == sizeof hdr;
  std::vector<uint8_t> row((size_t)w * 4u);
  for(uint32_t y = 0; ok && y < h; ++y)
  {
    const uint8_t* src = rgba + (size_t)y * w * 4u;
    for(uint32_t x = 0; x < w; ++x)
    {
      row[4u * x + 0u] = src[4u * x + 2u];
      row[4u * x + 1u] = src[4u * x + 1u];
      row[4u * x + 2u] = src[4u * x + 0u];
      row[4u * x + 3u] = src[4u * x + 3u];
    }
    ok = fwrite(row.data(), 1, row.size(), f) == row.size();
  }
  ok = (fclose(f) == 0) && ok;
  if(!ok)
    return hostFail(hipErrorInvalidValue, "nvproPyramidWriteTga: writing %s failed", name.c_str());
  return 0;
}

}  // namespace

extern "C" int nvproPyramidWriteTga(const void* hostChain, uint32_t width, uint32_t height, uint32_t levelCount,
                                    const char* baseFilename)
{
  const char* who = "nvproPyramidWriteTga";
  if(!hostChain || !baseFilename)
    return hostFail(hipErrorInvalidValue, "%s: null argument", who);
  if(width == 0 || height == 0 || width > 65535u || height > 65535u)
    return hostFail(hipErrorInvalidValue, "%s: %ux%u is outside TGA's 1..65535", who, width, height);
  const uint32_t full = nvproPyramidFullLevelCount(width, height);
  if(levelCount == 0)
    levelCount = full;
  if(levelCount > full || levelCount > NVPRO_PYRAMID_MAX_LEVELS)
    return hostFail(hipErrorInvalidValue, "%s: %u levels for %ux%u", who, levelCount, width, height);
  uint64_t offsets[NVPRO_PYRAMID_MAX_LEVELS];
  if(int e = nvproPyramidPackedLayout(width, height, levelCount, 4u, offsets, nullptr))
    return e;
  const uint8_t* base = (const uint8_t*)hostChain;
  for(uint32_t l = 0; l < levelCount; ++l)
  {
    const uint32_t lw = width >> l ? width >> l : 1u, lh = height >> l ? height >> l : 1u;
    if(int e = writeTga(levelFilename(baseFilename, l), base + offsets[l], lw, lh))
      return e;
  }
  hostClearError();
  return 0;
}

extern "C" int nvproPyramidReadTga(const char* filename, void* rgba, uint64_t capacity, uint32_t* width,
                                   uint32_t* height)
{
  const char* who = "nvproPyramidReadTga";
  if(!filename || !width || !height)
    return hostFail(hipErrorInvalidValue, "%s: null argument", who);
  FILE* f = fopen(filename, "rb");
  if(!f)
    return hostFail(hipErrorInvalidValue, "%s: cannot open %s", who, filename);
  std::vector<uint8_t> data;
  uint8_t buf[65536];
  size_t n;
  while((n = fread(buf, 1, sizeof buf, f)) > 0)
    data.insert(data.end(), buf, buf + n);
  fclose(f);
  if(data.size() < 18)
    return hostFail(hipErrorInvalidValue, "%s: %s is not a TGA file", who, filename);
  const uint8_t* h   = data.data();
  const uint32_t idLen = h[0], cmapType = h[1], type = h[2], bpp = h[16], desc = h[17];
  const uint32_t w = h[12] | (uint32_t)h[13] << 8, hh = h[14] | (uint32_t)h[15] << 8;
  if(cmapType != 0 || (type != 2 && type != 10) || (bpp != 24 && bpp != 32) || w == 0 || hh == 0)
    return hostFail(hipErrorInvalidValue, "%s: %s: only 24/32-bit true-colour TGA is read", who, filename);
  *width  = w;
  *height = hh;
  const uint64_t need = (uint64_t)w * hh * 4u;
  if(!rgba)
    return 0;
  if(capacity < need)
    return hostFail(hipErrorInvalidValue, "%s: %llu bytes needed", who, (unsigned long long)need);
  const uint32_t px = bpp / 8u;
  size_t pos        = 18u + idLen;
  std::vector<uint8_t> texels((size_t)w * hh * px);
  if(type == 2)
  {
    if(data.size() < pos + texels.size())
      return hostFail(hipErrorInvalidValue, "%s: %s is truncated", who, filename);
    memcpy(texels.data(), data.data() + pos, texels.size());
  }
  else
  {
    size_t out = 0;
    while(out < texels.size())
    {
      if(pos >= data.size())
        return hostFail(hipErrorInvalidValue, "%s: %s is truncated", who, filename);
      const uint8_t c    = data[pos++];
      const size_t count = (size_t)(c & 0x7fu) + 1u;
      if(out + count * px > texels.size())
        return hostFail(hipErrorInvalidValue, "%s: %s: run past the image", who, filename);
      if(c & 0x80u)
      {
        if(pos + px > data.size())
          return hostFail(hipErrorInvalidValue, "%s: %s is truncated", who, filename);
        for(size_t k = 0; k < count; ++k)
          memcpy(texels.data() + out + k * px, data.data() + pos, px);
        pos += px;
      }
      else
      {
        if(pos + count * px > data.size())
          return hostFail(hipErrorInvalidValue, "%s: %s is truncated", who, filename);
        memcpy(texels.data() + out, data.data() + pos, count * px);
        pos += count * px;
      }
      out += count * px;
    }
  }
  const bool topDown = (desc & 0x20u) != 0;
  uint8_t* dst       = (uint8_t*)rgba;
  for(uint32_t y = 0; y < hh; ++y)
  {
    const uint8_t* src = texels.data() + (size_t)(topDown ? y : hh - 1u - y) * w * px;
    for(uint32_t x = 0; x < w; ++x)
    {
      uint8_t* d = dst + ((size_t)y * w + x) * 4u;
      d[0]       = src[px * x + 2u];
      d[1]       = src[px * x + 1u];
      d[2]       = src[px * x + 0u];
      d[3]       = px == 4u ? src[px * x + 3u] : 255u;
    }
  }
  hostClearError();
  return 0;
}
