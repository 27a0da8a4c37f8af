// Level-0 conversion on the device: premultiplied alpha.
//
// The reference premultiplies while staging a file on the host
// (include/scoped_image.hpp:233-252): alpha = a / 255, rgb = srgbFromLinear(
// linearFromSrgb(c) * alpha), alpha byte unchanged.  Here the same formula runs
// on level 0 in device memory, one thread per texel, with the functors' exact
// conversions (the FP32 rounding of the exact sRGB decode; the exact encode)
// in place of the host's powf.
#include "functors.hpp"
#include "pyramid_internal.hpp"

namespace nvpro_pyramid {

__global__ void __launch_bounds__(256) premultiplySrgba8Kernel(uint8_t* base, uint32_t width, uint32_t height,
                                                               uint32_t pitch)
{
  __shared__ Srgba8::Tables tab;
  Srgba8::fill<256>(tab, 1.0f);
  __syncthreads();
  const uint64_t n      = (uint64_t)width * height;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
  {
    const uint32_t y = (uint32_t)(t / width), x = (uint32_t)(t - (uint64_t)y * width);
    uint32_t* p      = reinterpret_cast<uint32_t*>(base + (uint64_t)y * pitch + (uint64_t)x * 4u);
    const uint32_t c = *p;
    const float4 v   = Srgba8::load<false>(tab, c);  // linear RGB
    const float a    = (float)(c >> 24) * (1.0f / 255.0f);
    const uint32_t r = Srgba8::encodeChannel(tab, v.x * a);
    const uint32_t g = Srgba8::encodeChannel(tab, v.y * a);
    const uint32_t b = Srgba8::encodeChannel(tab, v.z * a);
    *p               = r | g << 8 | b << 16 | (c & 0xff000000u);
  }
}

// RGBA16F: rgb * a in FP32, rounded to binary16; alpha unchanged.
__global__ void __launch_bounds__(256) premultiplyRgba16fKernel(uint8_t* base, uint32_t width, uint32_t height,
                                                                uint32_t pitch)
{
  const uint64_t n      = (uint64_t)width * height;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
  {
    const uint32_t y = (uint32_t)(t / width), x = (uint32_t)(t - (uint64_t)y * width);
    uint2* p         = reinterpret_cast<uint2*>(base + (uint64_t)y * pitch + (uint64_t)x * 8u);
    const uint2 q    = *p;
    const float a    = Rgba16f::h2f(q.y >> 16);
    const uint32_t r = Rgba16f::f2h(Rgba16f::h2f(q.x & 0xffffu) * a);
    const uint32_t g = Rgba16f::f2h(Rgba16f::h2f(q.x >> 16) * a);
    const uint32_t b = Rgba16f::f2h(Rgba16f::h2f(q.y & 0xffffu) * a);
    *p               = make_uint2(r | g << 16, b | (q.y & 0xffff0000u));
  }
}

}  // namespace nvpro_pyramid

using namespace nvpro_pyramid;

extern "C" int nvproPyramidPremultiplyAlpha(hipStream_t stream, uint32_t format, const NvproPyramidImage* image)
{
  const char* who = "nvproPyramidPremultiplyAlpha";
  if(!image || !image->data)
    return hostFail(hipErrorInvalidValue, "%s: null image", who);
  if(format != NVPRO_PYRAMID_FORMAT_SRGBA8 && format != NVPRO_PYRAMID_FORMAT_RGBA16F)
    return hostFail(hipErrorInvalidValue, "%s: format %u not supported", who, format);
  if(image->width == 0 || image->height == 0)
    return hostFail(hipErrorInvalidValue, "%s: zero image size", who);
  const uint32_t bpt   = nvproPyramidBytesPerTexel(format);
  const uint64_t pitch = image->rowPitches ? image->rowPitches[0] : (uint64_t)image->width * bpt;
  const uint64_t off   = image->levelOffsets ? image->levelOffsets[0] : 0u;
  if(pitch < (uint64_t)image->width * bpt || pitch > 0xffffffffull || pitch % bpt != 0 ||
     (((uintptr_t)image->data + off) % bpt) != 0)
    return hostFail(hipErrorInvalidValue, "%s: level 0 layout not aligned to %u bytes", who, bpt);
  uint8_t* base       = (uint8_t*)image->data + off;
  const uint64_t n    = (uint64_t)image->width * image->height;
  const uint64_t want = (n + 255u) / 256u;
  const uint32_t cap  = deviceComputeUnits() * 8u;
  const uint32_t grid = (uint32_t)(want < cap ? want : cap);
  if(format == NVPRO_PYRAMID_FORMAT_SRGBA8)
    hipLaunchKernelGGL(premultiplySrgba8Kernel, dim3(grid), dim3(256), 0, stream, base, image->width, image->height,
                       (uint32_t)pitch);
  else
    hipLaunchKernelGGL(premultiplyRgba16fKernel, dim3(grid), dim3(256), 0, stream, base, image->width, image->height,
                       (uint32_t)pitch);
  hipError_t e = hipGetLastError();
  if(e != hipSuccess)
    return hostFail(e, "%s: launch failed: %s", who, hipGetErrorString(e));
  return 0;
}
