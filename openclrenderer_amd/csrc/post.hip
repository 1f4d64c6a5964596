// post.hip — passes that consume a rendered frame's depth and colour.
//
// blend_screens_with_depth (cl2.cl:7842-7875), the kernel behind
// engine::blend_with_depth (engine.cpp:2550-2570): composite one frame onto
// another of the same size wherever the source is at least as near, with
// alpha_blend (cl2.cl:1721-1728) on the RGBA8 colours.  HBM-bound: 16 bytes
// read and up to 4 written per pixel, four pixels per work item with 16-byte
// loads.
#pragma clang fp contract(off)

#include "ocr_internal.h"

namespace ocr {

__device__ __forceinline__ uint32_t blend_pixel(uint32_t d1, uint32_t d2, uint32_t s, uint32_t d) {
  if (d2 < d1) return d;   // the destination is nearer
  float sc[4], dc[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    // c / 255 correctly rounded (div_rcp is exact for these integers)
    sc[c] = div_rcp((float)((s >> (8 * c)) & 0xFFu), 255.f, kInv255);
    dc[c] = div_rcp((float)((d >> (8 * c)) & 0xFFu), 255.f, kInv255);
  }
  const float inv = 1.f - sc[3];
  uint32_t out = quant8(sc[3]) << 24;
#pragma unroll
  for (int c = 0; c < 3; c++) out |= quant8(fmaf(sc[c], sc[3], dc[c] * inv)) << (8 * c);
  return out;
}

__global__ void __launch_bounds__(256) k_blend_depth(const uint32_t* __restrict__ src_depth,
                                                     const uint32_t* __restrict__ dst_depth,
                                                     const uint32_t* __restrict__ src_rgba,
                                                     uint32_t* __restrict__ dst_rgba, uint32_t npix) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t nvec = npix / 4;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nvec; v += stride) {
    const uint4 d1 = reinterpret_cast<const uint4*>(src_depth)[v];
    const uint4 d2 = reinterpret_cast<const uint4*>(dst_depth)[v];
    const uint4 s = reinterpret_cast<const uint4*>(src_rgba)[v];
    const uint4 d = reinterpret_cast<const uint4*>(dst_rgba)[v];
    uint4 o;
    o.x = blend_pixel(d1.x, d2.x, s.x, d.x);
    o.y = blend_pixel(d1.y, d2.y, s.y, d.y);
    o.z = blend_pixel(d1.z, d2.z, s.z, d.z);
    o.w = blend_pixel(d1.w, d2.w, s.w, d.w);
    if (o.x != d.x || o.y != d.y || o.z != d.z || o.w != d.w) reinterpret_cast<uint4*>(dst_rgba)[v] = o;
  }
  // the last npix % 4 pixels
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t p = nvec * 4 + t;
  if (t < npix % 4) dst_rgba[p] = blend_pixel(src_depth[p], dst_depth[p], src_rgba[p], dst_rgba[p]);
}

hipError_t launch_blend_depth(const uint32_t* src_depth, const uint32_t* dst_depth, const uint32_t* src_rgba,
                              uint32_t* dst_rgba, uint32_t npix, hipStream_t s) {
  const uint32_t nvec = npix / 4;
  const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((nvec + 255) / 256, (uint32_t)device_cu_count() * 8));
  hipLaunchKernelGGL(k_blend_depth, dim3(blocks), dim3(256), 0, s, src_depth, dst_depth, src_rgba, dst_rgba, npix);
  return hipGetLastError();
}

}  // namespace ocr

using namespace ocr;

extern "C" int ocr_frame_blend_with_depth(ocr_frame* src, ocr_frame* dst, void* stream) {
  if (!src || !dst || src == dst) return fail(OCR_ERR_INVALID, "ocr_frame_blend_with_depth: two distinct frames needed");
  if (src->ctx != dst->ctx) return fail(OCR_ERR_INVALID, "ocr_frame_blend_with_depth: frames of different contexts");
  if (src->width != dst->width || src->height != dst->height)
    return fail(OCR_ERR_INVALID, "ocr_frame_blend_with_depth: frame sizes differ");
  OCR_HIP_CHECK(hipSetDevice(dst->ctx->device));
  OCR_HIP_CHECK(launch_blend_depth(src->depth, dst->depth, src->rgba, dst->rgba, dst->width * dst->height,
                                   (hipStream_t)stream));
  return OCR_OK;
}
