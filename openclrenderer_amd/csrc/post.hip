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

// decode_normal(ushort2), cl2.cl:5611-5646: short_to_float is
// (s / 65535) * 2 - 1 (the POW2e16-1 macro binds as 65536 - 1).
__device__ __forceinline__ f3 decode_normal_u16(uint32_t packed) {
  const float vx = div_rcp((float)(packed & 0xFFFFu), 65535.f, kInv65535) * 2.f - 1.f;
  const float vy = div_rcp((float)(packed >> 16), 65535.f, kInv65535) * 2.f - 1.f;
  const float z = (vx * vx + vy * vy) * 2.f - 1.f;
  const float r = inv_sqrt_cr(vx * vx + vy * vy);
  const float k = sqrtf(fmaxf(1.f - z * z, 0.f));
  return mk3((vx * r) * k, (vy * r) * k, z);
}

__device__ __forceinline__ f3 unorm_rgb(uint32_t c) {
  return mk3(div_rcp((float)(c & 0xFFu), 255.f, kInv255), div_rcp((float)((c >> 8) & 0xFFu), 255.f, kInv255),
             div_rcp((float)((c >> 16) & 0xFFu), 255.f, kInv255));
}

// do_pseudo_aa, cl2.cl:6438-6657: one work item per interior pixel.  The
// eight neighbours are split by depth difference > 100 and by a normal
// turned more than 20 degrees (cos_limit); the first test (then the second)
// with exactly one differing x neighbour, one differing y neighbour and at
// least one differing corner sets the pixel to 0.65 * mean(same side) +
// 0.35 * mean(other side), alpha 1.  Colours come from `in`, the frame's
// colours before the pass; the reference's avg_depth / avg_normal feed only
// its disabled REDUCED_AA exit and are not formed.
__global__ void __launch_bounds__(256) k_pseudo_aa(const uint32_t* __restrict__ depth,
                                                   const uint32_t* __restrict__ normals,
                                                   const uint32_t* __restrict__ in, uint32_t* __restrict__ rgba,
                                                   uint32_t W, uint32_t H, float cos_limit) {
  const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x < 1 || y < 1 || x + 1 >= W || y + 1 >= H) return;
  const uint32_t p = y * W + x;
  const uint32_t my_raw = depth[p];
  if ((float)my_raw == kMulint) return;
  const f3 mn = fast_normalize(decode_normal_u16(normals[p]));
  const float my_depth = ((float)my_raw / kMulint) * kDepthFar;
  int nx[2] = {0, 0}, ny[2] = {0, 0}, nc[2] = {0, 0};
  f3 mine[2] = {mk3(0, 0, 0), mk3(0, 0, 0)}, theirs[2] = {mk3(0, 0, 0), mk3(0, 0, 0)};
  float ms[2] = {0.f, 0.f}, ts[2] = {0.f, 0.f};
#pragma unroll
  for (int j = -1; j < 2; j++)
#pragma unroll
    for (int i = -1; i < 2; i++) {
      if (i == 0 && j == 0) continue;
      const uint32_t q = (uint32_t)((int)p + j * (int)W + i);
      const float d = ((float)depth[q] / kMulint) * kDepthFar;
      const f3 fnrm = decode_normal_u16(normals[q]);
      const f3 val = unorm_rgb(in[q]);
      const bool tests[2] = {fabsf(d - my_depth) > 100.f, dot(mn, fnrm) < cos_limit};
#pragma unroll
      for (int k = 0; k < 2; k++) {
        if (tests[k]) {
          if (i == j || i == -j) nc[k]++;
          else if (i == 1 || i == -1) nx[k]++;
          else ny[k]++;
          theirs[k] = theirs[k] + val;
          ts[k] += 1.f;
        } else {
          mine[k] = mine[k] + val;
          ms[k] += 1.f;
        }
      }
    }
#pragma unroll
  for (int k = 0; k < 2; k++) {
    if (nx[k] == 1 && ny[k] == 1 && nc[k] >= 1) {
      const f3 a = mine[k] / ms[k], b = theirs[k] / ts[k];
      const f3 acc = a * 0.65f + b * 0.35f;
      rgba[p] = quant8(acc.x) | (quant8(acc.y) << 8) | (quant8(acc.z) << 16) | (255u << 24);
      return;
    }
  }
}

hipError_t launch_pseudo_aa(const uint32_t* depth, const uint32_t* normals, const uint32_t* in, uint32_t* rgba,
                            uint32_t W, uint32_t H, hipStream_t s) {
  // cos(20 * 2 * M_PI / 360) in single precision (argument 0.34906587f)
  const float cos_limit = (float)cos((double)(20.f * 2.f * 3.14159265f / 360.f));
  hipLaunchKernelGGL(k_pseudo_aa, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, s, depth, normals, in, rgba, W, H,
                     cos_limit);
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

// engine::do_pseudo_aa (engine.cpp:1513-1516): the frame's colours are
// snapshotted first (the reference reads and writes the same image), then
// rewritten where an edge qualifies.
extern "C" int ocr_frame_pseudo_aa(ocr_frame* f, void* stream) {
  if (!f) return fail(OCR_ERR_INVALID, "ocr_frame_pseudo_aa: null frame");
  OCR_HIP_CHECK(hipSetDevice(f->ctx->device));
  const size_t npix = (size_t)f->width * f->height;
  if (!f->post_in) {
    hipError_t e = hipMalloc((void**)&f->post_in, 4 * npix);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      f->post_in = nullptr;
      return fail(OCR_ERR_OOM, std::string("ocr_frame_pseudo_aa: ") + hipGetErrorString(e));
    }
  }
  hipStream_t st = (hipStream_t)stream;
  OCR_HIP_CHECK(hipMemcpyAsync(f->post_in, f->rgba, 4 * npix, hipMemcpyDeviceToDevice, st));
  OCR_HIP_CHECK(launch_pseudo_aa(f->depth, f->normals, f->post_in, f->rgba, f->width, f->height, st));
  return OCR_OK;
}
