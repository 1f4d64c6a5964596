// atlas.hip — texture atlas upload and mip generation.
//
// update_gpu_tex (cl2.cl:923-953) writes a source RGBA8 image into its atlas
// slot (UNORM8 -> float -> *255 -> truncate, rows flipped y -> width - y);
// generate_mips / generate_mip_mips (cl2.cl:1070-1189) build level L+1 from
// level L with a 3x3 (1 2 1) Gaussian on premultiplied alpha.  Slots are
// addressed through the reference's descriptors: nums[id] = slice << 16 |
// which, sizes[slice] = slot size, tile origin (which % (2048/size),
// which / (2048/size)) * size (read_tex_array / write_tex_array,
// cl2.cl:785-889).
#pragma clang fp contract(off)

#include "ocr_internal.h"

namespace ocr {

struct AtlasRef {
  uint8_t* pages;
  const uint32_t* nums;
  const uint32_t* sizes;
};

// read_tex_array, cl2.cl:785-818.
__device__ __forceinline__ f4 atlas_read(const AtlasRef& a, float cx, float cy, uint32_t tid) {
  int nv = (int)a.nums[tid];
  int slice = nv >> 16, which = nv & 0xFFFF;
  float width = (float)a.sizes[slice];
  float hnum = floorf(2048.f / width);
  float tnumy = floorf((float)which / hnum);
  float tnumx = fmaf(-tnumy, hnum, (float)which);
  cx = clampf(cx, 0.001f, width - 0.001f);
  cy = clampf(cy, 0.001f, width - 0.001f);
  int ix = (int)fmaf(tnumx, width, cx);
  int iy = (int)fmaf(tnumy, width, cy);
  uint32_t c = *reinterpret_cast<const uint32_t*>(a.pages + ((size_t)slice * 2048 * 2048 + (size_t)iy * 2048 + ix) * 4);
  return f4{(float)(c & 0xFF), (float)((c >> 8) & 0xFF), (float)((c >> 16) & 0xFF), (float)(c >> 24)};
}

// write_tex_array, cl2.cl:856-889 (the uchar4 store keeps the low byte).
__device__ __forceinline__ void atlas_write(const AtlasRef& a, const uint32_t c[4], float cx, float cy,
                                            uint32_t tid) {
  int nv = (int)a.nums[tid];
  int slice = nv >> 16, which = nv & 0xFFFF;
  float width = (float)a.sizes[slice];
  float hnum = floorf(2048.f / width);
  float tnumy = floorf((float)which / hnum);
  float tnumx = fmaf(-tnumy, hnum, (float)which);
  float tx = tnumx * width, ty = tnumy * width;
  cx = clampf(fmodf(cx, width), 0.001f, width - 0.001f);
  cy = clampf(fmodf(cy, width), 0.001f, width - 0.001f);
  int ix = (int)(tx + cx), iy = (int)(ty + cy);
  uint32_t v = (c[0] & 0xFF) | ((c[1] & 0xFF) << 8) | ((c[2] & 0xFF) << 16) | ((c[3] & 0xFF) << 24);
  *reinterpret_cast<uint32_t*>(a.pages + ((size_t)slice * 2048 * 2048 + (size_t)iy * 2048 + ix) * 4) = v;
}

__global__ void __launch_bounds__(256) k_atlas_upload(AtlasRef a, uint32_t tid, const uint8_t* __restrict__ src,
                                                      uint32_t w, uint32_t h) {
  const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15);
  const uint32_t y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= w || y >= h) return;
  const uint8_t* s = src + ((size_t)y * w + x) * 4;
  uint32_t c[4];
#pragma unroll
  for (int k = 0; k < 4; k++) c[k] = sat_u32(((float)s[k] / 255.0f) * 255.f);
  float width = (float)a.sizes[a.nums[tid] >> 16];
  atlas_write(a, c, (float)x, width - (float)(int)y, tid);
}

__global__ void __launch_bounds__(256) k_atlas_mip(AtlasRef a, uint32_t src_id, uint32_t dst_id, uint32_t gw,
                                                   uint32_t gh) {
  const int x = (int)(blockIdx.x * 16 + (threadIdx.x & 15));
  const int y = (int)(blockIdx.y * 16 + (threadIdx.x >> 4));
  if (x >= (int)gw || y >= (int)gh) return;
  float width = (float)a.sizes[a.nums[src_id] >> 16];
  if ((float)x >= width || (float)y >= width) return;
  const float gauss[3][3] = {{1, 2, 1}, {2, 4, 2}, {1, 2, 1}};
  float ax = 0, ay = 0, az = 0, aw = 0, div = 0.f;
#pragma unroll
  for (int j = -1; j <= 1; j++)
#pragma unroll
    for (int i = -1; i <= 1; i++) {
      f4 col = atlas_read(a, (float)(x * 2 + i), (float)(y * 2 + j), src_id);
      col.w /= 255.f;
      col.x *= col.w; col.y *= col.w; col.z *= col.w;
      float g = gauss[j + 1][i + 1];
      ax += col.x * g; ay += col.y * g; az += col.z * g; aw += col.w * g;
      div += g;
    }
  ax /= div; ay /= div; az /= div; aw /= div;
  if (aw > 0.00000001f) { ax /= aw; ay /= aw; az /= aw; }
  aw *= 255.f;
  float nwidth = (float)a.sizes[a.nums[dst_id] >> 16];
  float yx = ((float)(x * 2) / width) * nwidth;
  float yy = ((float)(y * 2) / width) * nwidth;
  if (yx >= nwidth || yy >= nwidth) return;
  uint32_t c[4] = {sat_u32(ax), sat_u32(ay), sat_u32(az), sat_u32(aw)};
  atlas_write(a, c, yx, yy, dst_id);
}

hipError_t launch_atlas_upload(const Atlas& at, uint32_t tid, const uint8_t* src, uint32_t w, uint32_t h,
                               hipStream_t s) {
  AtlasRef a{at.pages, at.nums, at.sizes};
  hipLaunchKernelGGL(k_atlas_upload, dim3((w + 15) / 16, (h + 15) / 16), dim3(256), 0, s, a, tid, src, w, h);
  return hipGetLastError();
}

hipError_t launch_mip(const Atlas& at, uint32_t src_id, uint32_t dst_id, uint32_t gw, uint32_t gh,
                      hipStream_t s) {
  AtlasRef a{at.pages, at.nums, at.sizes};
  hipLaunchKernelGGL(k_atlas_mip, dim3((gw + 15) / 16, (gh + 15) / 16), dim3(256), 0, s, a, src_id, dst_id,
                     gw, gh);
  return hipGetLastError();
}

}  // namespace ocr
