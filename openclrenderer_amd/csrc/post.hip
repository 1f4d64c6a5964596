// post.hip — passes that consume a rendered frame's depth and colour.
//
// blend_screens_with_depth (cl2.cl:7842-7875), the kernel behind
// engine::blend_with_depth (engine.cpp:2550-2570): composite one frame onto
// another of the same size wherever the source is at least as near, with
// alpha_blend (cl2.cl:1721-1728) on the RGBA8 colours.  HBM-bound: 16 bytes
// read and up to 4 written per pixel, four pixels per work item with 16-byte
// loads.
#pragma clang fp contract(off)

#include <string.h>

#include "ocr_internal.h"

namespace ocr {

__device__ __forceinline__ uint32_t blend_pixel(uint32_t d1, uint32_t d2, uint32_t s, uint32_t d) {
  if (d2 < d1) return d;   // the destination is nearer
  float sc[4], dc[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    // c / 255 correctly rounded
    sc[c] = unorm8((s >> (8 * c)) & 0xFFu);
    dc[c] = unorm8((d >> (8 * c)) & 0xFFu);
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
  const float k = sqrt_cr(fmaxf(1.f - z * z, 0.f));
  return mk3((vx * r) * k, (vy * r) * k, z);
}

__device__ __forceinline__ f3 unorm_rgb(uint32_t c) {
  return mk3(unorm8(c & 0xFFu), unorm8((c >> 8) & 0xFFu), unorm8((c >> 16) & 0xFFu));
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
  const float my_depth = depth_of(my_raw);
  int nx[2] = {0, 0}, ny[2] = {0, 0}, nc[2] = {0, 0};
  f3 mine[2] = {mk3(0, 0, 0), mk3(0, 0, 0)}, theirs[2] = {mk3(0, 0, 0), mk3(0, 0, 0)};
  float ms[2] = {0.f, 0.f}, ts[2] = {0.f, 0.f};
#pragma unroll
  for (int j = -1; j < 2; j++)
#pragma unroll
    for (int i = -1; i < 2; i++) {
      if (i == 0 && j == 0) continue;
      const uint32_t q = (uint32_t)((int)p + j * (int)W + i);
      const float d = depth_of(depth[q]);
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

// ---- do_motion_blur (cl2.cl:6715-6858) ----------------------------------
//
// Per covered pixel: un-project with the quantised depth and the current
// camera, carry the point into object space with the object's current pose,
// out again with the pose of the object's last motion-blur pass and project
// it with the previous camera (and with the current one, for the object-only
// motion); blur along the screen vector from there to the pixel with up to
// 50 bilinear samples of the colours as they were before the pass.  The pass
// also records each drawn object's current pose in the slot the next frame
// parity reads (cl2.cl:6770-6785).
struct MotionArgs {
  const uint32_t* depth;
  const unsigned long long* idkey;
  const CutTri* cuts;
  const uint32_t* counters;
  const ObjDesc* objs;
  float* motion;            // 16 floats per object
  uint32_t nobj;
  const uint32_t* in;       // colour snapshot
  uint32_t* rgba;
  uint32_t W, H;
  float fov;
  float cpos[3], cpos_old[3];
  SinCos cam, cam_old;
  uint32_t frame_id;
  float strength, camera_contribution;
};

__device__ __forceinline__ f4 unorm4(uint32_t c) {
  return f4{unorm8(c & 0xFFu), unorm8((c >> 8) & 0xFFu), unorm8((c >> 16) & 0xFFu), unorm8(c >> 24)};
}

// read_imagef with CLK_FILTER_LINEAR, unnormalised coordinates (u, v)
// (OpenCL 1.2 section 8.2): i0 = floor(u - 0.5), a = frac(u - 0.5), the four
// texels weighted (1-a)(1-b), a(1-b), (1-a)b, ab; texels beyond the image
// edge read the edge.
__device__ __forceinline__ f4 sample_linear(const uint32_t* __restrict__ img, uint32_t W, uint32_t H, float u,
                                            float v) {
  const float tx = u - 0.5f, ty = v - 0.5f;
  const float fx = floorf(tx), fy = floorf(ty);
  const float a = tx - fx, b = ty - fy;
  const int i0 = min(max((int)fx, 0), (int)W - 1), i1 = min(max((int)fx + 1, 0), (int)W - 1);
  const int j0 = min(max((int)fy, 0), (int)H - 1), j1 = min(max((int)fy + 1, 0), (int)H - 1);
  const f4 t00 = unorm4(img[(uint32_t)j0 * W + (uint32_t)i0]), t10 = unorm4(img[(uint32_t)j0 * W + (uint32_t)i1]);
  const f4 t01 = unorm4(img[(uint32_t)j1 * W + (uint32_t)i0]), t11 = unorm4(img[(uint32_t)j1 * W + (uint32_t)i1]);
  const float w00 = (1.f - a) * (1.f - b), w10 = a * (1.f - b), w01 = (1.f - a) * b, w11 = a * b;
  return f4{((w00 * t00.x + w10 * t10.x) + w01 * t01.x) + w11 * t11.x,
            ((w00 * t00.y + w10 * t10.y) + w01 * t01.y) + w11 * t11.y,
            ((w00 * t00.z + w10 * t10.z) + w01 * t01.z) + w11 * t11.z,
            ((w00 * t00.w + w10 * t10.w) + w01 * t01.w) + w11 * t11.w};
}

// float -> int as the hardware converts: NaN to 0, saturating
__device__ __forceinline__ int to_int_sat(float m) {
  if (!(m == m)) return 0;
  if (m >= 2147483648.f) return 0x7FFFFFFF;
  if (m <= -2147483648.f) return (int)0x80000000;
  return (int)m;
}

__global__ void __launch_bounds__(256) k_motion_blur(const MotionArgs a) {
  const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  const uint32_t W = a.W, H = a.H;
  bool covered = x < W && y < H && a.counters[kOverflow] == 0u;
  const uint32_t pix = y * W + x;
  uint32_t ft = 0xFFFFFFFFu, oid = 0;
  if (covered) {
    ft = a.depth[pix];
    const unsigned long long key = a.idkey[pix];
    covered = ft != 0xFFFFFFFFu && key != 0ull;
    if (covered) {
      oid = a.cuts[(uint32_t)key].oid;
      covered = oid < a.nobj;
    }
  }
  // Slots: an even frame id reads pose 1 and records pose 2, an odd one the
  // reverse.  Every pixel of an object records the same pose, so one lane
  // per object and wave stores it.
  const bool even = (a.frame_id & 1u) == 0u;
  const uint32_t rd_pos = even ? 0u : 4u, rd_q = even ? 8u : 12u;
  const uint32_t wr_pos = even ? 4u : 0u, wr_q = even ? 12u : 8u;
  unsigned long long pending = __ballot(covered);
  while (pending) {
    const int lead = __ffsll(pending) - 1;
    const uint32_t o = __shfl(oid, lead);
    pending &= ~__ballot(covered && oid == o);
    if ((int)__lane_id() == lead) {
      const ObjDesc G = a.objs[o];
      float* m = a.motion + 16 * (size_t)o;
      m[wr_pos] = G.wp[0];
      m[wr_pos + 1] = G.wp[1];
      m[wr_pos + 2] = G.wp[2];
      *reinterpret_cast<float4*>(m + wr_q) = make_float4(G.q[0], G.q[1], G.q[2], G.q[3]);
    }
  }
  if (!covered) return;

  const ObjDesc G = a.objs[oid];
  const float* m = a.motion + 16 * (size_t)oid;
  const f3 old_pos = mk3(m[rd_pos], m[rd_pos + 1], m[rd_pos + 2]);
  const f4 old_q = f4{m[rd_q], m[rd_q + 1], m[rd_q + 2], m[rd_q + 3]};
  const float fov = a.fov;
  const float actual_depth = depth_of(ft);
  const bool fov_fast = rcp_safe(fov);
  const float ifov = fov_fast ? 1.f / fov : 0.f;
  auto by_fov = [&](float n) { return fov_fast ? div_rcp(n, fov, ifov) : n / fov; };
  const f3 local_position = mk3(by_fov(((float)x - (float)W / 2.0f) * actual_depth),
                                by_fov(((float)y - (float)H / 2.0f) * actual_depth), actual_depth);
  const f3 cpos = mk3(a.cpos[0], a.cpos[1], a.cpos[2]);
  const f3 global_position = back_rot(local_position, a.cam) + cpos;
  const f3 object_local =
      back_rot_quat(global_position - mk3(G.wp[0], G.wp[1], G.wp[2]), f4{G.q[0], G.q[1], G.q[2], G.q[3]});
  const f3 last = rot_quat(object_local, old_q) + old_pos;
  const f3 last_nc = depth_project(rot(last, cpos, a.cam), (float)W, (float)H, fov);
  const f3 last_p = depth_project(rot(last, mk3(a.cpos_old[0], a.cpos_old[1], a.cpos_old[2]), a.cam_old), (float)W,
                                  (float)H, fov);
  if (last_p.z < (float)kIcutoff) {
    // the linear sample at the pixel's centre is the pixel itself
    a.rgba[pix] = a.in[pix];
    return;
  }
  const float fx = (float)x, fy = (float)y;
  const float cc = a.camera_contribution, ncc = 1.f - cc;
  float vx = cc * (fx - last_p.x) + ncc * (fx - last_nc.x);
  float vy = cc * (fy - last_p.y) + ncc * (fy - last_nc.y);
  vx = vx * a.strength;
  vy = vy * a.strength;
  int n = to_int_sat(fmaxf(fabsf(vx), fabsf(vy)) + 1.f);
  constexpr int kBound = 50;
  if (n > kBound) {
    vx = (vx / (float)n) * (float)kBound;
    vy = (vy / (float)n) * (float)kBound;
    n = kBound;
  }
  float dx = 0.f, dy = 0.f;
  if (n != 0) {
    dx = vx / (float)n;
    dy = vy / (float)n;
  }
  float cx = fx - vx / 2.f, cy = fy - vy / 2.f;
  f4 acc = f4{0.f, 0.f, 0.f, 0.f};
  float count = 0.f;
  for (int i = 0; i < n; i++, cx += dx, cy += dy) {
    if (cx < 0.f || cx >= (float)W || cy < 0.f || cy >= (float)H) continue;
    const f4 c = sample_linear(a.in, W, H, cx + 0.5f, cy + 0.5f);
    acc = f4{acc.x + c.x, acc.y + c.y, acc.z + c.z, acc.w + c.w};
    count += 1.f;
  }
  if (count != 0.f) acc = f4{acc.x / count, acc.y / count, acc.z / count, acc.w / count};
  a.rgba[pix] = quant8(acc.x) | (quant8(acc.y) << 8) | (quant8(acc.z) << 16) | (quant8(acc.w) << 24);
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

// engine::do_motion_blur (engine.cpp:1518-1536): the frame's last render of
// `s` seen with `view`, blurred along each pixel's motion since the previous
// camera (c_pos_old, c_rot_old) and its object's last recorded pose.
extern "C" int ocr_frame_motion_blur(ocr_frame* f, ocr_scene* s, const ocr_view* view, const float c_pos_old[4],
                                     const float c_rot_old[4], uint32_t frame_id, float strength,
                                     float camera_contribution, void* stream) {
  if (!f || !s || !view || !c_pos_old || !c_rot_old) return fail(OCR_ERR_INVALID, "ocr_frame_motion_blur: bad argument");
  if (f->ctx != s->ctx) return fail(OCR_ERR_INVALID, "ocr_frame_motion_blur: frame and scene of different contexts");
  if (!f->rendered) return fail(OCR_ERR_INVALID, "ocr_frame_motion_blur: render the frame with ocr_render first");
  if (view->width != f->width || view->height != f->height)
    return fail(OCR_ERR_INVALID, "ocr_frame_motion_blur: view size differs from the frame's");
  const float fov = view->fov_const;
  if (!(fov >= 0x1p-60f && fov <= 0x1p60f))
    return fail(OCR_ERR_INVALID, "ocr_frame_motion_blur: fov_const must be in [2^-60, 2^60]");
  OCR_HIP_CHECK(hipSetDevice(f->ctx->device));
  const size_t npix = (size_t)f->width * f->height;
  if (!f->post_in) {
    hipError_t e = hipMalloc((void**)&f->post_in, 4 * npix);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      f->post_in = nullptr;
      return fail(OCR_ERR_OOM, std::string("ocr_frame_motion_blur: ") + hipGetErrorString(e));
    }
  }
  hipStream_t st = (hipStream_t)stream;
  // reads the scene's current descriptors (the version of its last
  // stream-ordered update), registered as their reader below
  std::lock_guard<std::mutex> lk(s->ver_mu);
  OCR_HIP_CHECK(s->wait_geometry(st));
  OCR_HIP_CHECK(hipMemcpyAsync(f->post_in, f->rgba, 4 * npix, hipMemcpyDeviceToDevice, st));
  MotionArgs a;
  memset(&a, 0, sizeof(a));
  a.depth = f->depth;
  a.idkey = f->idkey;
  a.cuts = f->out[f->cur].cuts;
  a.counters = f->out[f->cur].counters;
  a.objs = s->objs;
  a.motion = s->motion;
  a.nobj = s->nobj;
  a.in = f->post_in;
  a.rgba = f->rgba;
  a.W = f->width;
  a.H = f->height;
  a.fov = fov;
  for (int k = 0; k < 3; k++) {
    a.cpos[k] = view->c_pos[k];
    a.cpos_old[k] = c_pos_old[k];
  }
  sincos_of(view->c_rot, a.cam);
  sincos_of(c_rot_old, a.cam_old);
  a.frame_id = frame_id;
  a.strength = strength;
  a.camera_contribution = camera_contribution;
  hipLaunchKernelGGL(k_motion_blur, dim3((a.W + 15) / 16, (a.H + 15) / 16), dim3(256), 0, st, a);
  OCR_HIP_CHECK(hipGetLastError());
  OCR_HIP_CHECK(s->note_read(st));
  return OCR_OK;
}
