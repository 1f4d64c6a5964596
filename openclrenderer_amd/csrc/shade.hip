// shade.hip — deferred shading, one work item per pixel (kernel3,
// cl2.cl:5792-6408): un-project the quantised depth, object-space
// barycentrics, mip-mapped atlas fetch (texture_filter_diff), SSAO, point
// lights with cube-map shadows (generate_hard_occlusion) and Cook-Torrance
// specular; writes RGBA8 and the encoded screen normal.
//
// Pixels are processed in 256-pixel tiles, 16x16 by default (the reference's
// 16x16 work groups, engine.cpp:2018-2022); the tile shape is a tuning knob.
#pragma clang fp contract(off)

#include "ocr_internal.h"

namespace ocr {

// Pointers in the global address space: loads through them become
// global_load with a scalar base and a 32-bit vector offset (no per-lane
// 64-bit address arithmetic, no flat-address counters).
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

template <typename T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
  return (gptr<T>)(uintptr_t)p;
}

// Element `idx` with a 32-bit byte offset (every buffer the shading pass
// reads is far below 4 GiB), which the scalar-base addressing form takes.
template <typename T>
__device__ __forceinline__ T ldg(gptr<const T> base, uint32_t idx) {
  gptr<const char> p = (gptr<const char>)base + idx * (uint32_t)sizeof(T);
  if constexpr (sizeof(T) <= 8) {
    return *(gptr<const T>)p;
  } else if constexpr (sizeof(T) % 16 == 0 && alignof(T) >= 16) {
    // records: whole 16-byte words, assembled in registers
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    u32x4 w[sizeof(T) / 16];
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 16; i++) w[i] = ((gptr<const u32x4>)p)[i];
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
  } else {
    static_assert(sizeof(T) % 8 == 0, "record size");
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    u32x2 w[sizeof(T) / 8];
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 8; i++) w[i] = ((gptr<const u32x2>)p)[i];
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
  }
}

// What one shading block reads: its view, its scene and the frame's lists.
struct ShadeArgs {
  ViewParams v;
  gptr<const uint32_t> depth;
  gptr<const unsigned long long> idkey;
  gptr<const CutTri> cuts;
  gptr<const TriPos> tri_pos;
  gptr<const TriAttr> tri_attr;
  gptr<const ObjDesc> objs;
  gptr<const uint32_t> pages;     // RGBA8 texels as u32
  gptr<const uint32_t> nums;
  gptr<const uint32_t> sizes;
  uint32_t mip_start;
  gptr<const ocr_light> lights;
  gptr<const float> light_lin;
  uint32_t nlights;
  gptr<const uint32_t> shadow_maps;
  gptr<const uint32_t> static_maps;
  uint32_t L;
  gptr<const uint32_t> counters;
  gptr<uint32_t> rgba;
  gptr<uint32_t> normals;
  const SinCos* face;
  uint32_t* tex_touch;      // count mode: 1 bit per atlas texel
  uint32_t* shadow_touch;   // count mode: 1 bit per dynamic cube-map texel
  uint32_t* static_touch;   // count mode: 1 bit per static cube-map texel
};

// Kernel arguments of the shading launch.
struct ShadeLaunch {
  PassArgs p;
  SinCos face[6];
  uint32_t* tex_touch;
  uint32_t* shadow_touch;
  uint32_t* static_touch;
};

template <bool kCount>
__device__ __forceinline__ void touch(uint32_t* bits, size_t idx) {
  if constexpr (kCount) atomicOr(&bits[idx >> 5], 1u << (idx & 31));
}

// read_tex_array_all_precalculated, cl2.cl:820-851.
template <bool kCount>
__device__ __forceinline__ f4 read_tex_pre(const ShadeArgs& a, float cx, float cy, int which, int slice,
                                           float width) {
  const float imax_tex_size = 1.f / 2048;
  float ihnum = width * imax_tex_size;
  float tnumy = floorf((float)which * ihnum);
  // tnumy / ihnum: atlas tile sizes are powers of two 2^e (16 .. 2048, set
  // by the page plan), so ihnum = 2^(e-11) and the quotient is the exact
  // product by 2^(11-e), whose bits are (265 << 23) minus those of 2^e
  const float inv_ihnum = __int_as_float((265 << 23) - __float_as_int(width));
  float tnumx = (float)which - tnumy * inv_ihnum;
  cx = clampf(cx, 0.001f, width - 0.001f);
  cy = clampf(cy, 0.001f, width - 0.001f);
  int ix = (int)fmaf(tnumx, width, cx);
  int iy = (int)fmaf(tnumy, width, cy);
  const uint32_t t = (uint32_t)slice * (2048u * 2048u) + (uint32_t)iy * 2048u + (uint32_t)ix;
  touch<kCount>(a.tex_touch, t);
  uint32_t c = ldg(a.pages, t);
  return f4{(float)(c & 0xFF), (float)((c >> 8) & 0xFF), (float)((c >> 16) & 0xFF), (float)(c >> 24)};
}

// return_bilinear_col_all_precalculated, cl2.cl:1426-1455.
template <bool kCount>
__device__ __forceinline__ f4 bilinear_pre(const ShadeArgs& a, float mx, float my, int which, int slice,
                                           float width) {
  float px = floorf(mx), py = floorf(my);
  f4 c0 = read_tex_pre<kCount>(a, px, py, which, slice, width);
  f4 c1 = read_tex_pre<kCount>(a, px + 1, py, which, slice, width);
  f4 c2 = read_tex_pre<kCount>(a, px, py + 1, which, slice, width);
  f4 c3 = read_tex_pre<kCount>(a, px + 1, py + 1, which, slice, width);
  float ux = mx - px, uy = my - py;
  float bx = 1.f - ux, by = 1.f - uy;
  f4 r;
  r.x = fmaf(c0.x, bx, c1.x * ux) * by + fmaf(c2.x, bx, c3.x * ux) * uy;
  r.y = fmaf(c0.y, bx, c1.y * ux) * by + fmaf(c2.y, bx, c3.y * ux) * uy;
  r.z = fmaf(c0.z, bx, c1.z * ux) * by + fmaf(c2.z, bx, c3.z * ux) * uy;
  r.w = fmaf(c0.w, bx, c1.w * ux) * by + fmaf(c2.w, bx, c3.w * ux) * uy;
  return r;
}

// texture_mod, cl2.cl:1457-1468.
__device__ __forceinline__ float tmod(float v) {
  v = v >= 1 ? 1.0f - (v - floorf(v)) : v;
  v = v < 0 ? 1.0f + fabsf(v) - fabsf(floorf(v)) : v;
  return v;
}

// texture_filter_diff, cl2.cl:1511-1573.
template <bool kCount>
__device__ __forceinline__ f4 texture_filter_diff(const ShadeArgs& a, float vtx, float vty, float dx,
                                                  float dy, int tid2) {
  int nv = (int)ldg(a.nums, (uint32_t)tid2);
  int tsize = (int)ldg(a.sizes, (uint32_t)(nv >> 16));
  float vmx = tmod(vtx), vmy = tmod(vty);
  float tx = dx * (float)tsize, ty = dy * (float)tsize;
  float worst = sqrt_cr(tx * tx + ty * ty);
  float wf = fmaxf(log2_approx(worst), 0.f);
  float mip_lower = clampf(floorf(wf), 0.f, (float)kMipLevels);
  float fmd = wf - mip_lower;
  int tid_lower = mip_lower == 0 ? tid2 : (int)(mip_lower - 1 + (float)a.mip_start + (float)(tid2 * kMipLevels));
  int tid_higher = (int)(clampf(mip_lower, 0.f, kMipLevels - 1.f) + (float)a.mip_start + (float)(tid2 * kMipLevels));
  int lnv = (int)ldg(a.nums, (uint32_t)tid_lower);
  float sl = (float)ldg(a.sizes, (uint32_t)(lnv >> 16));
  f4 c1 = bilinear_pre<kCount>(a, vmx * sl, vmy * sl, lnv & 0xFFFF, lnv >> 16, sl);
  const float i255 = 1.f / 255.f;
  // From mip level 4 on both fetches read the last level: c2 = c1, the
  // blend c1 + 0 * fmd is c1 (fmd is finite), and a wavefront whose lanes
  // are all there skips the second fetch
  if (__all(tid_lower == tid_higher)) return f4{c1.x * i255, c1.y * i255, c1.z * i255, c1.w * i255};
  int hnv = (int)ldg(a.nums, (uint32_t)tid_higher);
  float sh = (float)ldg(a.sizes, (uint32_t)(hnv >> 16));
  f4 c2 = bilinear_pre<kCount>(a, vmx * sh, vmy * sh, hnv & 0xFFFF, hnv >> 16, sh);
  return f4{(c1.x + (c2.x - c1.x) * fmd) * i255, (c1.y + (c2.y - c1.y) * fmd) * i255,
            (c1.z + (c2.z - c1.z) * fmd) * i255, (c1.w + (c2.w - c1.w) * fmd) * i255};
}

// generate_ssao, cl2.cl:2194-2260.  The 5x5 taps sit at offsets
// round(k * world_rad), k = -2..2 on each axis (clamped to the frame).  Where
// round(world_rad) is 0 the offsets are (-m, 0, 0, 0, m), m = round(2 *
// world_rad), so the 25 taps read at most 9 distinct depths, each counted
// 1, 3 or 1 times per axis; where m is 0 as well they all read one depth.
// A wavefront whose every lane is in such a case counts the distinct taps
// with their multiplicities (the same sum of the same per-tap counts); the
// others run the 25 taps.  Far geometry (world_rad = (2 .. 2.5) * fov /
// depth below 0.5) takes the short forms.
struct SsaoTaps {
  float thr[5];   // depth + z, z = -2..2, as the reference adds them
  const ShadeArgs* a;
  // passing depth tests of the tap at (wx, wy): thr[] is non-decreasing
  // (rounding is monotone), so the count comes from the middle threshold and
  // the pair on its side
  __device__ __forceinline__ uint32_t at(float wx, float wy) const {
    const float d2 = depth_of(ldg(a->depth, (uint32_t)(int)wy * a->v.width + (uint32_t)(int)wx));
    const bool hi = d2 > thr[2];
    const float ta = hi ? thr[3] : thr[0], tb = hi ? thr[4] : thr[1];
    return (hi ? 3u : 0u) + (d2 > ta ? 1u : 0u) + (d2 > tb ? 1u : 0u);
  }
};

__device__ __forceinline__ float ssao(const ShadeArgs& a, int sx, int sy) {
  const uint32_t W = a.v.width, H = a.v.height;
  uint32_t seed2 = xorshift(wang_hash((uint32_t)sx + W * H * (uint32_t)sy));
  float foffset = (float)seed2 / kMulint;
  float depth = depth_of(ldg(a.depth, (uint32_t)sy * W + (uint32_t)sx));
  float rad = kSsaoRad;
  rad += foffset / 2.f;
  // (rad * fov) / depth.  Depths of drawn pixels lie in [20, 350000]; with
  // rad * fov in [2^-59, 2^62] the quotient by the correctly rounded
  // reciprocal is then the correctly rounded one.  A degenerate two-sided
  // triangle can leave depth 0 (its plane is NaN): a wavefront holding one
  // divides.
  float world_rad;
  if (__all(depth >= 1.f)) world_rad = div_rcp(rad * a.v.fov, depth, rcp_rn(depth));
  else world_rad = rad * a.v.fov / depth;
  float fsx = (float)sx, fsy = (float)sy;
  SsaoTaps T;
  T.a = &a;
#pragma unroll
  for (int z = 0; z < 5; z++) T.thr[z] = depth + (float)(z - 2);
  const float xmax = (float)W - 2.f, ymax = (float)H - 2.f;
  const float m1 = roundf(world_rad), m2 = roundf(2.f * world_rad);
  // the reference sums 1.0f per passing test; the count (<= 125) is exact
  // in binary32, so it is kept as an integer and converted once
  uint32_t count = 0;
  if (__all(m2 == 0.f)) {
    count = 25u * T.at(clampf(fsx, 1.f, xmax), clampf(fsy, 1.f, ymax));
  } else if (__all(m1 == 0.f)) {
    const float xs[3] = {clampf(fsx - m2, 1.f, xmax), clampf(fsx, 1.f, xmax), clampf(fsx + m2, 1.f, xmax)};
    const float ys[3] = {clampf(fsy - m2, 1.f, ymax), clampf(fsy, 1.f, ymax), clampf(fsy + m2, 1.f, ymax)};
    const uint32_t mult[3] = {1u, 3u, 1u};
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) count += mult[j] * mult[i] * T.at(xs[i], ys[j]);
  } else {
#pragma unroll
    for (int y = -2; y <= 2; y++) {
      float oy = roundf((float)y * world_rad);
      float wy = clampf(fsy + oy, 1.f, ymax);
#pragma unroll
      for (int x = -2; x <= 2; x++) {
        float ox = roundf((float)x * world_rad);
        count += T.at(clampf(fsx + ox, 1.f, xmax), wy);
      }
    }
  }
  float acc = (float)count;
  acc = div_rcp(acc, 125.f, kInv125);
  return 1.f - div_rcp(1.f - acc, kSsaoDiv, kInvSsaoDiv);
}

// bilinear_interpolate, cl2.cl:2286-2302.
__device__ __forceinline__ float bilerp(float cx, float cy, float v0, float v1, float v2, float v3) {
  float mx = cx - 0.5f, my = cy - 0.5f;
  float ux = mx - floorf(mx), uy = my - floorf(my);
  float bx = 1.0f - ux, by = 1.0f - uy;
  return (v0 * bx + v1 * ux) * by + (v2 * bx + v3 * ux) * uy;
}

// generate_hard_occlusion, cl2.cl:2536-2701 (SMOOTH_SHADOWS), in two
// parts: where the pixel falls in the light's cube face and the depth bias
// (the same for the static and the dynamic map of a light), then the taps.
struct ShadowSetup {
  float px, py;      // clamped position on the face
  float dpth;        // the pixel's depth from the light
  float bias;
  uint32_t face_off; // (iy - 1) * L + (ix - 1) within the face
  int face;
};

// float(pi/4) rounded up: for every binary32 a in [this, rational_acos(0.05)]
// tan_f(a) >= 1, so the bias saturates at SHADOWBIAS without evaluating tan.
constexpr float kTanOneFrom = 0x1.921fb6p-1f;

__device__ __forceinline__ ShadowSetup shadow_setup(const ShadeArgs& a, f3 lpos, f3 normal, f3 ptl_n, f3 gpos) {
  ShadowSetup s;
  s.face = cubeface(gpos, lpos);
  const uint32_t L = a.L;
  f3 local_pos = rot(gpos, lpos, a.face[s.face]);
  float lf = (float)L;
  f3 post = depth_project(local_pos, lf, lf, lf / 2.0f);
  s.dpth = post.z;
  s.px = clampf(post.x, 3.f, lf - 4.f);
  s.py = clampf(post.y, 3.f, lf - 4.f);
  const int ix = (int)s.px, iy = (int)s.py;
  s.face_off = (uint32_t)(iy - 1) * L + (uint32_t)(ix - 1);
  float acos_res = rational_acos(clampf(dot(normal, ptl_n), 0.05f, 0.95f));
  if (acos_res >= kTanOneFrom) {
    s.bias = (float)kShadowBias;
  } else {
    float bias = kShadowBias * tan_f(acos_res);
    s.bias = clampf(bias, 0.1f * kShadowBias, (float)kShadowBias);
  }
  return s;
}

// Returns the shadow term (0 = lit).  A 4x4 block of depth compares feeds 9
// bilinear filters of 0/1 values; when all 16 compares agree the filters are
// evaluated with those constants (all zero: the sum is exactly 0).
template <bool kCount>
__device__ __forceinline__ float shadow_taps(const ShadeArgs& a, const ShadowSetup& s,
                                             gptr<const uint32_t> buf, uint32_t* touch_bits, int shnum) {
  const uint32_t L = a.L;
  const uint32_t map_off = ((uint32_t)s.face + (uint32_t)shnum * 6u) * L * L + s.face_off;
  uint32_t cond = 0;   // 4x4 taps, bit (y+1)*4 + (x+1)
#pragma unroll
  for (int y = 0; y < 4; y++)
#pragma unroll
    for (int x = 0; x < 4; x++) {
      const uint32_t t = map_off + (uint32_t)y * L + (uint32_t)x;
      touch<kCount>(touch_bits, t);
      float ldp = depth_of(ldg(buf, t));
      cond |= (s.dpth > ldp + s.bias ? 1u : 0u) << (y * 4 + x);
    }
  if (cond == 0u) return 0.f;
  float shadow = 0.f;
  if (cond == 0xFFFFu) {
#pragma unroll
    for (int y = -1; y <= 1; y++)
#pragma unroll
      for (int x = -1; x <= 1; x++)
        shadow += bilerp(s.px + 0.5f + (float)x, s.py + 0.5f + (float)y, 1.f, 1.f, 1.f, 1.f);
  } else {
#pragma unroll
    for (int y = -1; y <= 1; y++)
#pragma unroll
      for (int x = -1; x <= 1; x++) {
        float v0 = (float)((cond >> ((y + 1) * 4 + x + 1)) & 1u);
        float v1 = (float)((cond >> ((y + 1) * 4 + x + 2)) & 1u);
        float v2 = (float)((cond >> ((y + 2) * 4 + x + 1)) & 1u);
        float v3 = (float)((cond >> ((y + 2) * 4 + x + 2)) & 1u);
        shadow += bilerp(s.px + 0.5f + (float)x, s.py + 0.5f + (float)y, v0, v1, v2, v3);
      }
  }
  return div_rcp(shadow, 9.f, kInv9);
}

__device__ __forceinline__ f4 vcol(uint32_t c) {
  return f4{unorm8(c >> 24), unorm8((c >> 16) & 0xFF), unorm8((c >> 8) & 0xFF), unorm8(c & 0xFF)};
}

// The view and scene records reach the kernel through a reference that may
// point into a device table, so the compiler cannot see that their fields
// are the same for the whole wavefront.  Moving them into scalar registers
// lets light records come through the scalar cache and every buffer access
// use a scalar base with a 32-bit vector offset.
template <typename T>
__device__ __forceinline__ gptr<T> rfl(gptr<T> p) {
  const uint64_t v = (uint64_t)(uintptr_t)p;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return (gptr<T>)(uintptr_t)((uint64_t)hi << 32 | lo);
}
__device__ __forceinline__ uint32_t rfl(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ float rfl(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ void uniform(ShadeArgs& a) {
  a.v.width = rfl(a.v.width);
  a.v.height = rfl(a.v.height);
  a.v.fov = rfl(a.v.fov);
  a.v.ifov = rfl(a.v.ifov);
  a.v.linear = rfl(a.v.linear);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    a.v.cpos[i] = rfl(a.v.cpos[i]);
    a.v.cam.s[i] = rfl(a.v.cam.s[i]);
    a.v.cam.c[i] = rfl(a.v.cam.c[i]);
  }
#pragma unroll
  for (int i = 0; i < 4; i++) a.v.clear[i] = rfl(a.v.clear[i]);
  a.depth = rfl(a.depth);
  a.idkey = rfl(a.idkey);
  a.cuts = rfl(a.cuts);
  a.tri_pos = rfl(a.tri_pos);
  a.tri_attr = rfl(a.tri_attr);
  a.objs = rfl(a.objs);
  a.pages = rfl(a.pages);
  a.nums = rfl(a.nums);
  a.sizes = rfl(a.sizes);
  a.mip_start = rfl(a.mip_start);
  a.lights = rfl(a.lights);
  a.light_lin = rfl(a.light_lin);
  a.nlights = rfl(a.nlights);
  a.shadow_maps = rfl(a.shadow_maps);
  a.static_maps = rfl(a.static_maps);
  a.L = rfl(a.L);
  a.counters = rfl(a.counters);
  a.rgba = rfl(a.rgba);
  a.normals = rfl(a.normals);
}

// The shading tile of this block.  Workgroups go to the eight XCDs in
// turn (block b to XCD b mod 8), each XCD with an L2 of its own.  Groups of
// kXcdRun consecutive tiles (a strip of a tile row) go to one XCD, the
// groups to the XCDs in turn: neighbouring tiles of a strip share the depth
// and id lines that SSAO's 5x5 taps and the gathers read in one L2, while
// the load of busy screen regions stays spread over all eight.  A bijection
// of [0, n) whatever the dispatch order.
constexpr uint32_t kXcdRun = 8;
__device__ __forceinline__ uint32_t shade_tile_index() {
  const uint32_t b = blockIdx.x, n = gridDim.x;
  const uint32_t full = n / (8u * kXcdRun) * (8u * kXcdRun);   // blocks in whole rounds
  if (b >= full) return b;
  const uint32_t k = b & 7u, j = b >> 3;
  return ((j / kXcdRun) * 8u + k) * kXcdRun + j % kXcdRun;
}

__device__ __forceinline__ const DevView& shade_view(const PassArgs& p, uint32_t tile, uint32_t& vi) {
  vi = 0;
  if (p.views) {
    uint32_t lo = 0, hi = p.nviews;
    while (hi - lo > 1) {
      uint32_t mid = (lo + hi) >> 1;
      if (p.views[mid].tile_begin <= tile) lo = mid; else hi = mid;
    }
    vi = (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
    return p.views[vi];
  }
  return p.view0;
}

// kShadows = false: no light of the pass casts shadows (every light.shadow
// is 0), so the occlusion code is compiled out.
template <bool kCount, bool kShadows>
__device__ __forceinline__ void shade_pixel(const ShadeLaunch& sl) {
  uint32_t vi;
  const uint32_t block_tile = shade_tile_index();
  const DevView& V = shade_view(sl.p, block_tile, vi);
  const DevScene& S = sl.p.scenes ? sl.p.scenes[V.scene] : sl.p.scene0;
  ShadeArgs a;
  a.v = V.v;
  a.depth = to_global<const uint32_t>(V.depth);
  a.idkey = to_global<const unsigned long long>(V.idkey);
  a.cuts = to_global<const CutTri>(sl.p.lists.cuts);
  a.tri_pos = to_global<const TriPos>(S.tri_pos);
  a.tri_attr = to_global<const TriAttr>(S.tri_attr);
  a.objs = to_global<const ObjDesc>(S.objs);
  a.pages = to_global<const uint32_t>(reinterpret_cast<const uint32_t*>(S.pages));
  a.nums = to_global<const uint32_t>(S.nums);
  a.sizes = to_global<const uint32_t>(S.sizes);
  a.mip_start = S.mip_start;
  a.lights = to_global<const ocr_light>(S.lights);
  a.light_lin = to_global<const float>(S.light_lin);
  a.nlights = S.nlights;
  a.shadow_maps = to_global<const uint32_t>(S.shadow_maps);
  a.static_maps = to_global<const uint32_t>(S.static_maps);
  a.L = S.L;
  a.counters = to_global<const uint32_t>(sl.p.lists.counters);
  a.rgba = to_global<uint32_t>(V.rgba);
  a.normals = to_global<uint32_t>(V.normals);
  a.face = sl.face;
  a.tex_touch = sl.tex_touch;
  a.shadow_touch = sl.shadow_touch;
  a.static_touch = sl.static_touch;
  uniform(a);
  const uint32_t W = a.v.width, H = a.v.height;
  const uint32_t lw = sl.p.tile_w_log2, tw = 1u << lw, th = 256u >> lw;
  const uint32_t tiles_x = (W + tw - 1) / tw;
  const uint32_t tile = block_tile - V.tile_begin;
  const int x = (int)((tile % tiles_x) * tw + (threadIdx.x & (tw - 1)));
  const int y = (int)((tile / tiles_x) * th + (threadIdx.x >> lw));
  if (x >= (int)W || y >= (int)H) return;
  if (a.counters[kOverflow]) return;
  const uint32_t pix = (uint32_t)y * W + (uint32_t)x;
  const uint32_t ft = ldg(a.depth, pix);
  const unsigned long long key = ldg(a.idkey, pix);
  if (ft == 0xFFFFFFFFu || key == 0ull) {
    a.rgba[pix] = quant8(a.v.clear[0]) | (quant8(a.v.clear[1]) << 8) | (quant8(a.v.clear[2]) << 16) |
                  (quant8(a.v.clear[3]) << 24);
    a.normals[pix] = 0u;
    return;
  }
  // SSAO reads only the depth buffer: formed first, while little else is live
  const float occ_ssao = ssao(a, x, y);
  const float fov = a.v.fov;
  const CutTri ct = ldg(a.cuts, (uint32_t)key);
  const uint32_t tri = (uint32_t)(key >> 32) - 1u;
  const uint32_t tri_global = tri >> 1;
  const TriPos T = ldg(a.tri_pos, tri_global);
  const TriAttr TA = ldg(a.tri_attr, tri_global);
  const ObjDesc G = ldg(a.objs, ct.oid);
  const f3 camera_pos = mk3(a.v.cpos[0], a.v.cpos[1], a.v.cpos[2]);
  const f3 p1 = mk3(T.p[0], T.p[1], T.p[2]) * G.scale;
  const f3 p2 = mk3(T.p[3], T.p[4], T.p[5]) * G.scale;
  const f3 p3 = mk3(T.p[6], T.p[7], T.p[8]) * G.scale;
  const f3 wpos = mk3(G.wp[0], G.wp[1], G.wp[2]);
  // the object's quaternion as rot_quat normalises it, and its inverse as
  // back_rot_quat forms and rot_quat then normalises it (formed once per
  // descriptor upload with the same operations)
  const f4 qn = f4{G.qn[0], G.qn[1], G.qn[2], G.qn[3]};
  const f4 qb = f4{G.qb[0], G.qb[1], G.qb[2], G.qb[3]};

  const float actual_depth = depth_of(ft);
  // quotients by the view's fov from its correctly rounded reciprocal (the
  // view's fov is checked to lie in [2^-60, 2^60] on the host)
  const float ifov = a.v.ifov;
  auto by_fov = [&](float n) { return div_rcp(n, fov, ifov); };
  f3 local_position = mk3(by_fov(((float)x - (float)W / 2.0f) * actual_depth),
                          by_fov(((float)y - (float)H / 2.0f) * actual_depth), actual_depth);
  f3 global_position = back_rot(local_position, a.v.cam) + camera_pos;
  f3 object_local = rot_quat_n(global_position - wpos, qb);
  float l1, l2, l3;
  const BaryPre BP = bary_pre(p1, p2, p3);
  bary_eval(BP, object_local, l1, l2, l3);
  const float vtx = fmaf(TA.vt[0], l1, fmaf(TA.vt[2], l2, TA.vt[4] * l3));
  const float vty = fmaf(TA.vt[1], l1, fmaf(TA.vt[3], l2, TA.vt[5] * l3));
  f3 normal = mk3(fmaf(TA.n[0], l1, fmaf(TA.n[3], l2, TA.n[6] * l3)),
                  fmaf(TA.n[1], l1, fmaf(TA.n[4], l2, TA.n[7] * l3)),
                  fmaf(TA.n[2], l1, fmaf(TA.n[5], l2, TA.n[8] * l3)));
  normal = rot_quat_n(normal, qn);

  // get_vtdiff, cl2.cl:5691-5760.
  float vdx, vdy;
  {
    const float fx = (float)x, fy = (float)y;
    // 1/z plane of the triangle, built once by prearrange
    const float DA = ct.da, DB = ct.db, DC = ct.dc;
    float dmx = fmaf(DA, fx + 1, fmaf(DB, fy, DC));
    float dmy = fmaf(DA, fx, fmaf(DB, fy + 1, DC));
    // (n / fov) / d for both components, and 1 / d.  The numerators are
    // +0 or pixel offsets of at least 1/2 over fov, so where d lies in
    // [2^-60, 2^60] the quotients by the correctly rounded reciprocal are
    // the correctly rounded ones; other lanes (a plane extrapolated through
    // 0) divide, in a branch skipped when no lane needs it.
    const float n0x = by_fov(fx + 1 - (float)W / 2.f), n1x = by_fov(fy - (float)H / 2.f);
    const float n0y = by_fov(fx - (float)W / 2.f), n1y = by_fov(fy + 1 - (float)H / 2.f);
    const float yx = rcp_rn(dmx), yy = rcp_rn(dmy);
    f3 lmx = mk3(div_rcp(n0x, dmx, yx), div_rcp(n1x, dmx, yx), yx);
    f3 lmy = mk3(div_rcp(n0y, dmy, yy), div_rcp(n1y, dmy, yy), yy);
    if (!(rcp_safe(dmx) && rcp_safe(dmy))) {
      lmx = mk3(n0x / dmx, n1x / dmx, 1.f / dmx);
      lmy = mk3(n0y / dmy, n1y / dmy, 1.f / dmy);
    }
    f3 gmx = rot_quat_n((back_rot(lmx, a.v.cam) + camera_pos) - wpos, qb);
    f3 gmy = rot_quat_n((back_rot(lmy, a.v.cam) + camera_pos) - wpos, qb);
    float lx1, lx2, lx3, ly1, ly2, ly3;
    bary_eval(BP, gmx, lx1, lx2, lx3);
    bary_eval(BP, gmy, ly1, ly2, ly3);
    float vtxx = fmaf(TA.vt[0], lx1, fmaf(TA.vt[2], lx2, TA.vt[4] * lx3));
    float vtxy = fmaf(TA.vt[1], lx1, fmaf(TA.vt[3], lx2, TA.vt[5] * lx3));
    float vtyx = fmaf(TA.vt[0], ly1, fmaf(TA.vt[2], ly2, TA.vt[4] * ly3));
    float vtyy = fmaf(TA.vt[1], ly1, fmaf(TA.vt[3], ly2, TA.vt[5] * ly3));
    const float mip_bias = 1.f / kMipBias;
    vdx = (fabsf(vtxx - vtx) + fabsf(vtyx - vtx)) * mip_bias;
    vdy = (fabsf(vtxy - vty) + fabsf(vtyy - vty)) * mip_bias;
  }

  f4 col;
  if (TA.col[0] == 0u) {
    col = texture_filter_diff<kCount>(a, vtx, vty, vdx, vdy, (int)G.tid);
  } else {
    f4 c0 = vcol(TA.col[0]), c1 = vcol(TA.col[1]), c2 = vcol(TA.col[2]);
    col = f4{fmaf(c0.x, l1, fmaf(c1.x, l2, c2.x * l3)), fmaf(c0.y, l1, fmaf(c1.y, l2, c2.y * l3)),
             fmaf(c0.z, l1, fmaf(c1.z, l2, c2.z * l3)), fmaf(c0.w, l1, fmaf(c1.w, l2, c2.w * l3))};
  }
  const bool linear = a.v.linear != 0;
  if (linear) { col.x = gamma_fwd(col.x); col.y = gamma_fwd(col.y); col.z = gamma_fwd(col.z); }

  uint32_t seed2 = xorshift(wang_hash((uint32_t)x + (uint32_t)y * W * H));
  uint32_t seed3 = xorshift(seed2);
  uint32_t seed4 = xorshift(seed3);
  f3 rseed = mk3(((float)seed2 / kMulint - 0.5f) * 2, ((float)seed3 / kMulint - 0.5f) * 2,
                 ((float)seed4 / kMulint - 0.5f) * 2);

  const f3 l2p = fast_normalize(camera_pos - global_position);
  const int ff = G.flags;
  const bool two_sided = (ff & OCR_FEATURE_TWO_SIDED) > 0;
  const bool recv_dyn = !((ff & OCR_FEATURE_DOES_NOT_RECEIVE_DYNAMIC_SHADOWS) > 0);
  const bool is_front = front_facing(mk3(ct.x[0], ct.y[0], ct.z[0]), mk3(ct.x[1], ct.y[1], ct.z[1]),
                                     mk3(ct.x[2], ct.y[2], ct.z[2]));
  if (!is_front && two_sided) normal = -normal;
  normal = fast_normalize(normal);
  // rseed / 100: each component is 0 or of magnitude >= 2^-25, where the
  // short division by 100 is exact
  const f3 lighting_normal = fast_normalize(
      normal + mk3(div_rcp(rseed.x, 100.f, kInv100), div_rcp(rseed.y, 100.f, kInv100), div_rcp(rseed.z, 100.f, kInv100)));
  const float ambient = linear ? gamma_fwd(kAmbient) : kAmbient;

  int shnum = 0, static_num = 0;
  f3 diffuse_sum = mk3(0, 0, 0), specular_sum = mk3(0, 0, 0);
  for (uint32_t i = 0; i < a.nlights; i++) {
    const ocr_light l = ldg(a.lights, i);
    const f3 lpos = mk3(l.pos[0], l.pos[1], l.pos[2]);
    f3 ptl = lpos - global_position;
    float occlusion = 1;
    // the static and the dynamic map of a light share the face, position
    // and bias; both use the normalised point-to-light vector shading uses
    const bool stat = kShadows && l.shadow && l.is_static;
    // fast_length(ptl) and fast_normalize(ptl) form the same dot product and
    // correctly rounded root: formed once
    float distance, inv_len;
    length_and_inverse(dot(ptl, ptl), distance, inv_len);
    f3 ptl_n;
    ShadowSetup ss;
    if (stat) {
      ptl_n = ptl * inv_len;
      ss = shadow_setup(a, lpos, normal, ptl_n, global_position);
      occlusion = 1.f - shadow_taps<kCount>(a, ss, a.static_maps, a.static_touch, static_num);
      static_num++;
    }
    // 1/radius from the light upload (0 when the radius is out of range)
    const float inv_r = rfl(ldg(a.light_lin, 4 * i + 3));
    float t = (inv_r != 0.f ? div_rcp(distance, l.radius, inv_r) : distance / l.radius) + 1;
    // brightness / t^2: with the brightness +0 or of magnitude in
    // [2^-60, 2^60] and t^2 in [1, 2^60] the quotient lies in [2^-120, 2^60],
    // where the one from the correctly rounded reciprocal is correctly
    // rounded; a wavefront with a lane outside divides
    const float t2 = t * t;
    float illumination;
    if (bright_ok(l.brightness) && __all(t2 >= 1.f && t2 <= 0x1p60f))
      illumination = div_rcp(l.brightness, t2, rcp_rn(t2));
    else
      illumination = l.brightness / t2;
    const float cutoff = 0.1f;
    illumination -= cutoff;
    illumination *= 1.f / (1.f - cutoff);
    if (illumination <= 0) continue;
    f3 light_col = mk3(l.col[0], l.col[1], l.col[2]);
    if (linear) light_col = mk3(ldg(a.light_lin, 4 * i), ldg(a.light_lin, 4 * i + 1), ldg(a.light_lin, 4 * i + 2));
    if (!stat) ptl_n = ptl * inv_len;
    if (kShadows && l.shadow && recv_dyn) {
      if (!stat) ss = shadow_setup(a, lpos, normal, ptl_n, global_position);
      float dyn = 1.f - shadow_taps<kCount>(a, ss, a.shadow_maps, a.shadow_touch, shnum);
      occlusion = fminf(occlusion, dyn);
      shnum++;
    }
    ptl = ptl_n;
    float light = dot(ptl, lighting_normal);
    light *= occlusion;
    light = fmaxf(light, 0.f);
    float diffuse = (1.0f - ambient) * light;
    diffuse_sum = diffuse_sum + light_col * ((diffuse + ambient) * l.diffuse * G.diffuse * illumination);
    const f3 Hv = fast_normalize(l2p + ptl);
    const f3 N = normal;
    const float kS = 0.4f;
    float ndh = fmaxf(0.f, dot(N, Hv));
    float ndv = fmaxf(0.f, dot(N, l2p));
    float vdh = fmaxf(0.f, dot(l2p, Hv));
    float ndl = fmaxf(0.f, dot(N, ptl));
    const float F0 = 0.4f;
    float fresnel = F0 + (1 - F0) * powr5(1.f - vdh);
    const float gauss_constant = 0.8346f;
    float alpha = rational_acos_unit(ndh);
    // -alpha^2 / rough^2, rough = clamp(1 - specular, 0.001, 10), with rough^2
    // in [1e-6, 100] and its correctly rounded reciprocal from the object's
    // descriptor: for |alpha^2| in [2^-100, 2.5] the quotient is the correctly
    // rounded one; below, both quotients are zero or below 2^-80, whose exp
    // is 1 either way
    float microfacet = gauss_constant * exp_f(div_rcp(-alpha * alpha, G.rough2, G.inv_rough2));
    float sv = 2 * ndh / vdh;
    float c1 = sv * ndv;
    float c2 = sv * ndl;
    float geometric = min3f(1.f, c1, c2);
    float spec = (fresnel * microfacet * geometric) / (kPi * ndv);
    specular_sum = specular_sum + (light_col * (spec * kS * illumination)) * G.spec_mult;
    specular_sum = mk3(fmaxf(specular_sum.x, 0.f), fmaxf(specular_sum.y, 0.f), fmaxf(specular_sum.z, 0.f));
    specular_sum = specular_sum * occlusion;
  }
  specular_sum = specular_sum * occ_ssao;
  diffuse_sum = diffuse_sum * occ_ssao;
  const float rsc = 0.7f;
  f3 colclamp = (mk3(col.x, col.y, col.z) + mk3(0, 0, 0)) + specular_sum * rsc;
  f3 fc = mk3(fmaf(colclamp.x, diffuse_sum.x, specular_sum.x * (1.f - rsc)),
              fmaf(colclamp.y, diffuse_sum.y, specular_sum.y * (1.f - rsc)),
              fmaf(colclamp.z, diffuse_sum.z, specular_sum.z * (1.f - rsc)));
  if (linear) gamma_inv3(fc.x, fc.y, fc.z);
  a.rgba[pix] = quant8(fc.x) | (quant8(fc.y) << 8) | (quant8(fc.z) << 16) | (quant8(col.w) << 24);
  a.normals[pix] = encode_normal(normal);
}

template <bool kCount>
__global__ void __launch_bounds__(256) k_shade(ShadeLaunch sl) {
  shade_pixel<kCount, true>(sl);
}

// Passes whose scenes have no shadow-casting light: without the occlusion
// code the kernel fits 96 VGPRs, five wavefronts per SIMD instead of the
// three of the general kernel's 139, which hides more of its gathers and
// leaves room for the next render's raster passes (1080p, 1 M triangles,
// 4 lights: shading alone 124 -> 108 us).  With shadows, even four
// wavefronts per SIMD spill inside the light loop and run slower.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_shade_unshadowed(ShadeLaunch sl) {
  shade_pixel<false, false>(sl);
}

// The linear-space light colours (gamma_fwd of each light's colour, the same
// for every pixel), once per light upload.
__global__ void k_light_prep(const ocr_light* __restrict__ lights, uint32_t n, float* __restrict__ lin) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  lin[4 * i] = gamma_fwd(lights[i].col[0]);
  lin[4 * i + 1] = gamma_fwd(lights[i].col[1]);
  lin[4 * i + 2] = gamma_fwd(lights[i].col[2]);
  // the correctly rounded 1/radius the shading pass divides by
  const float r = lights[i].radius;
  lin[4 * i + 3] = rcp_safe(r) ? 1.f / r : 0.f;
}

hipError_t launch_light_prep(const ocr_light* lights, uint32_t n, float* lin, hipStream_t s) {
  hipLaunchKernelGGL(k_light_prep, dim3((n + 63) / 64), dim3(64), 0, s, lights, n, lin);
  return hipGetLastError();
}

hipError_t launch_shade(const PassArgs& p, const SinCos face[6], hipStream_t s) {
  ShadeLaunch sl;
  sl.p = p;
  for (int i = 0; i < 6; i++) sl.face[i] = face[i];
  sl.tex_touch = sl.shadow_touch = sl.static_touch = nullptr;
  if (p.unshadowed)
    hipLaunchKernelGGL(k_shade_unshadowed, dim3(p.tiles > 0 ? p.tiles : 1), dim3(256), 0, s, sl);
  else
    hipLaunchKernelGGL(k_shade<false>, dim3(p.tiles > 0 ? p.tiles : 1), dim3(256), 0, s, sl);
  return hipGetLastError();
}

hipError_t launch_shade_count(const PassArgs& p, const SinCos face[6], uint32_t* tex_bits,
                              uint32_t* shadow_bits, uint32_t* static_bits, hipStream_t s) {
  ShadeLaunch sl;
  sl.p = p;
  for (int i = 0; i < 6; i++) sl.face[i] = face[i];
  sl.tex_touch = tex_bits;
  sl.shadow_touch = shadow_bits;
  sl.static_touch = static_bits;
  hipLaunchKernelGGL(k_shade<true>, dim3(p.tiles > 0 ? p.tiles : 1), dim3(256), 0, s, sl);
  return hipGetLastError();
}

}  // namespace ocr
