// raster.hip — prearrange (transform, near clip, cull, fragment count), the
// depth pass (u32 atomicMin) and the id pass (u64 atomicMax on a packed key),
// for camera views and for cube shadow maps.
//
// Reference: prearrange cl2.cl:4271-4409, kernel1 cl2.cl:4985-5127,
// kernel2 cl2.cl:5390-5546, prearrange_realtime_shadowing cl2.cl:4419-4636,
// kernel1_realtime_shadowing cl2.cl:5129-5246.
//
// Work lists.  prearrange appends each surviving cut triangle (48 B CutTri)
// and its ceil(bbox_area / op_size) fragment records (8 B: cut index, chunk)
// to two device lists.  Offsets come from a wavefront-wide exclusive scan of
// the per-lane counts plus ONE atomicAdd per wavefront and list (the reference
// does two global atomics per triangle, cl2.cl:4342, :4388).  Capacity is
// checked per wavefront; on overflow nothing of that wavefront is written and
// the overflow word is set, which every later pass of the frame checks before
// touching the lists.
#pragma clang fp contract(off)

#include "ocr_internal.h"

namespace ocr {

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint32_t t = __shfl_up(v, d, kWave);
    if (lane >= d) v += t;
  }
  return v;
}

// Allocate `mine` slots per lane from a list counter with one atomic per
// wavefront.  Returns this lane's first slot and the wavefront's range end.
__device__ __forceinline__ uint32_t wave_alloc(uint32_t* counter, uint32_t mine, uint32_t& wave_end) {
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t incl = wave_incl_scan(mine);
  uint32_t total = __shfl(incl, kWave - 1, kWave);
  uint32_t base = 0;
  if (lane == 0 && total) base = atomicAdd(counter, total);
  base = __shfl(base, 0, kWave);
  wave_end = base + total;
  return base + incl - mine;
}

// ---------------------------------------------------------- setup helpers
struct Setup {
  float mm[4];        // min_x, max_x, min_y, max_y
  int width;
  float A, B, C;      // 1/z plane, cl2.cl:5044
  float iwidth;
  float xp[3], yp[3]; // rounded vertex coordinates
};

__device__ __forceinline__ void raster_setup(const CutTri& c, float ew, float eh, Setup& s) {
  calc_min_max(c.x, c.y, ew, eh, s.mm);
  s.width = (int)(s.mm[1] - s.mm[0]);
#pragma unroll
  for (int i = 0; i < 3; i++) { s.xp[i] = roundf(c.x[i]); s.yp[i] = roundf(c.y[i]); }
  f3 depths = mk3(1.0f / (c.z[0] / kDepthFar), 1.0f / (c.z[1] / kDepthFar), 1.0f / (c.z[2] / kDepthFar));
  plane_const(depths, mk3(s.xp[0], s.xp[1], s.xp[2]), mk3(s.yp[0], s.yp[1], s.yp[2]), c.rconst, s.A, s.B, s.C);
  s.iwidth = 1.f / (float)s.width;
}

// The fragment walk of kernel1 / kernel2 (cl2.cl:5046-5126): 501 steps of a
// running (x, y) cursor from pixel op_size*chunk of the bounding box, with
// the reference's fma row computation and running-modulo column.
template <typename Visit>
__device__ __forceinline__ void walk(const Setup& s, uint32_t chunk, int op_size, Visit&& visit) {
  const int width = s.width;
  if (width <= 0) return;
  const int pixel_along = op_size * (int)chunk;
  int pcount = -1;
  float x = (float)((pixel_along + 0) % width) + s.mm[0] - 1;
  float y = floorf((float)(pixel_along + pcount) / (float)width) + s.mm[2];
  float rwm = (float)((pixel_along + pcount) % width);
  while (pcount < op_size) {
    pcount++;
    x += 1;
    rwm += 1.f;
    if (rwm >= (float)width) rwm = 0;
    float ty = y;
    y = floorf(fmaf((float)(pixel_along + pcount), s.iwidth, s.mm[2]));
    x = y != ty ? rwm + s.mm[0] : x;
    if (y >= s.mm[3]) break;
    if (x >= s.mm[1]) continue;
    if (point_in_tri(x, y, s.xp[0], s.yp[0], s.xp[1], s.yp[1], s.xp[2], s.yp[2])) {
      float fd = fmaf(s.A, x, fmaf(s.B, y, s.C));
      visit((int)x, (int)y, sat_u32(kMulint / fd));
    }
  }
}

// Clip/cull/count of one projected triangle pair (cl2.cl:4344-4380).
__device__ __forceinline__ uint32_t cull_count(const f3 p[3], float ew, float eh, bool two_sided,
                                               int op_size, float& rconst) {
  bool valid = two_sided || front_facing(p[0], p[1], p[2]);
  bool cond = (p[0].x < 0 && p[1].x < 0 && p[2].x < 0) ||
              (p[0].x >= ew && p[1].x >= ew && p[2].x >= ew) ||
              (p[0].y < 0 && p[1].y < 0 && p[2].y < 0) ||
              (p[0].y >= eh && p[1].y >= eh && p[2].y >= eh);
  if (!valid || cond) return 0;
  f3 xpv = mk3(roundf(p[0].x), roundf(p[1].x), roundf(p[2].x));
  f3 ypv = mk3(roundf(p[0].y), roundf(p[1].y), roundf(p[2].y));
  rconst = rconstant(xpv, ypv);
  float xs[3] = {p[0].x, p[1].x, p[2].x}, ys[3] = {p[0].y, p[1].y, p[2].y};
  float mm[4];
  calc_min_max(xs, ys, ew, eh, mm);
  float area = (mm[1] - mm[0]) * (mm[3] - mm[2]);
  float n = ceilf(area / (float)op_size);
  return n > 0.f ? (uint32_t)n : 0u;
}

__device__ __forceinline__ void store_cut(CutTri* dst, const f3 p[3], float rconst, uint32_t key, uint32_t oid) {
  CutTri c;
#pragma unroll
  for (int i = 0; i < 3; i++) { c.x[i] = p[i].x; c.y[i] = p[i].y; c.z[i] = p[i].z; }
  c.rconst = rconst;
  c.key = key;
  c.oid = oid;
  *dst = c;
}

__device__ __forceinline__ f4 ld_quat(const ObjDesc& g) { return f4{g.q[0], g.q[1], g.q[2], g.q[3]}; }

// full_rotate_quat, cl2.cl:700-729.
__device__ __forceinline__ int transform_clip_project(const TriPos& T, const ObjDesc& G, f3 cpos,
                                                      const SinCos& sc, float fovc, float ew, float eh,
                                                      f3 out[2][3]) {
  f3 wp = mk3(G.wp[0], G.wp[1], G.wp[2]);
  f4 q = ld_quat(G);
  f3 pr[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    f3 v = mk3(T.p[3 * i], T.p[3 * i + 1], T.p[3 * i + 2]);
    pr[i] = rot(rot_quat(v * G.scale, q) + wp, cpos, sc);
  }
  f3 tris[2][3];
  int n = clip_near(pr, tris);
  for (int t = 0; t < n; t++)
#pragma unroll
    for (int i = 0; i < 3; i++) out[t][i] = depth_project(tris[t][i], ew, eh, fovc);
  return n;
}

// ------------------------------------------------------------------ clear
__global__ void __launch_bounds__(256) k_clear_frame(uint32_t* __restrict__ depth,
                                                     unsigned long long* __restrict__ idkey,
                                                     uint32_t* __restrict__ counters, size_t npix) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  if (i < kCounterWords) counters[i] = 0u;
  for (; i < npix; i += stride) {
    depth[i] = 0xFFFFFFFFu;
    idkey[i] = 0ull;
  }
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

static unsigned grid_for(size_t n, unsigned block, unsigned max_blocks) {
  size_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (unsigned)b;
}

hipError_t launch_clear_frame(Frame* f, hipStream_t s) {
  size_t npix = (size_t)f->width * f->height;
  hipLaunchKernelGGL(k_clear_frame, dim3(grid_for(npix, 256, 4096)), dim3(256), 0, s, f->depth,
                     f->idkey, f->counters, npix);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, p, value, n);
  return hipGetLastError();
}

// ------------------------------------------------------------- prearrange
__global__ void __launch_bounds__(256) k_prearrange(const TriPos* __restrict__ tris, uint32_t ntri,
                                                    const ObjDesc* __restrict__ objs, uint32_t nobj,
                                                    ViewParams v, CutTri* __restrict__ cuts,
                                                    uint32_t cap_cut, uint2* __restrict__ frags,
                                                    uint32_t cap_frag, uint32_t* __restrict__ counters) {
  const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
  const float ew = (float)v.width, eh = (float)v.height;
  f3 tp[2][3];
  float rc[2] = {0.f, 0.f};
  uint32_t nf[2] = {0u, 0u};
  uint32_t oid = 0;
  if (id < ntri) {
    const TriPos T = tris[id];
    oid = T.oid;
    if (oid < nobj) {
      const ObjDesc G = objs[oid];
      f3 cpos = mk3(v.cpos[0], v.cpos[1], v.cpos[2]);
      if (!(fast_length(mk3(G.wp[0], G.wp[1], G.wp[2]) - cpos) > kDepthFar)) {
        int num = transform_clip_project(T, G, cpos, v.cam, v.fov, ew, eh, tp);
        bool two_sided = (G.flags & OCR_FEATURE_TWO_SIDED) > 0;
        for (int i = 0; i < num; i++) nf[i] = cull_count(tp[i], ew, eh, two_sided, kOpSize, rc[i]);
      }
    }
  }
  // A cut triangle whose bbox has zero area emits no fragment and can never
  // write a pixel, so it is not stored either.
  const uint32_t k0 = nf[0] > 0, k1 = nf[1] > 0;
  uint32_t cut_end, frag_end;
  uint32_t cut0 = wave_alloc(&counters[kCutCount], k0 + k1, cut_end);
  uint32_t frag0 = wave_alloc(&counters[kFragCount], nf[0] + nf[1], frag_end);
  if (cut_end > cap_cut || frag_end > cap_frag) {
    if ((threadIdx.x & (kWave - 1)) == 0) atomicOr(&counters[kOverflow], 1u);
    return;
  }
  uint32_t c = cut0, f = frag0;
#pragma unroll
  for (int i = 0; i < 2; i++) {
    if (!nf[i]) continue;
    store_cut(&cuts[c], tp[i], rc[i], 2u * id + (uint32_t)i, oid);
    for (uint32_t a = 0; a < nf[i]; a++) frags[f + a] = make_uint2(c, a);
    f += nf[i];
    c++;
  }
}

hipError_t launch_prearrange(const Scene* sc, const ViewParams& v, Frame* f, hipStream_t s) {
  if (sc->ntri == 0) return hipSuccess;
  hipLaunchKernelGGL(k_prearrange, dim3((sc->ntri + 255) / 256), dim3(256), 0, s, sc->tri_pos,
                     sc->ntri, sc->objs, sc->nobj, v, f->cuts, f->cap_cut, f->frags, f->cap_frag,
                     f->counters);
  return hipGetLastError();
}

// ------------------------------------------------------------- depth pass
__global__ void __launch_bounds__(256) k_depth(const CutTri* __restrict__ cuts,
                                               const uint2* __restrict__ frags,
                                               const uint32_t* __restrict__ counters, uint32_t cap_frag,
                                               uint32_t W, uint32_t H, uint32_t* __restrict__ depth) {
  if (counters[kOverflow]) return;
  const uint32_t n = min(counters[kFragCount], cap_frag);
  const float ew = (float)W, eh = (float)H;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint2 fr = frags[i];
    Setup st;
    raster_setup(cuts[fr.x], ew, eh, st);
    walk(st, fr.y, kOpSize, [&](int x, int y, uint32_t d) {
      atomicMin(&depth[(size_t)y * W + (size_t)x], d);
    });
  }
}

hipError_t launch_depth(const ViewParams& v, Frame* f, hipStream_t s) {
  int ncu = f->ctx->prop.multiProcessorCount > 0 ? f->ctx->prop.multiProcessorCount : 256;
  hipLaunchKernelGGL(k_depth, dim3(ncu * 8), dim3(256), 0, s, f->cuts, f->frags, f->counters,
                     f->cap_frag, v.width, v.height, f->depth);
  return hipGetLastError();
}

// ---------------------------------------------------------------- id pass
// kernel2's "last writer wins" is made deterministic: among fragments inside
// the BUF_ERROR window the pixel keeps the largest (2*tri + sub) key, which
// is the fragment a sequential in-order execution writes last.
__global__ void __launch_bounds__(256) k_id(const CutTri* __restrict__ cuts, const uint2* __restrict__ frags,
                                            const uint32_t* __restrict__ counters, uint32_t cap_frag,
                                            uint32_t W, uint32_t H, const uint32_t* __restrict__ depth,
                                            unsigned long long* __restrict__ idkey) {
  if (counters[kOverflow]) return;
  const uint32_t n = min(counters[kFragCount], cap_frag);
  const float ew = (float)W, eh = (float)H;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint2 fr = frags[i];
    const CutTri c = cuts[fr.x];
    Setup st;
    raster_setup(c, ew, eh, st);
    const unsigned long long key = ((unsigned long long)(c.key + 1u) << 32) | fr.x;
    walk(st, fr.y, kOpSize, [&](int x, int y, uint32_t mydepth) {
      size_t p = (size_t)y * W + (size_t)x;
      uint32_t val = depth[p];
      if (mydepth > val - kBufError && mydepth < val + kBufError) atomicMax(&idkey[p], key);
    });
  }
}

hipError_t launch_id(const ViewParams& v, Frame* f, hipStream_t s) {
  int ncu = f->ctx->prop.multiProcessorCount > 0 ? f->ctx->prop.multiProcessorCount : 256;
  hipLaunchKernelGGL(k_id, dim3(ncu * 8), dim3(256), 0, s, f->cuts, f->frags, f->counters, f->cap_frag,
                     v.width, v.height, f->depth, f->idkey);
  return hipGetLastError();
}

// ---------------------------------------------------------------- shadows
struct FaceTable { SinCos f[6]; };

__global__ void __launch_bounds__(256) k_prearrange_shadow(const TriPos* __restrict__ tris, uint32_t ntri,
                                                           const ObjDesc* __restrict__ objs, uint32_t nobj,
                                                           float lx, float ly, float lz, uint32_t L,
                                                           int only_static, FaceTable ft,
                                                           CutTri* __restrict__ cuts, uint32_t cap_cut,
                                                           uint2* __restrict__ frags, uint32_t cap_frag,
                                                           uint32_t* __restrict__ counters) {
  const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
  const float ew = (float)L, eh = (float)L, efov = (float)L / 2.0f;
  const f3 lpos = mk3(lx, ly, lz);
  int touched = 0;
  TriPos T;
  ObjDesc G;
  bool live = false;
  if (id < ntri) {
    T = tris[id];
    if (T.oid < nobj) {
      G = objs[T.oid];
      bool is_static = (G.flags & OCR_FEATURE_IS_STATIC) > 0;
      live = only_static ? is_static : !is_static;
      f3 wp = mk3(G.wp[0], G.wp[1], G.wp[2]);
      if (live && fast_length(wp - lpos) > kDepthFar) live = false;
      if (live) {
        f4 q = ld_quat(G);
#pragma unroll
        for (int k = 0; k < 3; k++) {
          f3 r = rot_quat(mk3(T.p[3 * k], T.p[3 * k + 1], T.p[3 * k + 2]) * G.scale, q) + wp;
          touched |= 1 << cubeface(r, lpos);
        }
      }
    }
  }
  // Count pass over the touched faces, then allocate, then write.
  uint32_t ncut = 0, nfrag = 0;
  const bool two_sided = live && (G.flags & OCR_FEATURE_TWO_SIDED) > 0;
  for (int kk = 0; kk < 6; kk++) {
    if (!(touched >> kk & 1)) continue;
    f3 tp[2][3];
    int num = transform_clip_project(T, G, lpos, ft.f[kk], efov, ew, eh, tp);
    for (int i = 0; i < num; i++) {
      float rc;
      uint32_t n = cull_count(tp[i], ew, eh, two_sided, kOpSizeLight, rc);
      if (n) { ncut++; nfrag += n; }
    }
  }
  uint32_t cut_end, frag_end;
  uint32_t c = wave_alloc(&counters[kCutCount], ncut, cut_end);
  uint32_t f = wave_alloc(&counters[kFragCount], nfrag, frag_end);
  if (cut_end > cap_cut || frag_end > cap_frag) {
    if ((threadIdx.x & (kWave - 1)) == 0) atomicOr(&counters[kOverflow], 1u);
    return;
  }
  for (int kk = 0; kk < 6; kk++) {
    if (!(touched >> kk & 1)) continue;
    f3 tp[2][3];
    int num = transform_clip_project(T, G, lpos, ft.f[kk], efov, ew, eh, tp);
    for (int i = 0; i < num; i++) {
      float rc;
      uint32_t n = cull_count(tp[i], ew, eh, two_sided, kOpSizeLight, rc);
      if (!n) continue;
      store_cut(&cuts[c], tp[i], rc, (uint32_t)kk, T.oid);
      for (uint32_t a = 0; a < n; a++) frags[f + a] = make_uint2(c, a);
      f += n;
      c++;
    }
  }
}

__global__ void __launch_bounds__(256) k_depth_shadow(const CutTri* __restrict__ cuts,
                                                      const uint2* __restrict__ frags,
                                                      const uint32_t* __restrict__ counters,
                                                      uint32_t cap_frag, uint32_t L,
                                                      uint32_t* __restrict__ map) {
  if (counters[kOverflow]) return;
  const uint32_t n = min(counters[kFragCount], cap_frag);
  const float ew = (float)L;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint2 fr = frags[i];
    const CutTri c = cuts[fr.x];
    Setup st;
    raster_setup(c, ew, ew, st);
    uint32_t* face = map + (size_t)c.key * L * L;
    walk(st, fr.y, kOpSizeLight, [&](int x, int y, uint32_t d) {
      atomicMin(&face[(size_t)y * L + (size_t)x], d);
    });
  }
}

__global__ void k_zero_counters(uint32_t* c) {
  if (threadIdx.x < kCounterWords) c[threadIdx.x] = 0u;
}

hipError_t launch_shadow_light(Scene* sc, const float lpos[3], int only_static, uint32_t* map,
                               hipStream_t s) {
  FaceTable ft;
  for (int i = 0; i < 6; i++) ft.f[i] = sc->ctx->face_sc[i];
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, sc->sh_counters);
  if (sc->ntri)
    hipLaunchKernelGGL(k_prearrange_shadow, dim3((sc->ntri + 255) / 256), dim3(256), 0, s, sc->tri_pos,
                       sc->ntri, sc->objs, sc->nobj, lpos[0], lpos[1], lpos[2], sc->lbuf_dim, only_static,
                       ft, sc->sh_cuts, sc->sh_cap_cut, sc->sh_frags, sc->sh_cap_frag, sc->sh_counters);
  int ncu = sc->ctx->prop.multiProcessorCount > 0 ? sc->ctx->prop.multiProcessorCount : 256;
  hipLaunchKernelGGL(k_depth_shadow, dim3(ncu * 8), dim3(256), 0, s, sc->sh_cuts, sc->sh_frags,
                     sc->sh_counters, sc->sh_cap_frag, sc->lbuf_dim, map);
  return hipGetLastError();
}

}  // namespace ocr
