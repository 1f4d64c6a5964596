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

#include <string.h>

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
  int n = clip_near(pr[0], pr[1], pr[2], tris);
  // both slots are projected unconditionally (static indices keep the
  // arrays in registers); slot 1 is only read when n == 2
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int i = 0; i < 3; i++) out[t][i] = depth_project(tris[t][i], ew, eh, fovc);
  return n;
}

// ------------------------------------------------------------------ views
__device__ __forceinline__ const DevView& view_at(const PassArgs& a, uint32_t i) {
  return a.views ? a.views[i] : a.view0;
}
__device__ __forceinline__ const DevScene& scene_at(const PassArgs& a, uint32_t i) {
  return a.scenes ? a.scenes[i] : a.scene0;
}

// The view owning a 256-slot block: the last view whose range starts at or
// before `slot` (ranges are 256-aligned, so the answer is block-uniform).
enum class Range { Tri, Pix, Tile };
template <Range kR>
__device__ __forceinline__ uint32_t view_of(const PassArgs& a, uint32_t slot) {
  if (!a.views) return 0;
  uint32_t lo = 0, hi = a.nviews;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    const DevView& m = a.views[mid];
    uint32_t b = kR == Range::Tri ? m.tri_begin : (kR == Range::Pix ? m.pix_begin : m.tile_begin);
    if (b <= slot) lo = mid; else hi = mid;
  }
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
}

// ------------------------------------------------------------------ clear
__global__ void __launch_bounds__(256) k_clear(PassArgs a) {
  const uint32_t slot = blockIdx.x * 256u;
  if (blockIdx.x == 0 && threadIdx.x < kCounterWords) a.lists.counters[threadIdx.x] = 0u;
  const DevView& V = view_at(a, view_of<Range::Pix>(a, slot));
  const uint32_t i = slot - V.pix_begin + threadIdx.x;
  if (i >= V.v.width * V.v.height) return;
  V.depth[i] = 0xFFFFFFFFu;
  V.idkey[i] = 0ull;
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

hipError_t launch_clear(const PassArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_clear, dim3(a.pix_slots / 256 > 0 ? a.pix_slots / 256 : 1), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, p, value, n);
  return hipGetLastError();
}

// ------------------------------------------------------------- prearrange
__global__ void __launch_bounds__(256) k_prearrange(PassArgs a) {
  const uint32_t slot = blockIdx.x * 256u;
  const uint32_t vi = view_of<Range::Tri>(a, slot);
  const DevView& V = view_at(a, vi);
  const DevScene& S = scene_at(a, V.scene);
  const ViewParams& v = V.v;
  const uint32_t id = slot - V.tri_begin + threadIdx.x;
  const float ew = (float)v.width, eh = (float)v.height;
  f3 tp[2][3];
  float rc[2] = {0.f, 0.f};
  uint32_t nf[2] = {0u, 0u};
  uint32_t oid = 0;
  if (id < S.ntri) {
    const TriPos T = S.tri_pos[id];
    oid = T.oid;
    if (oid < S.nobj) {
      const ObjDesc G = S.objs[oid];
      f3 cpos = mk3(v.cpos[0], v.cpos[1], v.cpos[2]);
      if (!(fast_length(mk3(G.wp[0], G.wp[1], G.wp[2]) - cpos) > kDepthFar)) {
        int num = transform_clip_project(T, G, cpos, v.cam, v.fov, ew, eh, tp);
        bool two_sided = (G.flags & OCR_FEATURE_TWO_SIDED) > 0;
#pragma unroll
        for (int i = 0; i < 2; i++)
          if (i < num) nf[i] = cull_count(tp[i], ew, eh, two_sided, kOpSize, rc[i]);
      }
    }
  }
  // A cut triangle whose bbox has zero area emits no fragment and can never
  // write a pixel, so it is not stored either.
  const uint32_t k0 = nf[0] > 0, k1 = nf[1] > 0;
  uint32_t cut_end, frag_end;
  const Lists& L = a.lists;
  uint32_t cut0 = wave_alloc(&L.counters[kCutCount], k0 + k1, cut_end);
  uint32_t frag0 = wave_alloc(&L.counters[kFragCount], nf[0] + nf[1], frag_end);
  if (cut_end > L.cap_cut || frag_end > L.cap_frag) {
    if ((threadIdx.x & (kWave - 1)) == 0) atomicOr(&L.counters[kOverflow], 1u);
    return;
  }
  uint32_t c = cut0, f = frag0;
  const uint32_t vbits = vi << kViewShift;
#pragma unroll
  for (int i = 0; i < 2; i++) {
    if (!nf[i]) continue;
    store_cut(&L.cuts[c], tp[i], rc[i], 2u * id + (uint32_t)i, oid);
    for (uint32_t q = 0; q < nf[i]; q++) L.frags[f + q] = make_uint2(c, q | vbits);
    f += nf[i];
    c++;
  }
}

hipError_t launch_prearrange(const PassArgs& a, hipStream_t s) {
  if (a.tri_slots == 0) return hipSuccess;
  hipLaunchKernelGGL(k_prearrange, dim3(a.tri_slots / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

// ---------------------------------------------- load-balanced raster passes
//
// A fragment is the reference's unit of work: the samples q = pa .. pa + 500
// (pa = op_size * chunk) of its cut triangle's bounding box, visited by a
// running cursor (cl2.cl:5046-5126).  Fragment sizes range from 1 to 501
// samples, so one thread per fragment leaves most lanes of a wavefront idle.
// Here a wavefront loads 64 fragments, keeps their setup in LDS, prefix-sums
// their sample counts and sweeps all of their samples with all 64 lanes.
//
// Each sample is evaluated on its own with the cursor's closed form:
//   y(q) = floor(fma(q, iwidth, min_y))                  (the reference's row)
//   the walk stops at the first q with y(q) >= max_y      (y is monotone in q)
//   x(q) = (r % width) + min_x + (q - r)
// where r is the later of the fragment start pa and the first index of q's
// row run (the cursor is reset to running_width_mod + min_x exactly when y
// changes, and starts at pa % width + min_x).  Row-run starts are found from
// the nominal (Y - min_y) * width by stepping while y disagrees, so rows whose
// fma rounding lags or leads by a sample are reproduced exactly.

struct FragLds {
  float mm[4][kWave];      // min_x, max_x, min_y, max_y
  float pl[4][kWave];      // A, B, C (1/z plane), iwidth
  float ek[8][kWave];      // edge test: s0, s1, s2, t0, t1, t2, sign, 2.0001*A*sign
  int w[kWave];
  int pa[kWave];
  uint32_t pitch[kWave];   // row pitch of the target buffer
  uint32_t* dst[kWave];    // depth buffer / cube face
  unsigned long long* ids[kWave];
  unsigned long long key[kWave];
  uint32_t pre[kWave + 1]; // exclusive prefix of sample counts
};

__device__ __forceinline__ float row_of(int q, float iw, float min_y) {
  return floorf(fmaf((float)q, iw, min_y));
}

// First q >= 0 with row_of(q) >= Y.
__device__ __forceinline__ int run_start(float Y, int w, float iw, float min_y) {
  long long n0 = (long long)(Y - min_y) * (long long)w;
  int q = n0 < 0 ? 0 : (int)n0;
  for (int g = 0; g < 64 && q > 0 && row_of(q - 1, iw, min_y) >= Y; g++) q--;
  for (int g = 0; g < 64 && row_of(q, iw, min_y) < Y; g++) q++;
  return q;
}

enum class RasterMode { Depth, Id, Shadow };

struct ShadowTarget {
  uint32_t* map;   // 6 faces of L x L
  uint32_t L;
};

template <RasterMode kMode>
__global__ void __launch_bounds__(256) k_raster(PassArgs a, ShadowTarget sh, int op_size) {
  __shared__ FragLds lds_all[256 / kWave];
  const Lists& Ls = a.lists;
  if (Ls.counters[kOverflow]) return;
  const uint32_t n = min(Ls.counters[kFragCount], Ls.cap_frag);
  const int lane = threadIdx.x & (kWave - 1);
  FragLds& L = lds_all[threadIdx.x / kWave];
  const uint32_t wave0 = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  for (uint32_t base = wave0 * kWave; base < n; base += nwaves * kWave) {
    // ---- per-fragment setup, one lane per fragment
    const uint32_t fi = base + lane;
    uint32_t cnt = 0;
    if (fi < n) {
      const uint2 fr = Ls.frags[fi];
      const CutTri c = Ls.cuts[fr.x];
      uint32_t W, H;
      if constexpr (kMode == RasterMode::Shadow) {
        W = H = sh.L;
        L.dst[lane] = sh.map + (size_t)c.key * sh.L * sh.L;
      } else {
        const DevView& V = view_at(a, fr.y >> kViewShift);
        W = V.v.width;
        H = V.v.height;
        L.dst[lane] = V.depth;
        L.ids[lane] = V.idkey;
        L.key[lane] = ((unsigned long long)(c.key + 1u) << 32) | fr.x;
      }
      L.pitch[lane] = W;
      Setup st;
      raster_setup(c, (float)W, (float)H, st);
      const int pa = op_size * (int)(fr.y & kChunkMask);
      if (st.width > 0) {
        // samples of this fragment: q in [pa, pa + op_size] with y(q) < max_y
        const int q_end = run_start(st.mm[3], st.width, st.iwidth, st.mm[2]);
        int c2 = q_end - pa;
        c2 = c2 < 0 ? 0 : (c2 > op_size + 1 ? op_size + 1 : c2);
        cnt = (uint32_t)c2;
      }
#pragma unroll
      for (int k = 0; k < 4; k++) L.mm[k][lane] = st.mm[k];
      L.pl[0][lane] = st.A; L.pl[1][lane] = st.B; L.pl[2][lane] = st.C; L.pl[3][lane] = st.iwidth;
      // point_in_tri's loop-invariant sub-expressions (same operation order)
      const float p0x = st.xp[0], p0y = st.yp[0], p1x = st.xp[1], p1y = st.yp[1], p2x = st.xp[2], p2y = st.yp[2];
      const float A = 0.5f * (-p1y * p2x + p0y * (-p1x + p2x) + p0x * (p1y - p2y) + p1x * p2y);
      const float sign = A < 0 ? -1.f : 1.f;
      L.ek[0][lane] = p0y * p2x - p0x * p2y;
      L.ek[1][lane] = p2y - p0y;
      L.ek[2][lane] = p0x - p2x;
      L.ek[3][lane] = p0x * p1y - p0y * p1x;
      L.ek[4][lane] = p0y - p1y;
      L.ek[5][lane] = p1x - p0x;
      L.ek[6][lane] = sign;
      L.ek[7][lane] = 2.0001f * A * sign;
      L.w[lane] = st.width;
      L.pa[lane] = pa;
    }
    const uint32_t incl = wave_incl_scan(cnt);
    const uint32_t total = __shfl(incl, kWave - 1, kWave);
    L.pre[lane] = incl - cnt;
    if (lane == kWave - 1) L.pre[kWave] = total;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // ---- sweep every sample of the 64 fragments with all lanes
    for (uint32_t j = lane; j < total; j += kWave) {
      int k = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1)
        if (L.pre[k + step] <= j) k += step;
      const int w = L.w[k];
      const float min_x = L.mm[0][k], max_x = L.mm[1][k], min_y = L.mm[2][k];
      const float iw = L.pl[3][k];
      const int q = L.pa[k] + (int)(j - L.pre[k]);
      const float y = row_of(q, iw, min_y);
      int r = run_start(y, w, iw, min_y);
      r = r > L.pa[k] ? r : L.pa[k];
      const float x = (float)((uint32_t)r % (uint32_t)w) + min_x + (float)(q - r);
      if (x >= max_x) continue;
      const float sign = L.ek[6][k];
      const float s = (L.ek[0][k] + L.ek[1][k] * x + L.ek[2][k] * y) * sign;
      const float t = (L.ek[3][k] + L.ek[4][k] * x + L.ek[5][k] * y) * sign;
      if (!(s > -0.0001f && t > -0.0001f && (s + t) < L.ek[7][k])) continue;
      const float fd = fmaf(L.pl[0][k], x, fmaf(L.pl[1][k], y, L.pl[2][k]));
      const uint32_t d = sat_u32(kMulint / fd);
      const size_t p = (size_t)(int)y * L.pitch[k] + (size_t)(int)x;
      if constexpr (kMode == RasterMode::Id) {
        const uint32_t val = L.dst[k][p];
        if (d > val - kBufError && d < val + kBufError) atomicMax(&L.ids[k][p], L.key[k]);
      } else {
        atomicMin(&L.dst[k][p], d);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

int device_cu_count();

hipError_t launch_depth(const PassArgs& a, int blocks_per_cu, hipStream_t s) {
  hipLaunchKernelGGL(k_raster<RasterMode::Depth>, dim3(device_cu_count() * blocks_per_cu), dim3(256), 0, s, a,
                     ShadowTarget{nullptr, 0}, kOpSize);
  return hipGetLastError();
}

// kernel2's "last writer wins" is made deterministic: among fragments inside
// the BUF_ERROR window the pixel keeps the largest (2*tri + sub) key, which
// is the fragment a sequential in-order execution writes last.
hipError_t launch_id(const PassArgs& a, int blocks_per_cu, hipStream_t s) {
  hipLaunchKernelGGL(k_raster<RasterMode::Id>, dim3(device_cu_count() * blocks_per_cu), dim3(256), 0, s, a,
                     ShadowTarget{nullptr, 0}, kOpSize);
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
#pragma unroll
  for (int kk = 0; kk < 6; kk++) {
    if (!(touched >> kk & 1)) continue;
    f3 tp[2][3];
    int num = transform_clip_project(T, G, lpos, ft.f[kk], efov, ew, eh, tp);
#pragma unroll
    for (int i = 0; i < 2; i++) {
      if (i >= num) continue;
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
#pragma unroll
  for (int kk = 0; kk < 6; kk++) {
    if (!(touched >> kk & 1)) continue;
    f3 tp[2][3];
    int num = transform_clip_project(T, G, lpos, ft.f[kk], efov, ew, eh, tp);
#pragma unroll
    for (int i = 0; i < 2; i++) {
      if (i >= num) continue;
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
  PassArgs a;
  memset(&a, 0, sizeof(a));
  a.lists = Lists{sc->sh_cuts, sc->sh_frags, sc->sh_counters, sc->sh_cap_cut, sc->sh_cap_frag};
  hipLaunchKernelGGL(k_raster<RasterMode::Shadow>, dim3(device_cu_count() * 8), dim3(256), 0, s, a,
                     ShadowTarget{map, sc->lbuf_dim}, kOpSizeLight);
  return hipGetLastError();
}

}  // namespace ocr
