// raster.hip — prearrange (transform, near clip, cull, sample count), the
// depth pass (u32 atomicMin) and the id pass (u64 atomicMax on a packed key),
// for camera views and for cube shadow maps.
//
// Reference: prearrange cl2.cl:4271-4409, kernel1 cl2.cl:4985-5127,
// kernel2 cl2.cl:5390-5546, prearrange_realtime_shadowing cl2.cl:4419-4636,
// kernel1_realtime_shadowing cl2.cl:5129-5246.
//
// Work decomposition.  The reference cuts every triangle's bounding box into
// fragments of op_size pixels and runs one work item per fragment; a fragment
// walks the samples q = pa .. pa + op_size (pa = op_size * chunk) with a
// running cursor and stops at the first row >= max_y.  Fragments hold from 1
// to op_size + 1 samples, so one work item per fragment leaves most lanes of
// a wavefront idle.  Here prearrange computes each cut triangle's exact
// sample count over all of its fragments and appends the triangle with its
// range in one global sample space: one packed 64-bit atomicAdd per
// wavefront returns both the record index and the sample offset, so the cut
// list is sorted by sample offset and its ranges tile [0, total).  The raster
// passes split that sample space into equal chunks: a wavefront finds the
// chunk's first triangle with a 64-ary search, stages 64 triangles' setup in
// LDS and sweeps the chunk with all lanes.  No per-fragment records exist.
#pragma clang fp contract(off)

#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include "ocr_internal.h"
#include "tiles.h"
#include "walk.h"

namespace ocr {

// Prearrange blocks: one triangle per work item.  Cut records and sample
// ranges are allocated once per block (one 64-bit atomicAdd on the packed
// counter (cuts << 32 | samples) per block): every wavefront of the render
// appending to one counter would serialise on that address (~12 ns per
// same-address atomic).  The camera pass uses 512 work items: the next
// render's prearrange runs beside the current render's shading, and a
// 512-item block (8 wavefronts) fits on a CU beside shading wavefronts where
// a 1024-item one waits for the CU to drain (measured: 4083-4094 vs
// 4019-4024 frames/s on the 1080p, 1 M-triangle frame; 256 items pays for
// twice the atomics, 57 vs 40 us alone).  The shadow prearrange (104 VGPRs)
// also uses 512: two blocks per CU hide each other's loads, barriers and
// allocation atomic (16-light shadow pass over 4 M triangles 5.09 -> 4.82 ms).
constexpr int kPreThreads = 512;
constexpr int kShadowPreThreads = 512;

template <int kWaves>
struct BlockAllocLds {
  uint32_t cuts[kWaves], smps[kWaves], frags[kWaves];
  unsigned long long base;
  uint32_t ok;
};

// Returns false (for the whole block) when the block's records do not fit the
// list or the sample space; the overflow word then says which.  The block's
// sample total is checked before its atomic, and each atomic adds less than
// kMaxSamples, so the low half reaches the threshold and sets the flag before
// it could carry into the record count.
template <int kWaves>
__device__ __forceinline__ bool block_alloc(const Lists& L, BlockAllocLds<kWaves>& sh, uint32_t ncut, uint32_t nsmp,
                                            uint32_t nfrag, uint32_t& cut0, uint32_t& smp0) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const uint32_t ic = wave_incl_scan(ncut), is = wave_incl_scan(nsmp), ifr = wave_incl_scan(nfrag);
  if (lane == kWave - 1) {
    sh.cuts[wave] = ic;
    sh.smps[wave] = is;
    sh.frags[wave] = ifr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tc = 0, tf = 0;
    unsigned long long ts = 0;
    for (int w = 0; w < kWaves; w++) {
      const uint32_t c = sh.cuts[w], sm = sh.smps[w];
      sh.cuts[w] = tc;
      sh.smps[w] = (uint32_t)ts;
      tc += c;
      ts += sm;
      tf += sh.frags[w];
    }
    uint32_t bad = ts > kMaxSamples ? kOverflowSamples : 0u;
    unsigned long long base = 0;
    if (!bad && tc) {
      base = atomicAdd(reinterpret_cast<unsigned long long*>(L.counters + kPacked), ((unsigned long long)tc << 32) + ts);
      const unsigned long long end = base + ((unsigned long long)tc << 32) + ts;
      if ((end >> 32) > L.cap_cut) bad |= kOverflowCuts;
      if ((uint32_t)end > kMaxSamples) bad |= kOverflowSamples;
    }
    if (tf && L.count_frags) atomicAdd(&L.counters[kFragCount], tf);
    if (bad) atomicOr(&L.counters[kOverflow], bad);
    sh.base = base;
    sh.ok = bad == 0;
  }
  __syncthreads();
  const unsigned long long base = sh.base;
  cut0 = (uint32_t)(base >> 32) + sh.cuts[wave] + ic - ncut;
  smp0 = (uint32_t)base + sh.smps[wave] + is - nsmp;
  const bool ok = sh.ok != 0;
  __syncthreads();   // the LDS record is reused by the block's next allocation
  return ok;
}

// ---------------------------------------------------------- setup helpers

// Setup of one cut triangle for the raster passes (kernel1 cl2.cl:5000-5046).
__device__ __forceinline__ void raster_setup(const f3 p[3], float rc, float ew, float eh, RasterTri& t,
                                             float& max_y) {
  const float xs[3] = {p[0].x, p[1].x, p[2].x}, ys[3] = {p[0].y, p[1].y, p[2].y};
  float mm[4];
  calc_min_max(xs, ys, ew, eh, mm);
  t.min_x = mm[0];
  t.min_y = mm[2];
  max_y = mm[3];
  t.w = mm[1] - mm[0];
  t.iw = rcp_cr((float)(int)t.w);   // integer-valued
  const float xp[3] = {roundf(p[0].x), roundf(p[1].x), roundf(p[2].x)};
  const float yp[3] = {roundf(p[0].y), roundf(p[1].y), roundf(p[2].y)};
  // 1 / (z / depth_far) for z in [20, 350000]: the quotient by the constant
  // is exact for every finite z >= 0, its reciprocal lies in rcp_rn's range
  auto inv_rel = [](float z) { return rcp_rn(div_rcp(z, kDepthFar, kInvDepthFar)); };
  const f3 depths = mk3(inv_rel(p[0].z), inv_rel(p[1].z), inv_rel(p[2].z));
  plane_const(depths, mk3(xp[0], xp[1], xp[2]), mk3(yp[0], yp[1], yp[2]), rc, t.A, t.B, t.C);
  // point_in_tri (cl2.cl:4798-4807), loop-invariant terms in the same order
  const float p0x = xp[0], p0y = yp[0], p1x = xp[1], p1y = yp[1], p2x = xp[2], p2y = yp[2];
  const float A = 0.5f * (-p1y * p2x + p0y * (-p1x + p2x) + p0x * (p1y - p2y) + p1x * p2y);
  const float sign = A < 0 ? -1.f : 1.f;
  t.ek[0] = (p0y * p2x - p0x * p2y) * sign;
  t.ek[1] = (p2y - p0y) * sign;
  t.ek[2] = (p0x - p2x) * sign;
  t.ek[3] = (p0x * p1y - p0y * p1x) * sign;
  t.ek[4] = (p0y - p1y) * sign;
  t.ek[5] = (p1x - p0x) * sign;
  t.ek[6] = 2.0001f * A * sign;
}

// Samples of all fragments of a cut triangle.  Fragment a visits
// q in [op*a, op*a + op] up to the first q whose row is >= max_y (q_end), so
// it holds clamp(q_end - op*a, 0, op + 1) samples; the full fragments come
// first, then at most one partial one.
__device__ __forceinline__ uint32_t triangle_samples(const RasterTri& t, float max_y, int op, uint32_t nfrag) {
  if (nfrag == 0 || !(t.w > 0.f)) return 0;
  // q_end < 2^24 (frames hold at most 16,000,000 pixels): 32-bit arithmetic
  const uint32_t q_end = (uint32_t)run_start(max_y, t.w, t.iw, t.min_y);
  const uint32_t uop = (uint32_t)op;
  uint32_t m = q_end >= uop + 1 ? (q_end - (uop + 1)) / uop + 1 : 0;
  if (m > nfrag) m = nfrag;
  uint64_t s = (uint64_t)(uop + 1) * m;
  if (m < nfrag) {
    const int64_t r = (int64_t)q_end - (int64_t)uop * m;
    s += r < 0 ? 0 : (r > (int64_t)uop + 1 ? uop + 1 : (uint64_t)r);
  }
  return s > (uint64_t)kMaxTriSamples ? kMaxTriSamples : (uint32_t)s;
}

// Clip/cull/count of one projected triangle (cl2.cl:4344-4380): returns the
// reference's fragment count ceil(bbox_area / op_size), 0 if culled.
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
  float n = ceilf(div_rcp(area, (float)op_size, 1.0f / (float)op_size));   // area >= 0, op 300 or 500
  return n > 0.f ? (uint32_t)n : 0u;
}

template <bool kCamera>
__device__ __forceinline__ void store_cut(CutTri* dst, const f3 p[3], float rconst, uint32_t key, uint32_t oid,
                                          uint32_t view) {
  CutTri c;
#pragma unroll
  for (int i = 0; i < 3; i++) { c.x[i] = p[i].x; c.y[i] = p[i].y; c.z[i] = p[i].z; }
  c.rconst = rconst;
  c.key = key;
  c.oid = oid;
  c.view = view;
  c.da = c.db = c.dc = 0.f;
  if constexpr (kCamera) {
    // the plane kernel3's get_vtdiff builds per pixel (cl2.cl:5691-5760)
    const f3 xpv = mk3(roundf(p[0].x), roundf(p[1].x), roundf(p[2].x));
    const f3 ypv = mk3(roundf(p[0].y), roundf(p[1].y), roundf(p[2].y));
    const f3 depths = mk3(rcp_rn(p[0].z), rcp_rn(p[1].z), rcp_rn(p[2].z));   // z in [20, 350000]
    plane_const(depths, xpv, ypv, rconst, c.da, c.db, c.dc);
  }
  *dst = c;
}

// full_rotate_quat's object-to-world step (cl2.cl:700-729): the same for the
// camera and every light, so it is applied once per object update into the
// scene's world-space copy of its triangles.
__global__ void __launch_bounds__(256) k_world(const TriPos* __restrict__ tris, uint32_t ntri,
                                               const ObjDesc* __restrict__ objs, uint32_t nobj,
                                               TriPos* __restrict__ world) {
  const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= ntri) return;
  const TriPos T = tris[id];
  TriPos W;
  W.oid = T.oid;
  if (T.oid < nobj) {
    const ObjDesc G = objs[T.oid];
    const f3 wp = mk3(G.wp[0], G.wp[1], G.wp[2]);
    // rot_quat's normalised quaternion, formed once per descriptor
    const f4 qn = f4{G.qn[0], G.qn[1], G.qn[2], G.qn[3]};
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const f3 r = rot_quat_n(mk3(T.p[3 * i], T.p[3 * i + 1], T.p[3 * i + 2]) * G.scale, qn) + wp;
      W.p[3 * i] = r.x;
      W.p[3 * i + 1] = r.y;
      W.p[3 * i + 2] = r.z;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 9; i++) W.p[i] = 0.f;
  }
  world[id] = W;
}

hipError_t launch_world(Scene* sc, hipStream_t s) {
  if (sc->ntri)
    hipLaunchKernelGGL(k_world, dim3((sc->ntri + 255) / 256), dim3(256), 0, s, sc->tri_pos, sc->ntri, sc->objs,
                       sc->nobj, sc->world);
  return hipGetLastError();
}

// full_rotate_quat's rotation into a camera or cube face (cl2.cl:700-729),
// near clip and projection, from world-space vertices.
__device__ __forceinline__ int face_clip_project(const f3 w[3], f3 cpos, const SinCos& sc, float fovc, float ew,
                                                 float eh, f3 out[2][3]) {
  f3 pr[3];
#pragma unroll
  for (int i = 0; i < 3; i++) pr[i] = rot(w[i], cpos, sc);
  f3 tris[2][3];
  const int n = clip_near(pr[0], pr[1], pr[2], tris);
  // slot 1 exists only for triangles crossing the near plane with one
  // vertex behind it; static indices keep both slots in registers
#pragma unroll
  for (int i = 0; i < 3; i++) out[0][i] = depth_project_clipped(tris[0][i], ew, eh, fovc);
  if (n == 2) {
#pragma unroll
    for (int i = 0; i < 3; i++) out[1][i] = depth_project_clipped(tris[1][i], ew, eh, fovc);
  } else {
#pragma unroll
    for (int i = 0; i < 3; i++) out[1][i] = out[0][i];
  }
  return n;
}

// Appends this work item's cut triangles that hold samples (0..2) with their
// sample ranges; called by every work item of the block.  A block that would
// pass the list capacity or the sample limit writes nothing and sets the
// overflow word, which every later pass of the render checks before reading
// the list.
// For a tiled render (L.tb.bin set) also the records' tile boxes and their
// tiles' entry counts.
template <bool kCamera, int kWaves>
__device__ __forceinline__ void append_cuts(const Lists& L, BlockAllocLds<kWaves>& sh, const f3 tp[2][3],
                                            const float rc[2], RasterTri rt[2], const uint32_t nf[2],
                                            const uint32_t ns[2], uint32_t key0, uint32_t key_step, uint32_t oid,
                                            uint32_t view, const float* my = nullptr, uint32_t bin_begin = 0,
                                            uint32_t tiles_x = 0) {
  const uint32_t k0 = ns[0] > 0, k1 = ns[1] > 0;
  uint32_t c, smp;
  if (!block_alloc(L, sh, k0 + k1, ns[0] + ns[1], nf[0] * k0 + nf[1] * k1, c, smp)) return;
  uint4 box[2] = {};
#pragma unroll
  for (int i = 0; i < 2; i++) {
    if (!ns[i]) continue;
    const uint32_t key = key0 + key_step * (uint32_t)i;
    // shadow records are read only through their raster setup
    if constexpr (kCamera) store_cut<kCamera>(&L.cuts[c], tp[i], rc[i], key, oid, view);
    rt[i].key = key | (rt[i].key & kKeyBoxForm);
    rt[i].sample_begin = smp;
    L.rast[c] = rt[i];
    if constexpr (kCamera) {
      if (L.tb.bin) {
        box[i] = tile_box(rt[i], my[i], L.tb, bin_begin, tiles_x);
        L.tb.box[c] = box[i];
      }
    }
    smp += ns[i];
    c++;
  }
  if constexpr (kCamera) {
    if (L.tb.bin) {
      uint32_t* cnt = L.tb.cnt;
      const uint32_t ntiles = L.tb.ntiles;
      expand_tile_pairs(
          box[0], k0 != 0, box[1], k1 != 0,
          [cnt, ntiles](uint32_t t, uint32_t len) {
            if (t < ntiles) atomicAdd(&cnt[t], len);
            return 0u;
          },
          [](uint32_t, int, uint32_t, uint32_t) {});
    }
  }
}

// Cull, count and set up the (up to 2) clipped pieces of one triangle.
// my[i]: piece i's max_y (its box's end row).  A piece whose walk visits
// exactly its box gets the box form (rt.key = kKeyBoxForm; the caller adds
// the key): one sample per box pixel, mapped without the walk's fragment
// arithmetic.
__device__ __forceinline__ void cut_pieces(const f3 tp[2][3], int num, float ew, float eh, bool two_sided, int op,
                                           float rc[2], RasterTri rt[2], uint32_t nf[2], uint32_t ns[2],
                                           float my[2]) {
#pragma unroll
  for (int i = 0; i < 2; i++) {
    if (i >= num) continue;
    nf[i] = cull_count(tp[i], ew, eh, two_sided, op, rc[i]);
    if (!nf[i]) continue;
    raster_setup(tp[i], rc[i], ew, eh, rt[i], my[i]);
    // the box holds at most 16,000,000 pixels (frames and cube faces), an
    // exact product, 0 when the walk has no samples (w or rows 0); the
    // closed forms of walk_is_rect settle all but tall triangles near the
    // top row, which keep the walk's samples
    const float area = rt[i].w * (my[i] - rt[i].min_y);
    const bool box = area >= 1.f && area <= (float)kMaxTriSamples &&
                     walk_is_rect_closed(rt[i].min_y, my[i], rt[i].w, rt[i].iw);
    rt[i].key = box ? kKeyBoxForm : 0u;
    ns[i] = box ? (uint32_t)area : triangle_samples(rt[i], my[i], op, nf[i]);
  }
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
enum class Range { Tri, Pix };
template <Range kR>
__device__ __forceinline__ uint32_t view_of(const PassArgs& a, uint32_t slot) {
  if (!a.views) return 0;
  uint32_t lo = 0, hi = a.nviews;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    const DevView& m = a.views[mid];
    uint32_t b = kR == Range::Tri ? m.tri_begin : m.pix_begin;
    if (b <= slot) lo = mid; else hi = mid;
  }
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
}

// ------------------------------------------------------------------ clear
__global__ void __launch_bounds__(256) k_clear(PassArgs a) {
  const uint32_t slot = blockIdx.x * 256u;
  // (a plain render zeroes its counters on its side stream instead)
  if (a.lists.counters && blockIdx.x == 0 && threadIdx.x < kCounterWords) a.lists.counters[threadIdx.x] = 0u;
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
__global__ void __launch_bounds__(kPreThreads) k_prearrange(PassArgs a) {
  __shared__ BlockAllocLds<kPreThreads / kWave> sh;
  // per-view triangle ranges are 256-aligned: each wavefront has one view
  const uint32_t slot = blockIdx.x * (uint32_t)kPreThreads + (threadIdx.x & ~(uint32_t)(kWave - 1));
  const uint32_t vi = view_of<Range::Tri>(a, slot);
  const DevView& V = view_at(a, vi);
  const DevScene& S = scene_at(a, V.scene);
  const ViewParams& v = V.v;
  const uint32_t id = slot - V.tri_begin + (threadIdx.x & (kWave - 1));
  const float ew = (float)v.width, eh = (float)v.height;
  f3 tp[2][3];
  float rc[2] = {0.f, 0.f};
  RasterTri rt[2];
  uint32_t nf[2] = {0u, 0u}, ns[2] = {0u, 0u};
  float my[2] = {0.f, 0.f};
  uint32_t oid = 0;
  if (id < S.ntri) {
    const TriPos T = S.world[id];
    oid = T.oid;
    if (oid < S.nobj) {
      const ObjDesc G = S.objs[oid];
      f3 cpos = mk3(v.cpos[0], v.cpos[1], v.cpos[2]);
      if (!(fast_length(mk3(G.wp[0], G.wp[1], G.wp[2]) - cpos) > kDepthFar)) {
        const f3 w[3] = {mk3(T.p[0], T.p[1], T.p[2]), mk3(T.p[3], T.p[4], T.p[5]), mk3(T.p[6], T.p[7], T.p[8])};
        const int num = face_clip_project(w, cpos, v.cam, v.fov, ew, eh, tp);
        cut_pieces(tp, num, ew, eh, (G.flags & OCR_FEATURE_TWO_SIDED) > 0, kOpSize, rc, rt, nf, ns, my);
      }
    }
  }
  append_cuts<true>(a.lists, sh, tp, rc, rt, nf, ns, 2u * id, 1u, oid, vi, my, V.bin_begin, V.bin_tiles_x);
}

hipError_t launch_prearrange(const PassArgs& a, hipStream_t s) {
  if (a.tri_slots == 0) return hipSuccess;
  hipLaunchKernelGGL(k_prearrange, dim3((a.tri_slots + kPreThreads - 1) / kPreThreads), dim3(kPreThreads), 0, s, a);
  return hipGetLastError();
}

// ------------------------------------------------- sample-chunk raster passes
//
// Sample s of a cut triangle (0 <= s < nsamples) is sample o = s % (op + 1)
// of fragment a = s / (op + 1), i.e. q = op*a + o with pa = op*a.  Its cursor
// position is the closed form of the reference's running (x, y):
//   y(q) = floor(fma(q, 1/width, min_y))
//   x(q) = (r % width) + min_x + (q - r)
// where r is the later of pa and the first index of q's row run (the cursor
// is reset to running_width_mod + min_x exactly when y changes, and starts at
// pa % width + min_x).  Row-run starts come from the nominal
// (Y - min_y) * width by stepping while y disagrees, so rows whose fma
// rounding lags or leads by a sample are reproduced exactly.


// s_waitcnt immediate for vmcnt(0), expcnt(7), lgkmcnt(15) (gfx9 encoding:
// vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt_hi[15:14])
constexpr int kWaitVmcnt0 = 0x0F70;

enum class RasterMode { Depth, Id, Shadow };

struct ShadowTarget {
  uint32_t* map;   // 6 faces of L x L
  uint32_t L;
};

// Setups of 64 staged triangles, one column per lane.  The target buffer
// travels in the kernel arguments unless views differ per triangle (batch).
template <RasterMode kMode, bool kBatch>
struct TriLds {
  float min_x[kWave], min_y[kWave], w[kWave], iw[kWave];
  float pa[kWave], pb[kWave], pc[kWave];   // 1/z plane
  float ek[7][kWave];
  uint32_t key[kWave];
  uint32_t begin[kWave + 1];
  uint32_t start[kWave];   // start[p] == base: a staged triangle begins at sample base + p
  uint32_t pitch[kBatch ? kWave : 1];
  uint32_t* dst[kBatch ? kWave : 1];
  unsigned long long* ids[kBatch && kMode == RasterMode::Id ? kWave : 1];
};


// Whether sample_xy reduces, for every sample of a set-up triangle of n
// samples, to its short form
//   y = row_of(j),  x = min_x + (j - (y - min_y) * w):
// one fragment (n <= op + 1: a = 0, q = j) and rows starting exactly at
// the multiples of w (run_start(y) == (y - min_y) * w: row_of neither leads
// nor lags at any of the triangle's row boundaries, checked for up to 32
// rows).  Then r = (y - min_y) * w, floor(r / w) gives e = 0 after the
// corrections, and xo = j - r, the same binary32 values.
template <int kOp>
__device__ __forceinline__ bool short_form_ok(uint32_t n, float min_y, float w, float iw) {
  if (n == 0 || n > (uint32_t)kOp + 1 || !(w > 0.f)) return false;
  const float rows = row_of((float)(n - 1), iw, min_y) - min_y;
  if (!(rows <= 32.f)) return false;
  bool ok = true;
  for (float k = 1.f; k <= rows; k += 1.f) {
    const float n0 = k * w;
    ok = ok && row_of(n0 - 1.f, iw, min_y) == min_y + (k - 1.f) && row_of(n0, iw, min_y) == min_y + k;
  }
  return ok;
}

// Every sample of [s, stage_end) of the staged triangles with all lanes of
// the wavefront, one sample per lane per step (called wave-uniformly; two
// per lane measured slower: stages of small triangles end in part-filled
// steps).  my_begin is the first sample of the lane's staged triangle
// (stage_end or more for lanes without one); staged triangles hold at least
// one sample each, in lane order, and L.start holds no value in
// [s, stage_end) on entry.  A single view's target (width, depth, id keys)
// comes in view_*; shadow maps come in view_depth.
// Bit k of short_bits: staged triangle k maps by the short form (box-form
// records, and walk-form ones that short_form_ok certifies); lanes sampling
// such triangles take it, the others the walk's general form.
// ent (depth pass of a single view): the chunk's covered-sample entries,
// nent of them so far (see CoverEntries).
template <RasterMode kMode, int kOp, bool kBatch>
__device__ __forceinline__ void sweep_samples(TriLds<kMode, kBatch>& L, uint32_t view_width, uint32_t* view_depth,
                                              unsigned long long* view_ids, uint32_t shadow_dim, uint32_t my_begin,
                                              uint32_t s, uint32_t stage_end, uint32_t c0,
                                              unsigned long long short_bits, uint2* ent, uint32_t& nent) {
  const int lane = threadIdx.x & (kWave - 1);
  for (uint32_t base = s; base < stage_end; base += kWave) {
    // Lane p takes sample j = base + p.  Its staged triangle k is one less
    // than the number of staged triangles beginning at or before j: those
    // beginning at or before `base` (a prefix of the lanes, one ballot)
    // plus those beginning in (base, j] (flags written by the triangles'
    // lanes, counted below p with mbcnt).
    const unsigned long long before = __ballot(my_begin <= base);
    const uint32_t off = my_begin - base;
    if (my_begin > base && off < (uint32_t)kWave) L.start[off] = base;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bool starts_here = L.start[lane] == base;
    const unsigned long long starts = __ballot(starts_here);
    const uint32_t j = base + (uint32_t)lane;
    bool hit = false, rec = false;
    int k = 0;
    float x = 0.f, y = 0.f;
    if (j < stage_end) {
      k = __popcll(before) - 1 +
          (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(starts >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)starts, 0u)) +
          (starts_here ? 1 : 0);
      const float slf = (float)(j - L.begin[k]);
      // box-form triangles take the short form in every step (their sample
      // indices are box pixels, not walk indices)
      if ((short_bits >> k) & 1ull) {
        const float min_y = L.min_y[k], w = L.w[k];
        y = row_of(slf, L.iw[k], min_y);
        const float xo = slf - (y - min_y) * w;
        x = L.min_x[k] + xo;
        hit = !(xo >= w) &&
              inside_tri(x, y, L.ek[0][k], L.ek[1][k], L.ek[2][k], L.ek[3][k], L.ek[4][k], L.ek[5][k], L.ek[6][k]);
      } else {
        hit = sample_xy<kOp>(slf, L.min_x[k], L.min_y[k], L.w[k], L.iw[k], L.ek[0][k], L.ek[1][k], L.ek[2][k],
                             L.ek[3][k], L.ek[4][k], L.ek[5][k], L.ek[6][k], x, y);
      }
    }
    uint32_t d = 0, pix32 = 0, xy = 0;
    if (hit) {
      d = depth_quant(fmaf(L.pa[k], x, fmaf(L.pb[k], y, L.pc[k])));
      uint32_t pitch;
      uint32_t* dst;
      if constexpr (kMode == RasterMode::Shadow) {
        // view_depth: the maps, 6 faces of shadow_dim^2 per light
        pitch = shadow_dim;
        dst = view_depth + (size_t)L.key[k] * shadow_dim * shadow_dim;
      } else if constexpr (kBatch) {
        pitch = L.pitch[k];
        dst = L.dst[k];
      } else {
        pitch = view_width;
        dst = view_depth;
      }
      const uint32_t xi = (uint32_t)(int)x, yi = (uint32_t)(int)y;
      xy = xi | yi << 16;
      // y < 2^15 and the pitch < 2^16 (frames are at most 32768 wide and
      // tall, shadow faces 4000): the 24-bit multiply is exact
      pix32 = __umul24(yi, pitch) + xi;
      const size_t pix = (size_t)pix32;
      if constexpr (kMode == RasterMode::Id) {
        unsigned long long* ids = kBatch ? L.ids[kBatch ? k : 0] : view_ids;
        const uint32_t val = dst[pix];
        if (d > val - kBufError && d < val + kBufError)
          atomicMax(&ids[pix], ((unsigned long long)(L.key[k] + 1u) << 32) | (c0 + (uint32_t)k));
      } else {
        // The L2 performs atomics at ~26.6 G cache-line operations/s; a
        // plain load first drops the samples that cannot win (the stored
        // depth only decreases, so d >= any value a load returns loses).
        // A single view's buffer is addressed from its (uniform) base with
        // a 32-bit offset.
        uint32_t val;
        if constexpr (kMode == RasterMode::Depth && !kBatch) {
          typedef __attribute__((address_space(1))) char gchar;
          typedef __attribute__((address_space(1))) uint32_t gu32;
          gu32* g = (gu32*)((gchar*)(uintptr_t)dst + pix32 * 4u);   // 32-bit byte offset: frames < 2^30 px
          val = *g;
          if (d < val) __hip_atomic_fetch_min(g, d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          val = dst[pix];
          if (d < val) atomicMin(&dst[pix], d);
        }
        // a sample at or beyond val + BUF_ERROR stays outside the id
        // pass's window of the final depth (<= val) unless val + BUF_ERROR
        // wraps
        rec = val > 0xFFFFFFFFu - kBufError || d < val + kBufError;
      }
    }
    if constexpr (kMode == RasterMode::Depth) {
      if (ent) {
        const unsigned long long m = __ballot(rec);
        if (rec) {
          const uint32_t slot =
              nent + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
          if (slot < kEntCap) ent[slot] = uint2{xy, c0 + (uint32_t)k};
        }
        nent += (uint32_t)__popcll(m);
      }
    }
  }
}

// The id pass over the covered-sample entries the depth pass kept: the
// wavefront's chunks (wave0, wave0 + nwaves, ...) two at a time, their
// entries as one list (at most kEntCap per chunk; a chunk with more is swept
// by the caller), four of the lane's entries at a time: the entries, then
// their records' depth planes and keys, then the depths there, then the
// window tests.  Each entry's depth is formed again as the depth pass formed
// it.  Two chunks per round halve the rounds of dependent loads a wavefront
// waits through.
template <bool kBatch>
__device__ __forceinline__ void id_from_entries(const PassArgs& a, const Lists& Ls, uint32_t wave0, uint32_t nwaves,
                                                uint32_t nchunks) {
  const int lane = threadIdx.x & (kWave - 1);
  constexpr int kPer = 4;
  const uint4* R = reinterpret_cast<const uint4*>(Ls.rast);   // 4 x uint4 per record
  for (uint32_t ca = wave0; ca < nchunks; ca += 2 * nwaves) {
    const uint32_t cb = ca + nwaves;
    const uint32_t na0 = Ls.ent.cnt[ca], nb0 = cb < nchunks ? Ls.ent.cnt[cb] : 0u;
    const uint32_t na = na0 <= kEntCap ? na0 : 0u, nb = nb0 <= kEntCap ? nb0 : 0u;
    const uint32_t n = na + nb;
    const uint2* Ea = Ls.ent.ent + (size_t)ca * kEntCap;
    // entry e >= na of the list is entry e - na of chunk cb
    const uint2* Eb = Ls.ent.ent + (size_t)(cb < nchunks ? cb : ca) * kEntCap;
    for (uint32_t e0 = 0; e0 < n; e0 += kPer * kWave) {
      uint2 v[kPer];
      uint32_t d[kPer], key[kPer], pix[kPer], val[kPer];
      uint32_t* dep[kPer];
      unsigned long long* ids[kPer];
#pragma unroll
      for (int i = 0; i < kPer; i++) {
        const uint32_t e = e0 + (uint32_t)lane + (uint32_t)(i * kWave);
        if (e < n) v[i] = e < na ? Ea[e] : Eb[e - na];
      }
#pragma unroll
      for (int i = 0; i < kPer; i++) {
        const uint32_t e = e0 + (uint32_t)lane + (uint32_t)(i * kWave);
        dep[i] = a.view0.depth;
        ids[i] = a.view0.idkey;
        uint32_t pitch = a.view0.v.width;
        if (e < n) {
          // RasterTri words 1 (key) and 5-7 (the 1/z plane A, B, C)
          const uint4 w0 = R[4 * (size_t)v[i].y], w1 = R[4 * (size_t)v[i].y + 1];
          if constexpr (kBatch) {
            // a batch's entries hold pixels of their record's view
            const DevView& V = a.views[Ls.cuts[v[i].y].view];
            dep[i] = V.depth;
            ids[i] = V.idkey;
            pitch = V.v.width;
          }
          const float x = (float)(v[i].x & 0xFFFFu), y = (float)(v[i].x >> 16);
          d[i] = depth_quant(fmaf(__uint_as_float(w1.y), x, fmaf(__uint_as_float(w1.z), y, __uint_as_float(w1.w))));
          key[i] = w0.y & ~kKeyBoxForm;
          pix[i] = (v[i].x >> 16) * pitch + (v[i].x & 0xFFFFu);
        }
      }
#pragma unroll
      for (int i = 0; i < kPer; i++) {
        const uint32_t e = e0 + (uint32_t)lane + (uint32_t)(i * kWave);
        if (e < n) val[i] = dep[i][pix[i]];
      }
#pragma unroll
      for (int i = 0; i < kPer; i++) {
        const uint32_t e = e0 + (uint32_t)lane + (uint32_t)(i * kWave);
        if (e < n && d[i] > val[i] - kBufError && d[i] < val[i] + kBufError)
          atomicMax(&ids[i][pix[i]], ((unsigned long long)(key[i] + 1u) << 32) | v[i].y);
      }
    }
  }
}

template <RasterMode kMode, int kOp, bool kBatch>
__global__ void __launch_bounds__(256) k_raster(PassArgs a, ShadowTarget sh) {
  __shared__ TriLds<kMode, kBatch> lds_all[256 / kWave];
  const Lists& Ls = a.lists;
  if (Ls.counters[kOverflow]) return;
  const int lane = threadIdx.x & (kWave - 1);
  // the list's records and samples; shadow lists are kShadowParts partitions
  // (lane p holds partition p's records, samples and inclusive chunk count)
  uint32_t ncut = 0, total = 0, nchunks;
  uint32_t part_ncut = 0, part_total = 0, part_chunks = 0;
  if constexpr (kMode == RasterMode::Shadow) {
    // (records << 40 | samples), up to the valid prefix
    const unsigned long long* pc = reinterpret_cast<const unsigned long long*>(Ls.counters + kPartWord(lane));
    const unsigned long long packed = min(pc[0], pc[1]);
    part_ncut = (uint32_t)(packed >> kShadowSampleBits);
    part_total = (uint32_t)(packed & kShadowSampleMask);   // at most 2^30
    part_chunks = wave_incl_scan((part_total + kChunk - 1) / kChunk);
    nchunks = __shfl(part_chunks, kWave - 1, kWave);
  } else {
    const unsigned long long packed = *reinterpret_cast<const unsigned long long*>(Ls.counters + kPacked);
    ncut = (uint32_t)(packed >> 32);
    total = (uint32_t)packed;
    nchunks = (total + kChunk - 1) / kChunk;
  }
  TriLds<kMode, kBatch>& L = lds_all[threadIdx.x / kWave];
  const uint32_t wave0 = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  L.start[lane] = 0xFFFFFFFFu;   // never a sample index (total < 2^31)
  if constexpr (kMode == RasterMode::Id) {
    if (Ls.ent.cnt) id_from_entries<kBatch>(a, Ls, wave0, nwaves, min(Ls.ent.chunks, nchunks));
  }
  for (uint32_t chunk = wave0; chunk < nchunks; chunk += nwaves) {
    const RasterTri* rast = Ls.rast;
    uint32_t s0 = chunk * kChunk;
    if constexpr (kMode == RasterMode::Shadow) {
      // the chunk's partition: the first whose inclusive chunk count passes it
      const uint32_t part = (uint32_t)__popcll(__ballot(part_chunks <= chunk));
      const uint32_t before = part ? (uint32_t)__shfl(part_chunks, (int)part - 1, kWave) : 0u;
      ncut = __shfl(part_ncut, (int)part, kWave);
      total = __shfl(part_total, (int)part, kWave);
      rast = Ls.rast + (size_t)part * (Ls.cap_cut / kShadowParts);
      s0 = (chunk - before) * kChunk;
      // sample indices are the partition's own, so the start flags of an
      // earlier chunk (another partition) may hold this chunk's indices
      L.start[lane] = 0xFFFFFFFFu;
    }
    const uint32_t s1 = min(s0 + kChunk, total);
    // covered-sample entries of the chunk: written by the depth pass of a
    // single view, read by its id pass instead of a second sweep
    uint2* ent = nullptr;
    uint32_t nent = 0;
    if constexpr (kMode != RasterMode::Shadow) {
      if (Ls.ent.cnt && chunk < Ls.ent.chunks) {
        if constexpr (kMode == RasterMode::Id) {
          // chunks whose entries the depth pass kept were done above
          if (Ls.ent.cnt[chunk] <= kEntCap) continue;
        } else {
          ent = Ls.ent.ent + (size_t)chunk * kEntCap;
        }
      }
    }
    // ---- the chunk's first triangle: last record with sample_begin <= s0
    // (record 0 begins at sample 0; 64-ary search, one probe per lane)
    uint32_t lo = 0, hi = ncut;
    while (hi - lo > 1) {
      const uint32_t p = lo + (uint32_t)(((unsigned long long)(hi - lo) * (uint32_t)lane) / kWave);
      const unsigned long long m = __ballot(rast[p].sample_begin <= s0);
      const int t = 63 - __clzll((long long)m);
      const uint32_t plo = __shfl(p, t, kWave);
      const uint32_t phi = __shfl(p, t < kWave - 1 ? t + 1 : t, kWave);
      lo = plo;
      hi = t < kWave - 1 ? phi : hi;
    }
    uint32_t c0 = lo;
    uint32_t s = s0;
    while (s < s1) {
      // ---- stage triangles c0 .. c0 + 63, one per lane
      const uint32_t ci = c0 + lane;
      uint32_t my_begin = total;   // this lane's staged triangle's first sample
      float t_min_y = 0.f, t_w = 0.f, t_iw = 0.f;
      bool t_box = false;
      if (ci < ncut) {
        const RasterTri t = rast[ci];
        my_begin = t.sample_begin;
        t_min_y = t.min_y;
        t_w = t.w;
        t_iw = t.iw;
        t_box = (t.key & kKeyBoxForm) != 0u;
        L.min_x[lane] = t.min_x;
        L.min_y[lane] = t.min_y;
        L.w[lane] = t.w;
        L.iw[lane] = t.iw;
        L.pa[lane] = t.A;
        L.pb[lane] = t.B;
        L.pc[lane] = t.C;
#pragma unroll
        for (int e = 0; e < 7; e++) L.ek[e][lane] = t.ek[e];
        L.key[lane] = t.key & ~kKeyBoxForm;
        L.begin[lane] = t.sample_begin;
        if constexpr (kBatch) {
          const DevView& V = a.views[Ls.cuts[ci].view];
          L.dst[lane] = V.depth;
          L.pitch[lane] = V.v.width;
          if constexpr (kMode == RasterMode::Id) L.ids[lane] = V.idkey;
        }
        if (lane == kWave - 1) L.begin[kWave] = ci + 1 < ncut ? rast[ci + 1].sample_begin : total;
      } else {
        L.begin[lane] = total;
        if (lane == kWave - 1) L.begin[kWave] = total;
      }
      // Every staging load has landed before the sweep.  vmcnt counts loads
      // and atomics in one in-order counter: without this unconditional wait
      // the compiler waits for my_begin inside the loop, i.e. for all the
      // atomics of the previous iteration, on every iteration.
      __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const uint32_t stage_end = min(s1, L.begin[kWave]);
      const unsigned long long short_bits =
          __ballot(ci < ncut && (t_box || short_form_ok<kOp>(L.begin[lane + 1] - my_begin, t_min_y, t_w, t_iw)));
      // ---- every sample of [s, stage_end) with all lanes
      sweep_samples<kMode, kOp, kBatch>(L, a.view0.v.width, kMode == RasterMode::Shadow ? sh.map : a.view0.depth,
                                        a.view0.idkey, sh.L, my_begin, s, stage_end, c0, short_bits, ent, nent);
      __builtin_amdgcn_wave_barrier();
      s = stage_end;
      c0 += kWave;
    }
    if (ent && lane == 0) Ls.ent.cnt[chunk] = nent;
  }
}

int device_cu_count();

template <RasterMode kMode>
static hipError_t launch_camera_raster(const PassArgs& a, int blocks_per_cu, hipStream_t s) {
  const dim3 grid(device_cu_count() * blocks_per_cu), block(256);
  if (a.views)
    hipLaunchKernelGGL((k_raster<kMode, kOpSize, true>), grid, block, 0, s, a, ShadowTarget{nullptr, 0});
  else
    hipLaunchKernelGGL((k_raster<kMode, kOpSize, false>), grid, block, 0, s, a, ShadowTarget{nullptr, 0});
  return hipGetLastError();
}

hipError_t launch_depth(const PassArgs& a, int blocks_per_cu, hipStream_t s) {
  return launch_camera_raster<RasterMode::Depth>(a, blocks_per_cu, s);
}

// kernel2's "last writer wins" is made deterministic: among fragments inside
// the BUF_ERROR window the pixel keeps the largest (2*tri + sub) key, which
// is the fragment a sequential in-order execution writes last.
hipError_t launch_id(const PassArgs& a, int blocks_per_cu, hipStream_t s) {
  return launch_camera_raster<RasterMode::Id>(a, blocks_per_cu, s);
}

// ---------------------------------------------------------------- shadows
struct FaceTable { SinCos f[6]; };

// prearrange_realtime_shadowing (cl2.cl:4419-4636) and kernel1's shadow
// raster (cl2.cl:4638-4700) for a group of lights in one launch.  Phase 1:
// one triangle per work item — its world-space vertices into LDS, read from
// HBM once for the whole group.  Then for each light of the group: the faces
// the vertices fall in (ret_cubeface), each (triangle, face) pair appended
// to a block-local work list in LDS, and phase 2: the pairs, one per work
// item per round, clipped, projected, culled and set up for their face.
// Pieces become records, key = slot * 6 + face with slot the light's cube
// map in the scene's map array, for one raster launch over the whole group;
// pieces of at most local_max samples, and any that no longer fit the list,
// are rasterised at once by the wavefront, their samples spread over its
// lanes.  (Looping over the six faces per
// triangle would leave most lanes idle: a triangle touches one to three
// faces.)
constexpr int kMaxShadowGroup = 16;
struct ShadowGroup {
  float pos[kMaxShadowGroup][3];
  uint32_t slot[kMaxShadowGroup];
  uint32_t n;
};

// Shadow-pass allocation for the lanes with `take` set, per wavefront: one
// atomicAdd on the packed counter of the block's partition (records
// [part * cap, part * cap + cap) of the list, cap = cap_cut / kShadowParts,
// and a sample space of its own).  An allocation that passes the
// partition's capacity or kShadowMaxSamples is not written: the wavefront
// rasterises those pieces itself and lowers the partition's valid-prefix
// word to its start, and every later allocation in the partition fails too
// (the counter only grows), so the raster launch reads each partition up to
// min(counter, valid prefix).
__device__ __forceinline__ bool wave_alloc_fit(const Lists& L, uint32_t ncut, uint32_t nsmp, uint32_t& cut0,
                                               uint32_t& smp0) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t ic = wave_incl_scan(ncut), is = wave_incl_scan(nsmp);
  const uint32_t tc = __shfl(ic, kWave - 1, kWave), ts = __shfl(is, kWave - 1, kWave);
  cut0 = smp0 = 0;
  if (tc == 0) return true;
  const uint32_t part = blockIdx.x % kShadowParts, cap = L.cap_cut / kShadowParts;
  uint32_t base_lo = 0, base_hi = 0, ok = 0;
  if (lane == 0) {
    unsigned long long* p = reinterpret_cast<unsigned long long*>(L.counters + kPartWord(part));
    const unsigned long long seen = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((seen >> kShadowSampleBits) < cap && (seen & kShadowSampleMask) < kShadowMaxSamples) {
      const unsigned long long old = atomicAdd(p, ((unsigned long long)tc << kShadowSampleBits) + ts);
      ok = (old >> kShadowSampleBits) + tc <= cap && (old & kShadowSampleMask) + ts <= kShadowMaxSamples;
      if (!ok) atomicMin(p + 1, old);   // the valid-prefix word, same line
      base_lo = (uint32_t)(old & kShadowSampleMask);   // below 2^30 when ok
      base_hi = (uint32_t)(old >> kShadowSampleBits);
    }
  }
  ok = __shfl(ok, 0, kWave);
  cut0 = part * cap + __shfl(base_hi, 0, kWave) + ic - ncut;
  smp0 = __shfl(base_lo, 0, kWave) + is - nsmp;
  return ok != 0;
}

using ShadowStage = TriLds<RasterMode::Shadow, false>;

// Rasterises the piece each lane holds (n samples, 0 for none) with the
// whole wavefront, straight into the maps (called wave-uniformly).  A map
// keeps the minimum depth of its samples, so the order in which samples
// arrive does not matter.
__device__ __forceinline__ void local_sweep(ShadowStage& W, const RasterTri& rt, uint32_t n, uint32_t key,
                                            const ShadowTarget& sh) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t incl = wave_incl_scan(n);   // n < 2^24 per lane: the sum fits 32 bits
  const uint32_t total = __shfl(incl, kWave - 1, kWave);
  if (total == 0) return;
  const unsigned long long m = __ballot(n > 0);
  const uint32_t pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  const uint32_t nloc = (uint32_t)__popcll(m);
  W.start[lane] = 0xFFFFFFFFu;   // sample indices of this sweep are below 2^31
  if (n) {
    W.min_x[pos] = rt.min_x;
    W.min_y[pos] = rt.min_y;
    W.w[pos] = rt.w;
    W.iw[pos] = rt.iw;
    W.pa[pos] = rt.A;
    W.pb[pos] = rt.B;
    W.pc[pos] = rt.C;
#pragma unroll
    for (int e = 0; e < 7; e++) W.ek[e][pos] = rt.ek[e];
    W.key[pos] = key;
    W.begin[pos] = incl - n;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint32_t my_begin = (uint32_t)lane < nloc ? W.begin[lane] : total;
  const uint32_t my_end = (uint32_t)lane + 1 < nloc ? W.begin[lane + 1] : total;
  // the box-form flag of staged triangle p to lane p (lanes without a
  // piece push theirs to distinct lanes >= nloc)
  const uint32_t to = n ? pos : nloc + ((uint32_t)lane - pos);
  const int box = __builtin_amdgcn_ds_permute((int)(to * 4u), n > 0 && (rt.key & kKeyBoxForm) != 0u ? 1 : 0);
  const unsigned long long short_bits =
      __ballot((uint32_t)lane < nloc &&
               (box != 0 || short_form_ok<kOpSizeLight>(my_end - my_begin, W.min_y[lane], W.w[lane], W.iw[lane])));
  uint32_t nent_unused = 0;
  sweep_samples<RasterMode::Shadow, kOpSizeLight, false>(W, 0u, sh.map, nullptr, sh.L, my_begin, 0u, total, 0u,
                                                         short_bits, nullptr, nent_unused);
  // the stage is rewritten by the wavefront's next sweep
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct ShadowLds {
  float wpos[kShadowPreThreads][9];
  // (local triangle << 4 | two_sided << 3 | face) work items, one list per
  // light, two lists alternating
  uint32_t items[2][kShadowPreThreads * 3];
  uint32_t n_items[kMaxShadowGroup];
};

__global__ void __launch_bounds__(kShadowPreThreads) k_prearrange_shadow(const TriPos* __restrict__ world, uint32_t ntri,
                                                           const ObjDesc* __restrict__ objs, uint32_t nobj,
                                                           ShadowGroup group, ShadowTarget target,
                                                           int only_static, uint32_t local_max, FaceTable ft,
                                                           Lists lists) {
  __shared__ ShadowLds S;
  __shared__ ShadowStage stage_all[kShadowPreThreads / kWave];
  ShadowStage& W = stage_all[threadIdx.x / kWave];
  const uint32_t t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const uint32_t wave = t / kWave;
  const uint32_t id = blockIdx.x * blockDim.x + t;
  const uint32_t L = target.L;
  const float ew = (float)L, eh = (float)L, efov = (float)L / 2.0f;
  // ---- phase 1: the triangle, once for every light of the group
  bool live = false;
  uint32_t two_sided = 0;
  f3 wp = mk3(0.f, 0.f, 0.f);
  if (id < ntri) {
    const TriPos T = world[id];
    if (T.oid < nobj) {
      const ObjDesc G = objs[T.oid];
      const bool is_static = (G.flags & OCR_FEATURE_IS_STATIC) > 0;
      live = only_static ? is_static : !is_static;
      wp = mk3(G.wp[0], G.wp[1], G.wp[2]);
      if (live) {
#pragma unroll
        for (int k = 0; k < 9; k++) S.wpos[t][k] = T.p[k];
        two_sided = (G.flags & OCR_FEATURE_TWO_SIDED) > 0 ? 1u : 0u;
      }
    }
  }
  if (t < (uint32_t)kMaxShadowGroup) S.n_items[t] = 0;
  __syncthreads();
  for (uint32_t li = 0; li < group.n; li++) {
    const f3 lpos = mk3(group.pos[li][0], group.pos[li][1], group.pos[li][2]);
    const uint32_t key0 = group.slot[li] * 6u;
    uint32_t* const items = S.items[li & 1u];
    // Light li's items go to list li & 1.  Every wavefront has finished
    // light li - 2's rounds on that list before it passed light li - 1's
    // barrier, and writes here come after that barrier.
    uint32_t touched = 0;
    if (live && !(fast_length(wp - lpos) > kDepthFar)) {
#pragma unroll
      for (int k = 0; k < 3; k++)
        touched |= 1u << cubeface(mk3(S.wpos[t][3 * k], S.wpos[t][3 * k + 1], S.wpos[t][3 * k + 2]), lpos);
    }
    const uint32_t mine = (uint32_t)__popc(touched);
    const uint32_t incl = wave_incl_scan(mine);
    const uint32_t wave_total = __shfl(incl, kWave - 1, kWave);
    uint32_t wbase = 0;
    if (lane == 0 && wave_total) wbase = atomicAdd(&S.n_items[li], wave_total);
    wbase = __shfl(wbase, 0, kWave);
    uint32_t slot = wbase + incl - mine;
    for (uint32_t f = 0; f < 6; f++)
      if ((touched >> f) & 1u) items[slot++] = t << 4 | two_sided << 3 | f;
    __syncthreads();
    const uint32_t n_items = S.n_items[li];
    // ---- phase 2: 64 work items per wavefront per round, the wavefronts
    // independent of each other
    for (uint32_t r0 = wave * kWave; r0 < n_items; r0 += kShadowPreThreads) {
      f3 tp[2][3];
      float rc[2] = {0.f, 0.f};
      RasterTri rt[2];
      uint32_t nf[2] = {0u, 0u}, ns[2] = {0u, 0u};
      uint32_t face = 0;
      if (r0 + lane < n_items) {
        const uint32_t it = items[r0 + lane];
        const uint32_t lt = it >> 4;
        face = it & 7u;
        f3 w[3];
#pragma unroll
        for (int k = 0; k < 3; k++) w[k] = mk3(S.wpos[lt][3 * k], S.wpos[lt][3 * k + 1], S.wpos[lt][3 * k + 2]);
        const int num = face_clip_project(w, lpos, ft.f[face], efov, ew, eh, tp);
        float my[2];
        cut_pieces(tp, num, ew, eh, ((it >> 3) & 1u) != 0u, kOpSizeLight, rc, rt, nf, ns, my);
      }
      const uint32_t key = key0 + face;
      // pieces above local_max samples to the list, in allocations of at
      // most kShadowWaveSamples samples (each takes the lanes whose running
      // sum fits, at least the first waiting lane); pieces of the list's
      // failed allocations and the small ones rasterised here
      const bool big0 = ns[0] > local_max, big1 = ns[1] > local_max;
      const uint32_t mine = (big0 ? ns[0] : 0u) + (big1 ? ns[1] : 0u);   // < 2^25
      bool waiting = mine > 0, fit = false;
      while (__ballot(waiting)) {
        const bool take = waiting && wave_incl_scan(waiting ? mine : 0u) <= kShadowWaveSamples;
        uint32_t c, smp;
        const bool ok = wave_alloc_fit(lists, take ? (big0 ? 1u : 0u) + (big1 ? 1u : 0u) : 0u, take ? mine : 0u,
                                       c, smp);
        if (take) {
          waiting = false;
          fit = ok;
          if (ok) {
#pragma unroll
            for (int i = 0; i < 2; i++) {
              if (!(i == 0 ? big0 : big1)) continue;
              rt[i].key = key | (rt[i].key & kKeyBoxForm);
              rt[i].sample_begin = smp;
              lists.rast[c] = rt[i];
              smp += ns[i];
              c++;
            }
          }
        }
      }
      uint32_t nl[2];
      nl[0] = fit && big0 ? 0u : ns[0];
      nl[1] = fit && big1 ? 0u : ns[1];
      local_sweep(W, rt[0], nl[0], key, target);
      local_sweep(W, rt[1], nl[1], key, target);
    }
  }
}

__global__ void k_zero_counters(uint32_t* c) {
  if (threadIdx.x < kCounterWords)
    c[threadIdx.x] = threadIdx.x == kValidEnd || threadIdx.x == kValidEnd + 1 ? 0xFFFFFFFFu : 0u;
}

// A shadow pass's counters: the render words and every partition's packed
// counter (0) and valid prefix (all ones).
__global__ void __launch_bounds__(256) k_zero_shadow_counters(uint32_t* c) {
  for (uint32_t i = threadIdx.x; i < (uint32_t)kShadowCounterWords; i += blockDim.x) {
    const uint32_t w = i < (uint32_t)kCounterWords ? 0u : (i - kCounterWords) % kPartLineWords;
    c[i] = i >= (uint32_t)kCounterWords && (w == 2u || w == 3u) ? 0xFFFFFFFFu : 0u;
  }
}

hipError_t launch_zero_counters(uint32_t* counters, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, counters);
  return hipGetLastError();
}

// Tuning of the shadow launches, read at every pass: pieces of at most
// OCR_SHADOW_LOCAL_SAMPLES samples (default 16) are rasterised by the
// prearrange launch itself, larger ones go to the list for the chip-wide
// raster launch; OCR_SHADOW_LIST_RECORDS lowers the list's capacity (the
// pieces that do not fit are rasterised by the prearrange launch too).
// Neither changes the maps, only where their samples are rasterised.
static uint32_t env_u32(const char* name, uint32_t dflt) {
  const char* e = getenv(name);
  if (!e || !*e) return dflt;
  char* end = nullptr;
  const unsigned long v = strtoul(e, &end, 10);
  return end && *end == 0 && v <= 0xFFFFFFFFul ? (uint32_t)v : dflt;
}
constexpr uint32_t kShadowLocalSamples = 16;

hipError_t launch_shadow_group(Scene* sc, const float (*lpos)[3], const uint32_t* slots, uint32_t n,
                               int only_static, uint32_t* maps, hipStream_t s) {
  if (n == 0 || n > (uint32_t)kMaxShadowGroup) return hipErrorInvalidValue;
  FaceTable ft;
  for (int i = 0; i < 6; i++) ft.f[i] = sc->ctx->face_sc[i];
  ShadowGroup g;
  memset(&g, 0, sizeof(g));
  for (uint32_t i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) g.pos[i][k] = lpos[i][k];
    g.slot[i] = slots[i];
  }
  g.n = n;
  // raster setups only: shadow records have no CutTri
  Lists lists{nullptr, sc->sh_rast, sc->sh_counters,
              std::min(sc->sh_cap_cut, env_u32("OCR_SHADOW_LIST_RECORDS", sc->sh_cap_cut)), 0u};
  const ShadowTarget target{maps, sc->lbuf_dim};
  hipLaunchKernelGGL(k_zero_shadow_counters, dim3(1), dim3(256), 0, s, sc->sh_counters);
  if (sc->ntri)
    hipLaunchKernelGGL(k_prearrange_shadow, dim3((sc->ntri + kShadowPreThreads - 1) / kShadowPreThreads),
                       dim3(kShadowPreThreads), 0, s, sc->world, sc->ntri, sc->objs, sc->nobj, g, target,
                       only_static, env_u32("OCR_SHADOW_LOCAL_SAMPLES", kShadowLocalSamples), ft, lists);
  PassArgs a;
  memset(&a, 0, sizeof(a));
  a.lists = lists;
  hipLaunchKernelGGL((k_raster<RasterMode::Shadow, kOpSizeLight, false>), dim3(device_cu_count() * 8), dim3(256), 0, s,
                     a, target);
  return hipGetLastError();
}

}  // namespace ocr
