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

// Prearrange blocks: 1024 work items, one triangle each.  Cut records and
// sample ranges are allocated once per block (one 64-bit atomicAdd on the
// packed counter (cuts << 32 | samples) per 1024 triangles): every wavefront
// of the render appending to one counter would serialise on that address.
constexpr int kPreThreads = 1024;
constexpr int kPreWaves = kPreThreads / kWave;

struct BlockAllocLds {
  uint32_t cuts[kPreWaves], smps[kPreWaves], frags[kPreWaves];
  unsigned long long base;
  uint32_t ok;
};

// Returns false (for the whole block) when the block's records do not fit the
// list or the sample space; the overflow word then says which.  The block's
// sample total is checked before its atomic, and each atomic adds less than
// kMaxSamples, so the low half reaches the threshold and sets the flag before
// it could carry into the record count.
__device__ __forceinline__ bool block_alloc(const Lists& L, BlockAllocLds& sh, uint32_t ncut, uint32_t nsmp,
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
    for (int w = 0; w < kPreWaves; w++) {
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
    if (tf) atomicAdd(&L.counters[kFragCount], tf);
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

// The walk's row of sample q: floor(mad(q, 1/width, min_y)) (cl2.cl:5083).
__device__ __forceinline__ float row_of(int q, float iw, float min_y) {
  return floorf(fmaf((float)q, iw, min_y));
}

// First q >= 0 with row_of(q) >= Y, stepping from the nominal (Y - min_y) * w.
__device__ __forceinline__ int run_start(float Y, int w, float iw, float min_y) {
  long long n0 = (long long)(Y - min_y) * (long long)w;
  int q = n0 < 0 ? 0 : (int)n0;
  for (int g = 0; g < 64 && q > 0 && row_of(q - 1, iw, min_y) >= Y; g++) q--;
  for (int g = 0; g < 64 && row_of(q, iw, min_y) < Y; g++) q++;
  return q;
}

// Samples of all fragments of a cut triangle.  Fragment a visits
// q in [op*a, op*a + op] up to the first q whose row is >= max_y (q_end), so
// it holds clamp(q_end - op*a, 0, op + 1) samples; the full fragments come
// first, then at most one partial one.
__device__ __forceinline__ uint32_t triangle_samples(const f3 p[3], float ew, float eh, int op, uint32_t nfrag) {
  if (nfrag == 0) return 0;
  const float xs[3] = {p[0].x, p[1].x, p[2].x}, ys[3] = {p[0].y, p[1].y, p[2].y};
  float mm[4];
  calc_min_max(xs, ys, ew, eh, mm);
  const int w = (int)(mm[1] - mm[0]);
  if (w <= 0) return 0;
  const float iw = 1.f / (float)w;
  const long long q_end = run_start(mm[3], w, iw, mm[2]);
  long long m = q_end >= op + 1 ? (q_end - (op + 1)) / op + 1 : 0;
  if (m > (long long)nfrag) m = nfrag;
  long long s = (long long)(op + 1) * m;
  if (m < (long long)nfrag) {
    const long long r = q_end - (long long)op * m;
    s += r < 0 ? 0 : (r > op + 1 ? op + 1 : r);
  }
  return s > (long long)kMaxTriSamples ? kMaxTriSamples : (uint32_t)s;
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
  float n = ceilf(area / (float)op_size);
  return n > 0.f ? (uint32_t)n : 0u;
}

__device__ __forceinline__ void store_cut(CutTri* dst, const f3 p[3], float rconst, uint32_t key, uint32_t oid,
                                          uint32_t sample_begin, uint32_t nsamples, uint32_t view) {
  CutTri c;
#pragma unroll
  for (int i = 0; i < 3; i++) { c.x[i] = p[i].x; c.y[i] = p[i].y; c.z[i] = p[i].z; }
  c.rconst = rconst;
  c.key = key;
  c.oid = oid;
  c.sample_begin = sample_begin;
  c.nsamples = nsamples;
  c.view = view;
  c._pad = 0;
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

// Appends this work item's cut triangles that hold samples (0..2) with their
// sample ranges; called by every work item of the block.  A block that would
// pass the list capacity or the sample limit writes nothing and sets the
// overflow word, which every later pass of the render checks before reading
// the list.
__device__ __forceinline__ void append_cuts(const Lists& L, BlockAllocLds& sh, const f3 tp[2][3],
                                            const float rc[2], const uint32_t nf[2], const uint32_t ns[2],
                                            uint32_t key0, uint32_t key_step, uint32_t oid, uint32_t view) {
  const uint32_t k0 = ns[0] > 0, k1 = ns[1] > 0;
  uint32_t c, smp;
  if (!block_alloc(L, sh, k0 + k1, ns[0] + ns[1], nf[0] * k0 + nf[1] * k1, c, smp)) return;
#pragma unroll
  for (int i = 0; i < 2; i++) {
    if (!ns[i]) continue;
    store_cut(&L.cuts[c], tp[i], rc[i], key0 + key_step * (uint32_t)i, oid, smp, ns[i], view);
    smp += ns[i];
    c++;
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
__global__ void __launch_bounds__(kPreThreads) k_prearrange(PassArgs a) {
  __shared__ BlockAllocLds sh;
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
  uint32_t nf[2] = {0u, 0u}, ns[2] = {0u, 0u};
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
          if (i < num) {
            nf[i] = cull_count(tp[i], ew, eh, two_sided, kOpSize, rc[i]);
            ns[i] = triangle_samples(tp[i], ew, eh, kOpSize, nf[i]);
          }
      }
    }
  }
  append_cuts(a.lists, sh, tp, rc, nf, ns, 2u * id, 1u, oid, vi);
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

constexpr uint32_t kChunk = kWave * 8;   // samples per wavefront work item

enum class RasterMode { Depth, Id, Shadow };

struct ShadowTarget {
  uint32_t* map;   // 6 faces of L x L
  uint32_t L;
};

// Setup of 64 staged triangles, one column per lane.  point_in_tri's
// loop-invariant sub-expressions are kept with the edge sign folded in
// (negation is exact, so (a + b*x + c*y) * sign has the same bits as
// a*sign + (b*sign)*x + (c*sign)*y).
template <RasterMode kMode>
struct TriLds {
  float min_x[kWave], max_x[kWave], min_y[kWave];
  float pa[kWave], pb[kWave], pc[kWave], iw[kWave];   // 1/z plane, 1/width
  float ek[7][kWave];   // s0, s1, s2, t0, t1, t2 (signed), 2.0001 * |area|
  int w[kWave];
  uint32_t key[kWave];
  uint32_t begin[kWave + 1];
  uint32_t pitch[kMode == RasterMode::Shadow ? 1 : kWave];
  uint32_t* dst[kWave];                                // depth buffer / cube face
  unsigned long long* ids[kMode == RasterMode::Id ? kWave : 1];
};

// x mod w for 0 <= x < 2^25, 1 <= w: a float quotient estimate (|error| <= 1
// for these ranges) corrected by one step each way.
__device__ __forceinline__ int mod_small(int x, int w, float invw) {
  const int qq = (int)((float)x * invw);
  int r = x - qq * w;
  if (r < 0) r += w;
  if (r >= w) r -= w;
  return r;
}

template <RasterMode kMode, int kOp>
__global__ void __launch_bounds__(256) k_raster(PassArgs a, ShadowTarget sh) {
  __shared__ TriLds<kMode> lds_all[256 / kWave];
  const Lists& Ls = a.lists;
  if (Ls.counters[kOverflow]) return;
  const unsigned long long packed = *reinterpret_cast<const unsigned long long*>(Ls.counters + kPacked);
  const uint32_t ncut = (uint32_t)(packed >> 32);
  const uint32_t total = (uint32_t)packed;
  const int lane = threadIdx.x & (kWave - 1);
  TriLds<kMode>& L = lds_all[threadIdx.x / kWave];
  const uint32_t wave0 = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  for (uint32_t s0 = wave0 * kChunk; s0 < total; s0 += nwaves * kChunk) {
    const uint32_t s1 = min(s0 + kChunk, total);
    // ---- the chunk's first triangle: last record with sample_begin <= s0
    // (record 0 begins at sample 0; 64-ary search, one probe per lane)
    uint32_t lo = 0, hi = ncut;
    while (hi - lo > 1) {
      const uint32_t p = lo + (uint32_t)(((unsigned long long)(hi - lo) * (uint32_t)lane) / kWave);
      const unsigned long long m = __ballot(Ls.cuts[p].sample_begin <= s0);
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
      if (ci < ncut) {
        const CutTri c = Ls.cuts[ci];
        uint32_t W, H;
        if constexpr (kMode == RasterMode::Shadow) {
          W = H = sh.L;
          L.dst[lane] = sh.map + (size_t)c.key * sh.L * sh.L;
        } else {
          const DevView& V = view_at(a, c.view);
          W = V.v.width;
          H = V.v.height;
          L.dst[lane] = V.depth;
          L.pitch[lane] = W;
          if constexpr (kMode == RasterMode::Id) L.ids[lane] = V.idkey;
        }
        Setup st;
        raster_setup(c, (float)W, (float)H, st);
        L.min_x[lane] = st.mm[0];
        L.max_x[lane] = st.mm[1];
        L.min_y[lane] = st.mm[2];
        L.pa[lane] = st.A;
        L.pb[lane] = st.B;
        L.pc[lane] = st.C;
        L.iw[lane] = st.iwidth;
        const float p0x = st.xp[0], p0y = st.yp[0], p1x = st.xp[1], p1y = st.yp[1], p2x = st.xp[2], p2y = st.yp[2];
        const float A = 0.5f * (-p1y * p2x + p0y * (-p1x + p2x) + p0x * (p1y - p2y) + p1x * p2y);
        const float sign = A < 0 ? -1.f : 1.f;
        L.ek[0][lane] = (p0y * p2x - p0x * p2y) * sign;
        L.ek[1][lane] = (p2y - p0y) * sign;
        L.ek[2][lane] = (p0x - p2x) * sign;
        L.ek[3][lane] = (p0x * p1y - p0y * p1x) * sign;
        L.ek[4][lane] = (p0y - p1y) * sign;
        L.ek[5][lane] = (p1x - p0x) * sign;
        L.ek[6][lane] = 2.0001f * A * sign;
        L.w[lane] = st.width;
        L.key[lane] = c.key;
        L.begin[lane] = c.sample_begin;
        if (lane == kWave - 1) L.begin[kWave] = c.sample_begin + c.nsamples;
      } else {
        L.begin[lane] = total;
        if (lane == kWave - 1) L.begin[kWave] = total;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const uint32_t stage_end = min(s1, L.begin[kWave]);
      // ---- every sample of [s, stage_end) with all lanes
      for (uint32_t j = s + lane; j < stage_end; j += kWave) {
        int k = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1)
          if (L.begin[k + step] <= j) k += step;
        const uint32_t sl = j - L.begin[k];
        const uint32_t fa = sl / (uint32_t)(kOp + 1);
        const int pa = kOp * (int)fa;
        const int q = pa + (int)(sl - fa * (uint32_t)(kOp + 1));
        const int w = L.w[k];
        const float min_y = L.min_y[k], iw = L.iw[k];
        const float y = row_of(q, iw, min_y);
        int r = run_start(y, w, iw, min_y);
        r = r > pa ? r : pa;
        const float x = (float)mod_small(r, w, iw) + L.min_x[k] + (float)(q - r);
        if (x >= L.max_x[k]) continue;
        const float s_ = L.ek[0][k] + L.ek[1][k] * x + L.ek[2][k] * y;
        const float t_ = L.ek[3][k] + L.ek[4][k] * x + L.ek[5][k] * y;
        if (!(s_ > -0.0001f && t_ > -0.0001f && (s_ + t_) < L.ek[6][k])) continue;
        const float fd = fmaf(L.pa[k], x, fmaf(L.pb[k], y, L.pc[k]));
        const uint32_t d = sat_u32(kMulint / fd);
        const uint32_t pitch = kMode == RasterMode::Shadow ? sh.L : L.pitch[kMode == RasterMode::Shadow ? 0 : k];
        const size_t pix = (size_t)(int)y * pitch + (size_t)(int)x;
        if constexpr (kMode == RasterMode::Id) {
          const uint32_t val = L.dst[k][pix];
          if (d > val - kBufError && d < val + kBufError)
            atomicMax(&L.ids[k][pix], ((unsigned long long)(L.key[k] + 1u) << 32) | (c0 + (uint32_t)k));
        } else {
          atomicMin(&L.dst[k][pix], d);
        }
      }
      __builtin_amdgcn_wave_barrier();
      s = stage_end;
      c0 += kWave;
    }
  }
}

int device_cu_count();

hipError_t launch_depth(const PassArgs& a, int blocks_per_cu, hipStream_t s) {
  hipLaunchKernelGGL((k_raster<RasterMode::Depth, kOpSize>), dim3(device_cu_count() * blocks_per_cu), dim3(256), 0,
                     s, a, ShadowTarget{nullptr, 0});
  return hipGetLastError();
}

// kernel2's "last writer wins" is made deterministic: among fragments inside
// the BUF_ERROR window the pixel keeps the largest (2*tri + sub) key, which
// is the fragment a sequential in-order execution writes last.
hipError_t launch_id(const PassArgs& a, int blocks_per_cu, hipStream_t s) {
  hipLaunchKernelGGL((k_raster<RasterMode::Id, kOpSize>), dim3(device_cu_count() * blocks_per_cu), dim3(256), 0,
                     s, a, ShadowTarget{nullptr, 0});
  return hipGetLastError();
}

// ---------------------------------------------------------------- shadows
struct FaceTable { SinCos f[6]; };

__global__ void __launch_bounds__(kPreThreads) k_prearrange_shadow(const TriPos* __restrict__ tris, uint32_t ntri,
                                                           const ObjDesc* __restrict__ objs, uint32_t nobj,
                                                           float lx, float ly, float lz, uint32_t L,
                                                           int only_static, FaceTable ft, Lists lists) {
  __shared__ BlockAllocLds sh;
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
  const bool two_sided = live && (G.flags & OCR_FEATURE_TWO_SIDED) > 0;
  // Faces one at a time; every work item of the block takes part in each
  // face's allocation (one that does not touch the face appends nothing).
  __shared__ int faces_touched;
  if (threadIdx.x == 0) faces_touched = 0;
  __syncthreads();
  if (touched) atomicOr(&faces_touched, touched);
  __syncthreads();
  const int any_touched = faces_touched;
#pragma unroll
  for (int kk = 0; kk < 6; kk++) {
    if (!((any_touched >> kk) & 1)) continue;
    f3 tp[2][3];
    float rc[2] = {0.f, 0.f};
    uint32_t nf[2] = {0u, 0u}, ns[2] = {0u, 0u};
    if ((touched >> kk) & 1) {
      int num = transform_clip_project(T, G, lpos, ft.f[kk], efov, ew, eh, tp);
#pragma unroll
      for (int i = 0; i < 2; i++)
        if (i < num) {
          nf[i] = cull_count(tp[i], ew, eh, two_sided, kOpSizeLight, rc[i]);
          ns[i] = triangle_samples(tp[i], ew, eh, kOpSizeLight, nf[i]);
        }
    }
    append_cuts(lists, sh, tp, rc, nf, ns, (uint32_t)kk, 0u, T.oid, 0u);
  }
}

__global__ void k_zero_counters(uint32_t* c) {
  if (threadIdx.x < kCounterWords) c[threadIdx.x] = 0u;
}

hipError_t launch_shadow_light(Scene* sc, const float lpos[3], int only_static, uint32_t* map,
                               hipStream_t s) {
  FaceTable ft;
  for (int i = 0; i < 6; i++) ft.f[i] = sc->ctx->face_sc[i];
  const Lists lists{sc->sh_cuts, sc->sh_counters, sc->sh_cap_cut};
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, sc->sh_counters);
  if (sc->ntri)
    hipLaunchKernelGGL(k_prearrange_shadow, dim3((sc->ntri + kPreThreads - 1) / kPreThreads), dim3(kPreThreads), 0, s, sc->tri_pos,
                       sc->ntri, sc->objs, sc->nobj, lpos[0], lpos[1], lpos[2], sc->lbuf_dim, only_static,
                       ft, lists);
  PassArgs a;
  memset(&a, 0, sizeof(a));
  a.lists = lists;
  hipLaunchKernelGGL((k_raster<RasterMode::Shadow, kOpSizeLight>), dim3(device_cu_count() * 8), dim3(256), 0, s,
                     a, ShadowTarget{map, sc->lbuf_dim});
  return hipGetLastError();
}

}  // namespace ocr
