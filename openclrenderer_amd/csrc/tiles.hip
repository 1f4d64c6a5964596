// tiles.hip — the tiled raster path: kernel1's depth pass and kernel2's id
// pass resolved per screen tile in LDS, from bins of the cut records whose
// bounding boxes touch each tile.
//
// Reference: split_into_tiled_chunks and tile_render (cl2.cl:3850-4210),
// run by engine::draw_tiled_deferred (engine.cpp:2447-2535).  That path
// bins triangles to tile display lists and keeps each tile's depth in a
// __local buffer before an atomic_min into the global one; its own coverage
// test differs from kernel1's and its chunk allocation can drop triangles.
// Here the tile path computes exactly what the fragment walk of kernel1 /
// kernel2 (cl2.cl:4985-5127, :5390-5546) computes, bit for bit:
//   bin_sums / bin_offsets  exclusive scan of the tiles' entry counts
//                           (counted by prearrange, raster.hip);
//   bin_scatter             each cut record's index into the bins of the
//                           tiles its box touches;
//   tile_raster             one workgroup per tile: for the tile's records,
//                           the walk's samples that fall in the tile (the
//                           box clipped to the tile when the walk is the box,
//                           walk_is_rect; the walk's own samples of the
//                           tile's rows otherwise), u32 atomicMin into an LDS
//                           depth tile, then the same samples again against
//                           the final depth with the ±BUF_ERROR window and a
//                           u64 atomicMax on (2*tri + piece + 1) << 32 |
//                           record into an LDS id tile; the tile's depth and
//                           ids are written once.  No clear pass is needed.
#pragma clang fp contract(off)

#include <algorithm>

#include "ocr_internal.h"
#include "tiles.h"
#include "walk.h"

namespace ocr {

constexpr int kScanThreads = 1024;
constexpr uint32_t kScanChunk = kScanThreads * 4;   // tiles per scan block

// ------------------------------------------------------------------ reset
// Zeroes the render's counters and every tile's entry count.
__global__ void __launch_bounds__(256) k_bin_reset(uint32_t* __restrict__ counters, uint32_t* __restrict__ cnt,
                                                   uint32_t ntiles) {
  if (blockIdx.x == 0 && threadIdx.x < kCounterWords) counters[threadIdx.x] = 0u;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < ntiles; i += gridDim.x * 256u) cnt[i] = 0u;
}

// ------------------------------------------------------------------- scan
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v, unsigned long long* sh) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  unsigned long long t = 0;
  for (int w = 0; w < kScanThreads / kWave; w++) t += sh[w];
  __syncthreads();
  return t;
}

// Entries of each chunk of kScanChunk tiles.
__global__ void __launch_bounds__(kScanThreads) k_bin_sums(TileBins tb, const uint32_t* __restrict__ counters) {
  __shared__ unsigned long long sh[kScanThreads / kWave];
  if (counters[kOverflow]) return;
  const uint32_t t0 = blockIdx.x * kScanChunk + threadIdx.x * 4u;
  unsigned long long v = 0;
#pragma unroll
  for (uint32_t i = 0; i < 4; i++)
    if (t0 + i < tb.ntiles) v += tb.cnt[t0 + i];
  v = block_sum_u64(v, sh);
  if (threadIdx.x == 0) reinterpret_cast<unsigned long long*>(tb.sums)[blockIdx.x] = v;
}

// Exclusive offsets of the tiles of one chunk (the chunks before it summed
// first); the last chunk records the total, and the overflow bit when the
// bins cannot hold it.
__global__ void __launch_bounds__(kScanThreads) k_bin_offsets(TileBins tb, uint32_t* __restrict__ counters) {
  __shared__ unsigned long long sh[kScanThreads / kWave];
  __shared__ unsigned long long wsum[kScanThreads / kWave];
  if (counters[kOverflow]) return;
  const unsigned long long* sums = reinterpret_cast<const unsigned long long*>(tb.sums);
  unsigned long long before = 0;
  for (uint32_t i = threadIdx.x; i < blockIdx.x; i += kScanThreads) before += sums[i];
  before = block_sum_u64(before, sh);
  const uint32_t t0 = blockIdx.x * kScanChunk + threadIdx.x * 4u;
  uint32_t c[4];
#pragma unroll
  for (uint32_t i = 0; i < 4; i++) c[i] = t0 + i < tb.ntiles ? tb.cnt[t0 + i] : 0u;
  const unsigned long long mine = (unsigned long long)c[0] + c[1] + c[2] + c[3];
  // inclusive scan over the block: within the wavefront, then across them
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  unsigned long long inc = mine;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const unsigned long long u = __shfl_up(inc, d, kWave);
    if (lane >= d) inc += u;
  }
  if (lane == kWave - 1) wsum[wave] = inc;
  __syncthreads();
  unsigned long long wbase = 0;
  for (int w = 0; w < wave; w++) wbase += wsum[w];
  unsigned long long run = before + wbase + inc - mine;
#pragma unroll
  for (uint32_t i = 0; i < 4; i++) {
    if (t0 + i < tb.ntiles) {
      const uint32_t o = run > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)run;
      tb.off[t0 + i] = o;
      tb.cur[t0 + i] = o;
    }
    run += c[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) {
    // run is now the total of every tile
    tb.off[tb.ntiles] = run > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)run;
    counters[kBinTotal] = run > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)run;
    if (run > tb.cap_bin) atomicOr(&counters[kOverflow], kOverflowBins);
  }
}

// ---------------------------------------------------------------- scatter
__global__ void __launch_bounds__(256) k_bin_scatter(TileBins tb, const uint32_t* __restrict__ counters,
                                                     uint32_t cap_cut) {
  if (counters[kOverflow]) return;
  const uint32_t ncut = min((uint32_t)(*reinterpret_cast<const unsigned long long*>(counters + kPacked) >> 32), cap_cut);
  const uint32_t stride = gridDim.x * 256u;
  // whole wavefronts iterate together (the pair loop needs every lane)
  for (uint32_t base = blockIdx.x * 256u + (threadIdx.x & ~(uint32_t)(kWave - 1)); base < ncut; base += stride) {
    const uint32_t c = base + (threadIdx.x & (kWave - 1));
    const bool has = c < ncut;
    const uint4 b = has ? tb.box[c] : uint4{0u, 0u, 0u, 0u};
    uint32_t* cur = tb.cur;
    uint32_t* bin = tb.bin;
    const uint32_t cap = tb.cap_bin, ntiles = tb.ntiles;
    expand_tile_pairs(
        b, has, b, false,
        [cur, ntiles](uint32_t t, uint32_t len) { return t < ntiles ? atomicAdd(&cur[t], len) : 0xFFFFFFFFu; },
        [bin, cap, base](uint32_t, int owner, uint32_t first, uint32_t rank) {
          const uint32_t slot = first + rank;
          if (first != 0xFFFFFFFFu && slot < cap) bin[slot] = base + (uint32_t)owner;
        });
  }
}

// ------------------------------------------------------------ tile raster
// Setups of up to 64 records of a tile, one column per lane.  mode 0: the
// walk is the box, samples are the box clipped to the tile, row-major
// (x0, y0, rw, inv_rw); mode 1: the walk's samples ja .. of the tile's rows
// (sample_xy maps them, the tile's columns are kept); mode 2: nothing.
struct TileStage {
  float x0[kWave], y0[kWave], rw[kWave], inv[kWave];
  float min_x[kWave], min_y[kWave], w[kWave], iw[kWave];
  float pa[kWave], pb[kWave], pc[kWave];
  float ek[7][kWave];
  uint32_t ja[kWave], mode[kWave], key[kWave], rec[kWave];
  uint32_t begin[kWave + 1];
  uint32_t start[kWave];
};

// Four wavefronts share a tile's records: a distant object's sub-pixel
// triangles put thousands of records in one tile, and those tiles are the
// long ones.  (A second launch giving tiles of more than 256 records 16
// wavefronts each measured slower: 286 vs 193 us on the 1080p, 1 M-triangle
// frame.)
constexpr uint32_t kTileWaves = 4;

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stages records [e0, e0 + 64) of the bin (up to e1) for the tile whose
// pixels are [tx0, tx0 + tw) x [ty0, ty0 + th); returns the chunk's samples.
__device__ __forceinline__ uint32_t stage_chunk(TileStage& S, const Lists& Ls, uint32_t e0, uint32_t e1, float tx0,
                                                float ty0, float tw, float th, uint32_t ncut, uint32_t total) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t e = e0 + (uint32_t)lane;
  uint32_t ns = 0;
  if (e < e1) {
    const uint32_t c = Ls.tb.bin[e];
    uint32_t mode = 2;
    if (c < ncut) {
      const RasterTri t = Ls.rast[c];
      const uint32_t bw = Ls.tb.box[c].w;
      const float max_y = (float)(bw & 0xFFFFu);
      const float max_x = t.min_x + t.w;
      const float x0 = fmaxf(t.min_x, tx0), x1 = fminf(max_x, tx0 + tw);
      const float y0 = fmaxf(t.min_y, ty0), y1 = fminf(max_y, ty0 + th);
      if (bw & kBoxRect) {
        const float rw = x1 - x0;
        ns = (uint32_t)rw * (uint32_t)(y1 - y0);
        S.x0[lane] = x0;
        S.y0[lane] = y0;
        S.rw[lane] = rw;
        S.inv[lane] = __builtin_amdgcn_rcpf(rw);
        mode = 0;
      } else {
        const uint32_t sb = t.sample_begin;
        const uint32_t n = (c + 1 < ncut ? Ls.rast[c + 1].sample_begin : total) - sb;
        const uint32_t qa = (uint32_t)run_start(y0, t.w, t.iw, t.min_y);
        const uint32_t qb = (uint32_t)run_start(y1, t.w, t.iw, t.min_y);
        const uint32_t ja = min(first_sample<kOpSize>(qa), n), jb = min(first_sample<kOpSize>(qb), n);
        ns = jb - ja;
        S.ja[lane] = ja;
        mode = 1;
      }
      S.min_x[lane] = t.min_x;
      S.min_y[lane] = t.min_y;
      S.w[lane] = t.w;
      S.iw[lane] = t.iw;
      S.pa[lane] = t.A;
      S.pb[lane] = t.B;
      S.pc[lane] = t.C;
#pragma unroll
      for (int k = 0; k < 7; k++) S.ek[k][lane] = t.ek[k];
      S.key[lane] = t.key & ~kKeyBoxForm;
      S.rec[lane] = c;
    }
    if (ns == 0) {
      ns = 1;   // one sample that draws nothing keeps every staged record's start distinct
      mode = 2;
    }
    S.mode[lane] = mode;
  }
  const uint32_t inc = wave_incl_scan(ns);
  S.begin[lane] = inc - ns;
  if (lane == kWave - 1) S.begin[kWave] = inc;
  S.start[lane] = 0xFFFFFFFFu;
  wave_sync_lds();
  return S.begin[kWave];
}

// Every staged sample with all lanes: kId false adds depths to the LDS
// depth tile, true tests them against it and adds id keys.
template <bool kId>
__device__ __forceinline__ void sweep_tile(TileStage& S, uint32_t nsmp, uint32_t nstaged, uint32_t* sd,
                                           unsigned long long* sid, float tx0, float ty0, float tw, float th,
                                           uint32_t twl) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t my_begin = (uint32_t)lane < nstaged ? S.begin[lane] : nsmp;
  for (uint32_t base = 0; base < nsmp; base += kWave) {
    // lane p takes sample j = base + p, of the last staged record beginning
    // at or before j (records hold at least one sample each)
    const unsigned long long before = __ballot(my_begin <= base);
    const uint32_t off = my_begin - base;
    if (my_begin > base && off < (uint32_t)kWave) S.start[off] = base;
    wave_sync_lds();
    const bool starts_here = S.start[lane] == base;
    const unsigned long long starts = __ballot(starts_here);
    const uint32_t j = base + (uint32_t)lane;
    if (j >= nsmp) continue;
    const int k = __popcll(before) - 1 +
                  (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(starts >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)starts, 0u)) +
                  (starts_here ? 1 : 0);
    const uint32_t sl = j - S.begin[k];
    const uint32_t mode = S.mode[k];
    float x, y;
    bool in;
    if (mode == 0) {
      // row-major position in the clipped box (sl < 2^12, exact)
      const float fs = (float)sl, rw = S.rw[k];
      const float ly = floorf((fs + 0.5f) * S.inv[k]);
      x = S.x0[k] + (fs - ly * rw);
      y = S.y0[k] + ly;
      in = inside_tri(x, y, S.ek[0][k], S.ek[1][k], S.ek[2][k], S.ek[3][k], S.ek[4][k], S.ek[5][k], S.ek[6][k]);
    } else if (mode == 1) {
      in = sample_xy<kOpSize>((float)(S.ja[k] + sl), S.min_x[k], S.min_y[k], S.w[k], S.iw[k], S.ek[0][k], S.ek[1][k],
                              S.ek[2][k], S.ek[3][k], S.ek[4][k], S.ek[5][k], S.ek[6][k], x, y) &&
           x >= tx0 && x < tx0 + tw && y >= ty0 && y < ty0 + th;
    } else {
      continue;
    }
    if (!in) continue;
    const uint32_t d = depth_quant(fmaf(S.pa[k], x, fmaf(S.pb[k], y, S.pc[k])));
    const uint32_t p = ((uint32_t)(y - ty0) << twl) | (uint32_t)(x - tx0);
    if constexpr (kId) {
      const uint32_t val = sd[p];
      if (d > val - kBufError && d < val + kBufError)
        atomicMax(&sid[p], ((unsigned long long)(S.key[k] + 1u) << 32) | S.rec[k]);
    } else {
      atomicMin(&sd[p], d);
    }
  }
}


// Depth and ids of one tile by the workgroup's kWaves wavefronts: its
// records in chunks of 64, the depth sweep of every chunk, a barrier, the id
// sweep of every chunk (restaged unless each wavefront had one chunk), a
// barrier, then the tile's pixels inside the view written once.
template <int kLogPix, int kWaves>
__device__ __forceinline__ void raster_tile(const PassArgs& a, uint32_t tile, bool ok, uint32_t* sd,
                                            unsigned long long* sid, TileStage* stage) {
  constexpr int kPix = 1 << kLogPix, kThreads = kWaves * kWave;
  const Lists& Ls = a.lists;
  const TileBins& tb = Ls.tb;
  // the view owning this tile: the last whose first tile is at or before it
  uint32_t vi = 0;
  if (a.views) {
    uint32_t lo = 0, hi = a.nviews;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (a.views[mid].bin_begin <= tile) lo = mid; else hi = mid;
    }
    vi = lo;
  }
  const DevView& V = a.views ? a.views[vi] : a.view0;
  const uint32_t twl = tb.tw_log2, thl = tb.th_log2;
  const uint32_t lt = tile - V.bin_begin;
  const uint32_t txi = lt % V.bin_tiles_x, tyi = lt / V.bin_tiles_x;
  const uint32_t px0 = txi << twl, py0 = tyi << thl;
  const float tx0 = (float)px0, ty0 = (float)py0, tw = (float)(1u << twl), th = (float)(1u << thl);
  for (int i = threadIdx.x; i < kPix; i += kThreads) {
    sd[i] = 0xFFFFFFFFu;
    sid[i] = 0ull;
  }
  __syncthreads();
  if (ok) {
    const unsigned long long packed = *reinterpret_cast<const unsigned long long*>(Ls.counters + kPacked);
    const uint32_t ncut = min((uint32_t)(packed >> 32), Ls.cap_cut), total = (uint32_t)packed;
    const uint32_t e0 = tb.off[tile], e1 = tb.off[tile + 1];
    const uint32_t nchunks = (e1 - e0 + kWave - 1) / kWave;
    const int wave = threadIdx.x / kWave;
    TileStage& S = stage[wave];
    const bool keep = nchunks <= (uint32_t)kWaves;
    for (uint32_t ch = wave; ch < nchunks; ch += kWaves) {
      const uint32_t n = stage_chunk(S, Ls, e0 + ch * kWave, e1, tx0, ty0, tw, th, ncut, total);
      sweep_tile<false>(S, n, min(e1 - (e0 + ch * kWave), (uint32_t)kWave), sd, sid, tx0, ty0, tw, th, twl);
    }
    __syncthreads();
    for (uint32_t ch = wave; ch < nchunks; ch += kWaves) {
      const uint32_t n = keep ? S.begin[kWave] : stage_chunk(S, Ls, e0 + ch * kWave, e1, tx0, ty0, tw, th, ncut, total);
      if (keep) {
        S.start[threadIdx.x & (kWave - 1)] = 0xFFFFFFFFu;
        wave_sync_lds();
      }
      sweep_tile<true>(S, n, min(e1 - (e0 + ch * kWave), (uint32_t)kWave), sd, sid, tx0, ty0, tw, th, twl);
    }
    __syncthreads();
  }
  const uint32_t W = V.v.width, H = V.v.height;
  for (int i = threadIdx.x; i < kPix; i += kThreads) {
    const uint32_t x = px0 + ((uint32_t)i & ((1u << twl) - 1u)), y = py0 + ((uint32_t)i >> twl);
    if (x < W && y < H) {
      const size_t o = (size_t)y * W + x;
      V.depth[o] = sd[i];
      V.idkey[o] = sid[i];
    }
  }
}

// One workgroup per tile.
template <int kLogPix>
__global__ void __launch_bounds__(kTileWaves * kWave) k_tile_raster(PassArgs a) {
  __shared__ uint32_t sd[1 << kLogPix];
  __shared__ unsigned long long sid[1 << kLogPix];
  __shared__ TileStage stage[kTileWaves];
  raster_tile<kLogPix, kTileWaves>(a, blockIdx.x, a.lists.counters[kOverflow] == 0u, sd, sid, stage);
}

// ---------------------------------------------------------------- launches
hipError_t launch_bin_reset(const TileBins& tb, uint32_t* counters, hipStream_t s) {
  const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>((tb.ntiles + 255) / 256, 1024u));
  hipLaunchKernelGGL(k_bin_reset, dim3(blocks), dim3(256), 0, s, counters, tb.cnt, tb.ntiles);
  return hipGetLastError();
}

uint32_t bin_scan_chunks(uint32_t ntiles) { return std::max<uint32_t>(1u, (ntiles + kScanChunk - 1) / kScanChunk); }

hipError_t launch_bins(const PassArgs& a, hipStream_t s) {
  const TileBins& tb = a.lists.tb;
  const uint32_t chunks = bin_scan_chunks(tb.ntiles);
  hipLaunchKernelGGL(k_bin_sums, dim3(chunks), dim3(kScanThreads), 0, s, tb, (const uint32_t*)a.lists.counters);
  hipLaunchKernelGGL(k_bin_offsets, dim3(chunks), dim3(kScanThreads), 0, s, tb, a.lists.counters);
  hipLaunchKernelGGL(k_bin_scatter, dim3(device_cu_count() * 4), dim3(256), 0, s, tb,
                     (const uint32_t*)a.lists.counters, a.lists.cap_cut);
  return hipGetLastError();
}

hipError_t launch_tile_raster(const PassArgs& a, hipStream_t s) {
  const TileBins& tb = a.lists.tb;
  const dim3 grid(tb.ntiles), block(kTileWaves * kWave);
  switch (tb.tw_log2 + tb.th_log2) {
    case 8: hipLaunchKernelGGL(k_tile_raster<8>, grid, block, 0, s, a); break;
    case 9: hipLaunchKernelGGL(k_tile_raster<9>, grid, block, 0, s, a); break;
    case 10: hipLaunchKernelGGL(k_tile_raster<10>, grid, block, 0, s, a); break;
    case 11: hipLaunchKernelGGL(k_tile_raster<11>, grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ocr
