// tiles.h — screen-tile bins of the tiled raster path: a cut record's tile
// rectangle, and the wavefront loop that visits every (record, tile) pair of
// its lanes, 64 pairs per step.
//
// The reference's tiled path (split_into_tiled_chunks, cl2.cl:3850-4026,
// driven by engine::draw_tiled_deferred, engine.cpp:2447-2535) appends each
// triangle to the display lists of the tiles around the tile corners it
// covers, in chunks of tile_chunk_size slots taken from a global counter.
// Here a record is listed in every tile its bounding box touches (the set
// of tiles whose pixels its walk can visit), the lists are one array
// ordered by tile (counted, scanned, scattered), and nothing is dropped.
#pragma once
#pragma clang fp contract(off)

#include "ocr_internal.h"
#include "walk.h"

namespace ocr {

// The TileBox of a set-up cut record of `view` (box [min_x, min_x + w) x
// [min_y, max_y), w >= 1, max_y > min_y).
__device__ __forceinline__ uint4 tile_box(const RasterTri& t, float max_y, const TileBins& tb, uint32_t bin_begin,
                                          uint32_t tiles_x) {
  const uint32_t x0 = (uint32_t)t.min_x, x1 = (uint32_t)(t.min_x + t.w) - 1u;
  const uint32_t y0 = (uint32_t)t.min_y, y1 = (uint32_t)max_y - 1u;
  uint4 b;
  b.x = (x0 >> tb.tw_log2) | (x1 >> tb.tw_log2) << 16;
  b.y = (y0 >> tb.th_log2) | (y1 >> tb.th_log2) << 16;
  b.z = bin_begin;
  b.w = (uint32_t)max_y | tiles_x << 16 | (walk_is_rect(t.min_y, max_y, t.w, t.iw) ? kBoxRect : 0u);
  return b;
}

__device__ __forceinline__ uint32_t box_tile(const uint4& b, uint32_t tx, uint32_t ty) {
  return b.z + ty * ((b.w >> 16) & 0x7FFFu) + tx;
}

// Visits the (record, tile) pairs of the wavefront's lanes, each lane's up
// to two boxes (has0 / has1; both of one view), as one list spread over the
// wavefront: step k gives lane p pair 64k + p of the concatenation of the
// lanes' pairs (each box in row-major tile order).  Neighbouring pairs of
// one step with the same tile form a run (small triangles of one object
// share tiles): the run's first lane calls head(tile, run length), one
// atomic for the run, and every lane of the run calls
// member(tile, owner lane, head's result, rank in the run).  Called by all
// lanes of the wavefront.
template <typename Head, typename Member>
__device__ __forceinline__ void expand_tile_pairs(const uint4& b0, bool has0, const uint4& b1, bool has1, Head head,
                                                  Member member) {
  const auto ntiles_of = [](const uint4& b) {
    return ((b.x >> 16) - (b.x & 0xFFFFu) + 1u) * ((b.y >> 16) - (b.y & 0xFFFFu) + 1u);
  };
  const uint32_t n0 = has0 ? ntiles_of(b0) : 0u, n1 = has1 ? ntiles_of(b1) : 0u;
  const uint32_t n = n0 + n1;
  const uint32_t incl = wave_incl_scan(n), excl = incl - n;
  const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
  const int lane = threadIdx.x & (kWave - 1);
  // both boxes of a lane are of one view: one tile base and row length
  const uint4 bz = has0 ? b0 : b1;
  for (uint32_t base = 0; base < total; base += kWave) {
    const uint32_t i = base + (uint32_t)lane;
    // owner: the last lane whose first pair is at or before i (it holds i:
    // a later lane with no pairs starts where the owner's pairs end)
    int lo = 0;
#pragma unroll
    for (int step = kWave / 2; step >= 1; step >>= 1) {
      const int mid = lo + step;
      if (__shfl(excl, mid, kWave) <= i) lo = mid;
    }
    const uint32_t o_excl = __shfl(excl, lo, kWave), o_n0 = __shfl(n0, lo, kWave);
    const uint32_t ax = __shfl(b0.x, lo, kWave), ay = __shfl(b0.y, lo, kWave);
    const uint32_t cx = __shfl(b1.x, lo, kWave), cy = __shfl(b1.y, lo, kWave);
    const uint32_t z = __shfl(bz.z, lo, kWave), tiles_x = (__shfl(bz.w, lo, kWave) >> 16) & 0x7FFFu;
    const bool active = i < total;
    uint32_t t = 0xFFFFFFFFu;
    if (active) {
      uint32_t l = i - o_excl;
      const bool first = l < o_n0;
      if (!first) l -= o_n0;
      const uint32_t bx = first ? ax : cx, by = first ? ay : cy;
      const uint32_t rw = (bx >> 16) - (bx & 0xFFFFu) + 1u;
      // l / rw for l < 2^16 (a frame has at most 62,500 tiles of 256 pixels
      // or more): the quotient of l + 1/2 is at least 1/(2 rw) from an
      // integer, far above the error of the reciprocal and the product
      const uint32_t q = (uint32_t)floorf(((float)l + 0.5f) * __builtin_amdgcn_rcpf((float)rw));
      t = z + ((by & 0xFFFFu) + q) * tiles_x + (bx & 0xFFFFu) + (l - q * rw);
    }
    // runs of equal tiles among the step's active lanes
    const uint32_t prev = __shfl_up(t, 1, kWave);
    const bool is_head = active && (lane == 0 || prev != t);
    const unsigned long long heads = __ballot(is_head);
    const unsigned long long act = __ballot(active);
    const unsigned long long upto = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1ull);
    const int hpos = 63 - __clzll((long long)(heads & upto));
    const unsigned long long after = heads & ~upto;
    const int next = after ? __ffsll((long long)after) - 1 : 64 - __clzll((long long)act);
    uint32_t r = 0;
    if (is_head) r = head(t, (uint32_t)(next - lane));
    r = (uint32_t)__shfl((int)r, hpos, kWave);
    if (active) member(t, lo, r, (uint32_t)(lane - hpos));
  }
}

}  // namespace ocr
