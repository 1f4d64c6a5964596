// walk.h — the reference's fragment walk (kernel1, cl2.cl:5000-5127) as
// closed forms shared by the sample-space raster passes (raster.hip) and
// the tiled raster path (tiles.hip), and small wavefront helpers.
#pragma once
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

// The walk's row of sample q: floor(mad(q, 1/width, min_y)) (cl2.cl:5083).
// Sample indices are integers below 2^24, exact in binary32.
__device__ __forceinline__ float row_of(float q, float iw, float min_y) {
  return floorf(fmaf(q, iw, min_y));
}

// First sample q >= 0 whose row is >= Y, from the nominal n0 = (Y - min_y) * w
// (exact: an integer below 2^24).  The fma rounds across a row boundary for at
// most a few samples, so the steps below run only on such rows.
__device__ __forceinline__ float run_start_steps(float Y, float q, float iw, float min_y) {
#pragma clang loop unroll(disable)
  for (int g = 0; g < 64 && q > 0.f && row_of(q - 1.f, iw, min_y) >= Y; g++) q -= 1.f;
#pragma clang loop unroll(disable)
  for (int g = 0; g < 64 && row_of(q, iw, min_y) < Y; g++) q += 1.f;
  return q;
}

__device__ __forceinline__ float run_start(float Y, float w, float iw, float min_y) {
  const float n0 = (Y - min_y) * w;
  const bool lead = n0 > 0.f && row_of(n0 - 1.f, iw, min_y) >= Y;
  const bool lag = row_of(n0, iw, min_y) < Y;
  return (lead || lag) ? run_start_steps(Y, n0, iw, min_y) : n0;
}

// point_in_tri (cl2.cl:4798-4807) with the triangle's loop-invariant terms.
__device__ __forceinline__ bool inside_tri(float x, float y, float e0, float e1, float e2, float e3, float e4,
                                           float e5, float e6) {
  const float s_ = e0 + e1 * x + e2 * y;
  const float t_ = e3 + e4 * x + e5 * y;
  return s_ > -0.0001f && t_ > -0.0001f && (s_ + t_) < e6;
}

// Sample `j` (index within its triangle, an exact binary32 integer) of a
// set-up triangle: its column x, row y and whether it lies inside.  Sample j
// is sample o = j % (op + 1) of fragment a = j / (op + 1), at walk index
// q = op*a + o; the quotient is a float estimate corrected by one step each
// way.
template <int kOp>
__device__ __forceinline__ bool sample_xy(float slf, float min_x, float min_y, float w, float iw, float e0, float e1,
                                          float e2, float e3, float e4, float e5, float e6, float& x, float& y) {
  float faf = floorf(slf * (1.0f / (float)(kOp + 1)));
  float of = fmaf(-faf, (float)(kOp + 1), slf);
  faf = of < 0.f ? faf - 1.f : (of >= (float)(kOp + 1) ? faf + 1.f : faf);
  of = fmaf(-faf, (float)(kOp + 1), slf);
  const float paf = faf * (float)kOp;
  const float qf = paf + of;
  y = row_of(qf, iw, min_y);
  // cursor column: (r % w) + min_x + (q - r), r = max(pa, row start)
  const float rs = run_start(y, w, iw, min_y);
  const float r = paf > rs ? paf : rs;
  const float m = floorf(r * iw);
  float e = fmaf(-m, w, r);
  e = e < 0.f ? e + w : e;
  e = e >= w ? e - w : e;
  const float xo = e + (qf - r);
  x = min_x + xo;
  if (xo >= w) return false;   // x >= max_x
  return inside_tri(x, y, e0, e1, e2, e3, e4, e5, e6);
}

// Whether the walk of a set-up triangle visits exactly its bounding box
// [min_x, max_x) x [min_y, max_y) (rows >= 1): true when every row boundary
// k*w (k = 1..rows) starts its row on time, row_of(k*w - 1) = min_y + k - 1
// and row_of(k*w) = min_y + k.  Then each row's run is [(y - min_y)*w,
// (y - min_y + 1)*w) with the cursor at min_x at its start, a fragment that
// starts inside a run puts the cursor where the run has it (its starting
// row comes from a correctly rounded quotient, exact for these sizes), and
// the walk ends at the first sample of row max_y.  Two closed-form bounds
// settle most triangles without the loop, with iw = RN(1/w) = (1 + d)/w,
// |d| <= 2^-24, and m = min_y + k:
//   a row never starts late when d >= 0 (the fma's exact value is >= m), or
//   when rows < min_y (a late start needs k|d| > half the spacing below m,
//   which is at least m 2^-25, i.e. k > min_y);
//   a row never starts early when w (max_y + rows) <= 2^24 - 1 (an early
//   start needs iw - k d <= half the spacing below m < m 2^-24, i.e.
//   w (m + k) > 2^24 - 1, and m + k <= max_y + rows).
// Others are checked boundary by boundary up to 64 rows; taller ones count
// as not certified.  Checked against the reference walk on the CPU
// (tests/tile_walk_check.c).  walk_is_rect_closed: the two bounds alone
// (a sufficient condition, no loop), which the raster passes' box-form
// records use.
__device__ __forceinline__ bool walk_is_rect_closed(float min_y, float max_y, float w, float iw) {
  const float rows = max_y - min_y;
  const bool no_lag = fmaf(w, iw, -1.f) >= 0.f || rows < min_y;
  const bool no_lead = w * (max_y + rows) <= 16777215.f;
  return no_lag && no_lead;
}

__device__ __forceinline__ bool walk_is_rect(float min_y, float max_y, float w, float iw) {
  const float rows = max_y - min_y;
  if (walk_is_rect_closed(min_y, max_y, w, iw)) return true;
  if (rows > 64.f) return false;
  bool ok = true;
#pragma clang loop unroll(disable)
  for (float k = 1.f; k <= rows; k += 1.f) {
    const float n0 = k * w;
    ok = ok && row_of(n0 - 1.f, iw, min_y) == min_y + (k - 1.f) && row_of(n0, iw, min_y) == min_y + k;
  }
  return ok;
}

// Number of a triangle's samples j (fragment a = j / (op + 1), walk index
// q(j) = op*a + j % (op + 1)) whose walk index is below q: q(j) does not
// decrease with j, and the walk index op*a is visited twice (as the last
// sample of fragment a - 1 and the first of fragment a).
template <int kOp>
__device__ __forceinline__ uint32_t first_sample(uint32_t q) {
  const uint32_t a = q / (uint32_t)kOp, r = q - a * (uint32_t)kOp;
  return q + a - (r == 0u && a > 0u ? 1u : 0u);
}

}  // namespace ocr
