// ocr_math.h — device arithmetic of the raster + shading path (gfx950).
//
// The reference kernels are built with -cl-fast-relaxed-math and FP_CONTRACT
// ON (ocl.h:227, cl2.cl:3), which leaves rounding implementation defined.
// This library fixes it so that the depth and triangle-id buffers are
// reproducible bit for bit on any IEEE host:
//   * binary32 round-to-nearest-even, no contraction (every file is built
//     with -ffp-contract=off and says so with the pragma below); the
//     reference's mad(a,b,c) is an explicit fmaf;
//   * division and sqrt correctly rounded (-fhip-fp32-correctly-rounded-
//     divide-sqrt);
//   * fast_normalize(v) = v * (1/sqrt(dot(v,v))), dot summed left to right;
//   * pow(x,2) = x*x; pow(x,4) and powr(x,5) are built in binary64 from the
//     exact square and rounded once; tan is evaluated in binary64 and
//     rounded once; exp by a fixed binary32 sequence (exp_f);
//   * float -> u32 conversions saturate, NaN -> 0.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ocr {

constexpr float kDepthFar = 350000.0f;       // cl2.cl:17
constexpr float kMulint = 4294967296.0f;     // (float)UINT_MAX, cl2.cl:19
constexpr int kIcutoff = 20;                 // depth_icutoff, cl2.cl:21-23
constexpr int kOpSize = 500;                 // cl2.cl:4247
constexpr int kOpSizeLight = 300;            // cl2.cl:4249
constexpr uint32_t kBufError = 20u;          // cl2.cl:5387
constexpr float kPi = 3.1415927f;            // cl2.cl:14
constexpr float kAmbient = 0.2f;             // main.cpp:81
constexpr float kSsaoRad = 2.f;              // main.cpp:82
constexpr float kSsaoDiv = 2.5f;             // cl2.cl:2253-2255
constexpr int kShadowBias = 50;              // ocl.h:234
constexpr int kMipLevels = 4;                // cl2.cl:5
constexpr float kMipBias = 1.1f;             // cl2.cl:5750-5752

struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// a / b, correctly rounded, from y = RN(1/b) (Markstein's correction: the
// residual a - b*q is exact under fma, and one fma step from a faithful
// quotient lands on the rounded one).  Valid when a is +0 or a, b and a/b
// are normal binary32 numbers; for the constant divisors below it is checked
// against IEEE division for a = 0 and every finite a >= 2^-100 (the values
// divided are integers, depths of at least 20, or sums of weights in
// [0, 9]).  Three instructions in place of the eleven of a full correctly
// rounded division.
__device__ __forceinline__ float div_rcp(float a, float b, float y) {
  const float q = a * y;
  const float r = fmaf(-q, b, a);
  return fmaf(r, y, q);
}
// RN(1/b) from the hardware reciprocal estimate and one Newton step (3
// instructions).  Equal to the IEEE quotient 1/b for every binary32 b with
// |b| in [2^-126, 2^126]: checked on the device for each such b
// (ocr_calib_rcp_check in calib.hip).  Callers keep b inside that range.
__device__ __forceinline__ float rcp_rn(float b) {
  const float r0 = __builtin_amdgcn_rcpf(b);
  const float e = fmaf(-b, r0, 1.0f);
  return fmaf(e, r0, r0);
}
// RN(1/b) for b that is +-0, +-inf, NaN or has |b| in [2^-126, 2^126] (no
// subnormals): rcp_rn, with the hardware estimate's exact +-inf / +-0 at
// the ends.  Integer-valued b and sqrt results qualify.
__device__ __forceinline__ float rcp_cr(float b) {
  const float r0 = __builtin_amdgcn_rcpf(b);
  const float r = fmaf(fmaf(-b, r0, 1.0f), r0, r0);
  return (b == 0.f || __builtin_isinf(b)) ? r0 : r;
}
constexpr float kInv9 = 1.0f / 9.0f;       // RN(1/9), folded by the compiler
constexpr float kInv125 = 1.0f / 125.0f;
constexpr float kInv100 = 1.0f / 100.0f;
constexpr float kInvSsaoDiv = 1.0f / kSsaoDiv;
constexpr float kInv255 = 1.0f / 255.0f;
// b / 255 correctly rounded for every byte b in two operations: b*hi + b*lo
// with hi + lo the two-term split of 1/255 (hi = RN(1/255), lo = RN(1/255 -
// hi)), the sum formed by one fma; exact for all 256 values (checked with
// rational arithmetic).
constexpr float kInv255Hi = 0x1.010102p-8f;
constexpr float kInv255Lo = -0x1.fdfdfep-33f;
__device__ __forceinline__ float unorm8(uint32_t b) {
  const float f = (float)b;
  return fmaf(f, kInv255Hi, f * kInv255Lo);
}
constexpr float kInvDepthFar = 1.0f / kDepthFar;
// ((float)u / kMulint) * kDepthFar (cl2.cl:5834, :2199, :2628) in one
// product.  (float)u is 0 or at least 1, so the division by 2^32 is exact
// and commutes with the rounding of the product: RN(RN(u) * 2^-32 * 350000)
// = RN(RN(u) * C) with C = 350000 * 2^-32, itself exact in binary32.
constexpr float kDepthScale = kDepthFar / kMulint;
__device__ __forceinline__ float depth_of(uint32_t u) { return (float)u * kDepthScale; }
constexpr float kInv65535 = 1.0f / 65535.0f;
// A uniform divisor d takes the fast path when it and 1/d are comfortably
// normal, which keeps every quotient the path forms in the normal range.
__device__ __forceinline__ bool rcp_safe(float d) { return fabsf(d) >= 0x1p-60f && fabsf(d) <= 0x1p60f; }

// The correctly rounded square root.  For x in [2^-100, FLT_MAX] the
// hardware estimate corrected once, s + (x - s*s) * rsq(x) / 2 (five
// instructions), is the correctly rounded root: checked on the device for
// every binary32 in that range (ocr_calib_rcp_check, calib.hip), which also
// finds it wrong below (residuals there underflow).  Other inputs (zero,
// tiny, infinite, negative, NaN) take the full sequence in a branch that a
// wavefront skips when none of its lanes needs it.
__device__ __forceinline__ bool sqrt_fast_ok(float x) {
  // x in [2^-100, FLT_MAX]: the bit pattern of a non-negative finite x is
  // its order, negative x and NaN are above every finite positive pattern
  return (uint32_t)__float_as_uint(x) - 0x0D800000u <= 0x7F7FFFFFu - 0x0D800000u;
}
__device__ __forceinline__ float sqrt_core(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float h = 0.5f * __builtin_amdgcn_rsqf(x);
  return fmaf(fmaf(-s, s, x), h, s);
}
__device__ __forceinline__ float sqrt_cr(float x) {
  float s = sqrt_core(x);
  if (!sqrt_fast_ok(x)) s = sqrtf(x);
  return s;
}
// 1 / sqrt(d), both correctly rounded (the reference's fast_normalize and
// fast_length, under the arithmetic contract).  On the fast path the root
// lies in [2^-50, 2^64], where rcp_rn is the correctly rounded reciprocal;
// other inputs divide.  Equal to 1.0f / sqrtf(d) for every binary32 d >= 0
// and NaN: checked on the device for each (calib.hip).
__device__ __forceinline__ float inv_sqrt_cr(float d) {
  float r = rcp_rn(sqrt_core(d));
  if (!sqrt_fast_ok(d)) r = 1.0f / sqrtf(d);
  return r;
}
// sqrt_cr(d) and inv_sqrt_cr(d) together, from one root.
__device__ __forceinline__ void length_and_inverse(float d, float& len, float& inv) {
  len = sqrt_core(d);
  inv = rcp_rn(len);
  if (!sqrt_fast_ok(d)) {
    len = sqrtf(d);
    inv = 1.0f / len;
  }
}
// A dividend whose quotients by divisors in [1, 2^60] the short division
// forms correctly rounded: +0, or a normal magnitude in [2^-60, 2^60].
__device__ __forceinline__ bool bright_ok(float b) {
  return __float_as_uint(b) == 0u || (fabsf(b) >= 0x1p-60f && fabsf(b) <= 0x1p60f);
}
__device__ __forceinline__ f3 fast_normalize(f3 v) { return v * inv_sqrt_cr(dot(v, v)); }
__device__ __forceinline__ float fast_length(f3 v) { return sqrt_cr(dot(v, v)); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float min3f(float a, float b, float c) { return fminf(fminf(a, b), c); }
__device__ __forceinline__ float max3f(float a, float b, float c) { return fmaxf(fmaxf(a, b), c); }

// convert_uint_sat: NaN and f <= 0 give 0, f >= 2^32 gives UINT_MAX, others
// truncate (4294967040 is the largest binary32 below 2^32).
__device__ __forceinline__ uint32_t sat_u32(float f) {
  const uint32_t v = (uint32_t)fminf(fmaxf(f, 0.0f), 4294967040.0f);
  return f >= 4294967296.0f ? 0xFFFFFFFFu : v;
}

// convert_uint_sat(UINT_MAX / fd), the quantised depth of kernel1/2
// (cl2.cl:5115): 2^32 * RN(1/fd) is RN(2^32/fd) wherever the saturated value
// depends on it (quotients in [1, 2^32)), and outside rcp_rn's range both
// saturate alike except for +0 and positive subnormal fd, where the
// quotient is +inf.
__device__ __forceinline__ uint32_t depth_quant(float fd) {
  const uint32_t d = sat_u32(kMulint * rcp_rn(fd));
  return (fd >= 0.f && fd < 0x1p-126f && !signbit(fd)) ? 0xFFFFFFFFu : d;
}

// tan in binary64: reduction by pi/2 + Taylor sin (deg 19) / cos (deg 18).
__device__ __forceinline__ double tan_d(double x) {
  const double kTwoOverPi = 6.36619772367581382433e-01;
  const double kPio2Hi = 1.57079632673412561417e+00;
  const double kPio2Lo = 6.07710050650619224932e-11;
  double k = floor(fma(x, kTwoOverPi, 0.5));
  double r = fma(-k, kPio2Lo, fma(-k, kPio2Hi, x));
  double r2 = r * r;
  double s = -1.0 / 121645100408832000.0;
  s = fma(s, r2, 1.0 / 355687428096000.0);
  s = fma(s, r2, -1.0 / 1307674368000.0);
  s = fma(s, r2, 1.0 / 6227020800.0);
  s = fma(s, r2, -1.0 / 39916800.0);
  s = fma(s, r2, 1.0 / 362880.0);
  s = fma(s, r2, -1.0 / 5040.0);
  s = fma(s, r2, 1.0 / 120.0);
  s = fma(s, r2, -1.0 / 6.0);
  s = fma(s * r2, r, r);
  double c = 1.0 / 6402373705728000.0;
  c = fma(c, r2, -1.0 / 20922789888000.0);
  c = fma(c, r2, 1.0 / 87178291200.0);
  c = fma(c, r2, -1.0 / 479001600.0);
  c = fma(c, r2, 1.0 / 3628800.0);
  c = fma(c, r2, -1.0 / 40320.0);
  c = fma(c, r2, 1.0 / 720.0);
  c = fma(c, r2, -1.0 / 24.0);
  c = fma(c, r2, 0.5);
  c = fma(-c, r2, 1.0);
  long long ki = (long long)k;
  return (ki & 1) ? -c / s : s / c;
}

// exp in binary32: k = rint(x log2 e), r = x - k ln2 (two-constant
// Cody-Waite: k * ln2_hi is exact for |k| < 2^11), e^r = 1 + r p(r) with a
// degree-6 polynomial fitted to e^r on [-ln2/2, ln2/2] (relative error
// 3.4e-9), scaled by ldexp.  Within one ulp of the correctly rounded exp
// over [-104, 88.75] (oracle pins); the same operations as the oracle's.
__device__ __forceinline__ float exp_f(float x) {
  if (!(x >= -104.0f)) return x != x ? x : 0.0f;
  if (x > 88.75f) return __builtin_inff();
  const float k = rintf(x * 1.44269502f);
  float r = fmaf(-k, 0.693145752f, x);
  r = fmaf(-k, 1.42860677e-06f, r);
  float p = 0.001399628003127873f;
  p = fmaf(p, r, 0.008374124765396118f);
  p = fmaf(p, r, 0.04166550189256668f);
  p = fmaf(p, r, 0.16666433215141296f);
  p = fmaf(p, r, 0.5000000596046448f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}
__device__ __forceinline__ float tan_f(float x) { return (float)tan_d((double)x); }
__device__ __forceinline__ float powr5(float x) { double d = x; double d2 = d * d; return (float)((d2 * d2) * d); }
__device__ __forceinline__ float pow4(float x) { double d = x; double d2 = d * d; return (float)(d2 * d2); }

// sin/cos of a rotation (camera or cube face), computed on the host.
struct SinCos { float s[3]; float c[3]; };

// rot, cl2.cl:220-236 (extrinsic XYZ about c_pos).
__device__ __forceinline__ f3 rot(f3 point, f3 c_pos, const SinCos& sc) {
  const float cx = sc.c[0], cy = sc.c[1], cz = sc.c[2];
  const float sx = sc.s[0], sy = sc.s[1], sz = sc.s[2];
  f3 rel = point - c_pos;
  float inner = fmaf(sz, rel.y, cz * rel.x);
  f3 r;
  r.x = fmaf(cy, inner, -sy * rel.z);
  float t = fmaf(cy, rel.z, sy * inner);
  float u = fmaf(cz, rel.y, -sz * rel.x);
  r.y = fmaf(sx, t, cx * u);
  r.z = fmaf(cx, t, -sx * u);
  return r;
}

// back_rot with c_pos = 0, cl2.cl:272-349.
__device__ __forceinline__ f3 back_rot(f3 rel, const SinCos& sc) {
  const float cx = sc.c[0], cy = sc.c[1], cz = sc.c[2];
  const float sx = sc.s[0], sy = sc.s[1], sz = sc.s[2];
  f3 r;
  r.x = cz * fmaf(cy, rel.x, fmaf(sx, sy * rel.y, cx * sy * rel.z)) + sz * (sx * rel.z - cx * rel.y);
  r.y = fmaf(sz, cy * rel.x, fmaf(fmaf(cx, cz, sx * sy * sz), rel.y, fmaf(-sx, cz, cx * sy * sz) * rel.z));
  r.z = fmaf(-sy, rel.x, cy * fmaf(sx, rel.y, cx * rel.z));
  return r;
}

// rot_quat, cl2.cl:350-357.
__device__ __forceinline__ f3 rot_quat(f3 p, f4 q) {
  float d = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  float inv = inv_sqrt_cr(d);
  f3 qv = mk3(q.x * inv, q.y * inv, q.z * inv);
  float qw = q.w * inv;
  f3 t = cross(qv, p) * 2.f;
  return (p + t * qw) + cross(qv, t);
}

// rot_quat's rotation by a quaternion it has already normalised (qv, qw as
// rot_quat forms them): the same operations, without forming the
// normalisation again.
__device__ __forceinline__ f3 rot_quat_n(f3 p, f4 qn) {
  const f3 qv = mk3(qn.x, qn.y, qn.z);
  f3 t = cross(qv, p) * 2.f;
  return (p + t * qn.w) + cross(qv, t);
}

// back_rot_quat, cl2.cl:359-370.
__device__ __forceinline__ f3 back_rot_quat(f3 p, f4 q) {
  f4 c{-q.x, -q.y, -q.z, q.w};
  float len_sq = c.x * c.x + c.y * c.y + c.z * c.z + c.w * c.w;
  return rot_quat(p, f4{c.x / len_sq, c.y / len_sq, c.z / len_sq, c.w / len_sq});
}

// depth_project, cl2.cl:535-545.
__device__ __forceinline__ f3 depth_project(f3 r, float width, float height, float fovc) {
  float f = fovc / r.z;
  return mk3(fmaf(r.x, f, width / 2.f), fmaf(r.y, f, height / 2.f), r.z);
}
// The same for vertices that passed clip_near, whose z lies in [20, 350000]
// (or is the rounding of a point on z = 20), with fovc normal (views are
// checked on the host, cube faces use LIGHTBUFFERDIM / 2).
__device__ __forceinline__ f3 depth_project_clipped(f3 r, float width, float height, float fovc) {
  float f = div_rcp(fovc, r.z, rcp_rn(r.z));
  return mk3(fmaf(r.x, f, width / 2.f), fmaf(r.y, f, height / 2.f), r.z);
}

__device__ __forceinline__ f3 sel3(int k, f3 a, f3 b, f3 c) {
  return k == 0 ? a : (k == 1 ? b : c);
}

// The point of edge a -> b at z = depth_icutoff (cl2.cl:636-637).
__device__ __forceinline__ f3 clip_edge(f3 a, f3 b) {
  return a + ((b - a) * ((float)kIcutoff - a.z)) / (b.z - a.z);
}

// generate_new_triangles, cl2.cl:577-663: near/far clip into 0..2
// triangles.  Vertex choices are selects, not array indexing, so the three
// vertices stay in registers.
__device__ __forceinline__ int clip_near(f3 p0, f3 p1, f3 p2, f3 out[2][3]) {
  const bool b0 = p0.z <= (float)kIcutoff || p0.z > kDepthFar;
  const bool b1 = p1.z <= (float)kIcutoff || p1.z > kDepthFar;
  const bool b2 = p2.z <= (float)kIcutoff || p2.z > kDepthFar;
  const int nb = (int)b0 + (int)b1 + (int)b2;
  if (nb > 2) return 0;
  if (nb == 0) { out[0][0] = p0; out[0][1] = p1; out[0][2] = p2; return 1; }
  if (nb == 1) {
    // g1 = the vertex behind, g2 / g3 the next two in cyclic order
    const int id = b0 ? 0 : (b1 ? 1 : 2);
    const f3 G1 = sel3(id, p0, p1, p2);
    const f3 G2 = sel3(id, p1, p2, p0);
    const f3 G3 = sel3(id, p2, p0, p1);
    const f3 c1 = clip_edge(G2, G1), c2 = clip_edge(G3, G1);
    out[0][0] = c1; out[0][1] = G2; out[0][2] = G3;
    out[1][0] = c1; out[1][1] = G3; out[1][2] = c2;
    return 2;
  }
  // two behind: g2 < g3 behind, g1 valid; each behind vertex is replaced
  const int v = !b0 ? 0 : (!b1 ? 1 : 2);
  const int first = b0 ? 0 : 1;                 // lower behind index
  const f3 G1 = sel3(v, p0, p1, p2);
  const f3 G2 = sel3(first, p0, p1, p2);
  const f3 G3 = b2 ? p2 : p1;                   // higher behind index
  const f3 c1 = clip_edge(G2, G1), c2 = clip_edge(G3, G1);
  out[0][0] = v == 0 ? G1 : c1;
  out[0][1] = v == 1 ? G1 : (first == 1 ? c1 : c2);
  out[0][2] = v == 2 ? G1 : c2;
  return 1;
}

// backface_cull_expanded, cl2.cl:488-494 (1 = keep).
__device__ __forceinline__ bool front_facing(f3 p0, f3 p1, f3 p2) {
  return cross(p1 - p0, p2 - p0).z < 0.f;
}

// calc_min_max, cl2.cl:420-441.
__device__ __forceinline__ void calc_min_max(const float xs[3], const float ys[3], float width,
                                             float height, float mm[4]) {
  float x0 = roundf(xs[0]), x1 = roundf(xs[1]), x2 = roundf(xs[2]);
  float y0 = roundf(ys[0]), y1 = roundf(ys[1]), y2 = roundf(ys[2]);
  mm[0] = clampf(min3f(x0, x1, x2) - 1, 0.0f, width - 1);
  mm[1] = clampf(max3f(x0, x1, x2), 0.0f, width - 1);
  mm[2] = clampf(min3f(y0, y1, y2) - 1, 0.0f, height - 1);
  mm[3] = clampf(max3f(y0, y1, y2), 0.0f, height - 1);
}

// calc_rconstant_v, cl2.cl:403-406, on rounded coordinates.
// The determinant of rounded (integer-valued) coordinates is integer-valued,
// so its reciprocal is rcp_cr's.
__device__ __forceinline__ float rconstant(f3 x, f3 y) {
  return rcp_cr(x.y * y.z + x.x * (y.y - y.z) - x.z * y.y + (x.z - x.y) * y.x);
}

// interpolate_get_const, cl2.cl:408-418.
__device__ __forceinline__ void plane_const(f3 f, f3 x, f3 y, float rc, float& A, float& B, float& C) {
  A = ((f.y * y.z + f.x * (y.y - y.z) - f.z * y.y + (f.z - f.y) * y.x) * rc);
  B = (-(f.y * x.z + f.x * (x.y - x.z) - f.z * x.y + (f.z - f.y) * x.x) * rc);
  C = f.x - A * x.x - B * y.x;
}

// point_in_tri, cl2.cl:4798-4807 (on rounded vertex coordinates).
struct EdgeSetup { float A2s; float s0, sx, sy, t0, tx, ty; };
__device__ __forceinline__ bool point_in_tri(float px, float py, float p0x, float p0y, float p1x,
                                             float p1y, float p2x, float p2y) {
  float A = 0.5f * (-p1y * p2x + p0y * (-p1x + p2x) + p0x * (p1y - p2y) + p1x * p2y);
  float sign = A < 0 ? -1.f : 1.f;
  float s = (p0y * p2x - p0x * p2y + (p2y - p0y) * px + (p0x - p2x) * py) * sign;
  float t = (p0x * p1y - p0y * p1x + (p0y - p1y) * px + (p1x - p0x) * py) * sign;
  return s > -0.0001f && t > -0.0001f && (s + t) < 2.0001f * A * sign;
}

// wang_hash / rand_xorshift, cl2.cl:1919-1936.
__device__ __forceinline__ uint32_t wang_hash(uint32_t seed) {
  seed = (seed ^ 61u) ^ (seed >> 16);
  seed *= 9u;
  seed = seed ^ (seed >> 4);
  seed *= 0x27d4eb2du;
  seed = seed ^ (seed >> 15);
  return seed;
}
__device__ __forceinline__ uint32_t xorshift(uint32_t s) {
  s ^= (s << 13);
  s ^= (s >> 17);
  s ^= (s << 5);
  return s;
}

// rational_acos, cl2.cl:2469-2477.
__device__ __forceinline__ float rational_acos(float x) {
  const float a = -0.939115566365855f;
  const float b = 0.9217841528914573f;
  const float c = -1.2845906244690837f;
  const float d = 0.295624144969963174f;
  return kPi / 2.f + (a * x + b * x * x * x) / (1.f + c * x * x + d * pow4(x));
}
// The same for x in [0, 1] (kernel3's n.h), with the quotient from the
// correctly rounded reciprocal of the denominator.  There the denominator
// lies in [0.0110, 1]; the numerator x (a + b x^2) is +0 at x = 0 (the
// quotient's +0 then equals IEEE's), has magnitude >= 0.017 x elsewhere, so
// for x >= 2^-30 it is >= 2^-36 and div_rcp is the correctly rounded
// quotient, and for 0 < x < 2^-30 both quotients are below 2^-29, which
// pi/2 absorbs (half an ulp of pi/2 is 2^-24): every x gives the bits of
// rational_acos(x).
__device__ __forceinline__ float rational_acos_unit(float x) {
  const float a = -0.939115566365855f;
  const float b = 0.9217841528914573f;
  const float c = -1.2845906244690837f;
  const float d = 0.295624144969963174f;
  const float den = 1.f + c * x * x + d * pow4(x);
  return kPi / 2.f + div_rcp(a * x + b * x * x * x, den, rcp_rn(den));
}

// log2_approx, cl2.cl:1498-1505.
__device__ __forceinline__ float log2_approx(float val) {
  int x = __float_as_int(val);
  float log_2 = (float)(((x >> 23) & 255) - 128);
  x &= ~(255 << 23);
  x += 127 << 23;
  float u = __int_as_float(x);
  log_2 += ((-0.3358287811f) * u + 2.0f) * u - 0.65871759316667f;
  return log_2;
}

// gamma_transform_approx / gamma_inverse_approx, cl2.cl:5763-5782.
__device__ __forceinline__ float gamma_fwd(float c) {
  return 0.012522878f * c + 0.682171111f * c * c + 0.305306011f * c * c * c;
}
__device__ __forceinline__ float gamma_inv(float c) {
  float s1 = sqrt_cr(c), s2 = sqrt_cr(s1), s3 = sqrt_cr(s2);
  return 0.585122381f * s1 + 0.783140355f * s2 - 0.368262736f * s3;
}
// gamma_inv of three channels.  When every channel of every lane lies in
// [2^-100, FLT_MAX] so do the nested roots (>= 2^-50), and the short root
// needs no per-root check.
__device__ __forceinline__ float gamma_inv_core(float c) {
  float s1 = sqrt_core(c), s2 = sqrt_core(s1), s3 = sqrt_core(s2);
  return 0.585122381f * s1 + 0.783140355f * s2 - 0.368262736f * s3;
}
__device__ __forceinline__ void gamma_inv3(float& x, float& y, float& z) {
  const bool ok = x >= 0x1p-100f && x <= 0x1.fffffep127f && y >= 0x1p-100f && y <= 0x1.fffffep127f &&
                  z >= 0x1p-100f && z <= 0x1.fffffep127f;
  if (__all(ok)) {
    x = gamma_inv_core(x);
    y = gamma_inv_core(y);
    z = gamma_inv_core(z);
  } else {
    x = gamma_inv(x);
    y = gamma_inv(y);
    z = gamma_inv(z);
  }
}

// ret_cubeface, cl2.cl:1745-1790.
__device__ __forceinline__ int cubeface(f3 point, f3 light) {
  f3 rel = point - light;
  float ax = fabsf(rel.x), ay = fabsf(rel.y), az = fabsf(rel.z);
  if (ax >= ay && ax >= az) return rel.x < 0 ? 4 : 5;
  if (ay > ax && ay >= az) return rel.y < 0 ? 1 : 3;
  if (az > ax && az > ay) { if (rel.z < 0) return 2; }
  return 0;
}

// get_barycentric (cl2.cl:5372-5384) split into the terms of the triangle,
// formed once per pixel, and the evaluation at a point.  v and w are the two
// numerators over denom: when denom lies in [2^-60, 2^60] and each numerator
// is +0 or of magnitude in [2^-60, 2^60] for every lane of the wavefront, the
// quotients by the correctly rounded reciprocal of denom are the correctly
// rounded ones; otherwise the wavefront divides.
struct BaryPre {
  f3 a, v0, v1;
  float d00, d01, d11, denom, rden;
  bool fast;
};
__device__ __forceinline__ BaryPre bary_pre(f3 a, f3 b, f3 c) {
  BaryPre B;
  B.a = a;
  B.v0 = b - a;
  B.v1 = c - a;
  B.d00 = dot(B.v0, B.v0);
  B.d01 = dot(B.v0, B.v1);
  B.d11 = dot(B.v1, B.v1);
  B.denom = B.d00 * B.d11 - B.d01 * B.d01;
  B.fast = rcp_safe(B.denom);
  B.rden = rcp_rn(B.denom);
  return B;
}
__device__ __forceinline__ bool quot_ok(float n) {
  return __float_as_uint(n) == 0u || (fabsf(n) >= 0x1p-60f && fabsf(n) <= 0x1p60f);
}
__device__ __forceinline__ void bary_eval(const BaryPre& B, f3 p, float& u, float& v, float& w) {
  const f3 v2 = p - B.a;
  const float d20 = dot(v2, B.v0), d21 = dot(v2, B.v1);
  const float nv = B.d11 * d20 - B.d01 * d21;
  const float nw = B.d00 * d21 - B.d01 * d20;
  if (__all(B.fast && quot_ok(nv) && quot_ok(nw))) {
    v = div_rcp(nv, B.denom, B.rden);
    w = div_rcp(nw, B.denom, B.rden);
  } else {
    v = nv / B.denom;
    w = nw / B.denom;
  }
  u = 1.0f - v - w;
}

// get_barycentric, cl2.cl:5372-5384.
__device__ __forceinline__ void barycentric(f3 p, f3 a, f3 b, f3 c, float& u, float& v, float& w) {
  f3 v0 = b - a, v1 = c - a, v2 = p - a;
  float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
  float d20 = dot(v2, v0), d21 = dot(v2, v1);
  float denom = d00 * d11 - d01 * d01;
  v = (d11 * d20 - d01 * d21) / denom;
  w = (d00 * d21 - d01 * d20) / denom;
  u = 1.0f - v - w;
}

// float_to_short / encode_normal, cl2.cl:5592-5630 ("POW2e16-1" expands
// without parentheses; the ushort conversion saturates).
__device__ __forceinline__ uint32_t to_short(float r) {
  float f = ((r + 1) / 2) * 65536 - 1;
  if (!(f > 0.f)) return 0u;
  if (f >= 65535.f) return 65535u;
  return (uint32_t)f;
}
__device__ __forceinline__ uint32_t encode_normal(f3 v) {
  if (v.x * v.x + v.y * v.y < 0.0001f) v.x = 0.01f;
  float inv = inv_sqrt_cr(v.x * v.x + v.y * v.y);
  float sq = sqrt_cr(fmaxf(v.z * 0.5f + 0.5f, 0.f));
  return to_short((v.x * inv) * sq) | (to_short((v.y * inv) * sq) << 16);
}

__device__ __forceinline__ uint32_t quant8(float v) {
  float c = clampf(v, 0.f, 1.f);
  return (uint32_t)floorf(c * 255.f + 0.5f);
}

}  // namespace ocr
