/*
 * ocr_oracle.c — CPU restatement of OpenCLRenderer's raster + shading path.
 *
 * TEST INFRASTRUCTURE ONLY (see ocr_oracle.h for who may use it and for the
 * arithmetic contract).  Every function cites the cl2.cl / host lines it
 * follows.  Built with gcc -ffp-contract=off: no expression below may be
 * contracted, every fused multiply-add is an explicit fmaf()/fma().
 */
#include "ocr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define DEPTH_FAR    350000.0f          /* cl2.cl:17 */
#define MULINT_F     4294967296.0f      /* (float)UINT_MAX, cl2.cl:19 */
#define ICUTOFF      20                 /* depth_icutoff, cl2.cl:21-23 / main.cpp:80 */
#define OP_SIZE      500                /* cl2.cl:4247 */
#define OP_SIZE_L    300                /* cl2.cl:4249 */
#define BUF_ERROR    20u                /* cl2.cl:5387 */
#define M_PI_F       3.1415927f         /* cl2.cl:14 */
#define AMBIENT_F    0.2f               /* main.cpp:81 */
#define SSAO_RAD_F   2.f                /* main.cpp:82 */
#define SSAO_DIV_F   2.5f               /* cl2.cl:2253-2255 */
#define SHADOWBIAS_I 50                 /* ocl.h:234 */
#define MIPL         4                  /* MIP_LEVELS, cl2.cl:5 */
#define MIP_BIAS_F   1.1f               /* cl2.cl:5750-5752 */

/* Planted mistakes for the mutation test (tests/test_oracle_pins.py builds
 * copies with -DOCR_ORACLE_PLANT=k and checks that each fails a pin).  The
 * oracle proper is built with 0: every PLANT(k) below is false. */
#ifndef OCR_ORACLE_PLANT
#define OCR_ORACLE_PLANT 0
#endif
#define PLANT(k) (OCR_ORACLE_PLANT == (k))

typedef struct { float x, y, z; } f3;
typedef struct { float x, y, z, w; } f4;

static double now_ms(void) {
    struct timespec ts; clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* ---------------------------------------------------------------- vectors */
static inline f3 V3(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 add3(f3 a, f3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub3(f3 a, f3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 mul3(f3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline f3 mul33(f3 a, f3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 div3(f3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
static inline f3 neg3(f3 a) { return V3(-a.x, -a.y, -a.z); }
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross3(f3 a, f3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline f3 fnorm3(f3 v) { float r = 1.0f / sqrtf(dot3(v, v)); return mul3(v, r); }
static inline float flen3(f3 v) { return sqrtf(dot3(v, v)); }
static inline f3 fma3(f3 a, float b, f3 c) {
    return V3(fmaf(a.x, b, c.x), fmaf(a.y, b, c.y), fmaf(a.z, b, c.z));
}
static inline f3 ld3(const float* p) { return V3(p[0], p[1], p[2]); }
/* OpenCL min/max/clamp on floats: fmin/fmax */
static inline float clampf_(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float min3f(float a, float b, float c) { return fminf(fminf(a, b), c); }   /* cl2.cl:44-47 */
static inline float max3f(float a, float b, float c) { return fmaxf(fmaxf(a, b), c); }   /* cl2.cl:49-52 */

static inline uint32_t sat_u32(float f) {
    if (!(f > 0.0f)) return 0u;                 /* negatives and NaN */
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}
static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* ------------------------------------------------------ binary64 helpers */
/* exp in binary64: Cody-Waite reduction by ln2, degree-13 Taylor polynomial
 * (|r| <= ln2/2, truncation error < 2e-17), scaled by 2^k. */
void oracle_exp_n(const float* x, float* y, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) y[i] = oracle_exp(x[i]);
}

/* tan in binary64: reduction by pi/2 (two-constant Cody-Waite, exact for the
 * small quotients that occur here), Taylor sin/cos to degree 19/18. */
static double ocrx_tan(double x) {
    const double TWO_OVER_PI = 6.36619772367581382433e-01;
    const double PIO2_HI = 1.57079632673412561417e+00;
    const double PIO2_LO = 6.07710050650619224932e-11;
    double k = floor(fma(x, TWO_OVER_PI, 0.5));
    double r = fma(-k, PIO2_LO, fma(-k, PIO2_HI, x));
    double r2 = r * r;
    double s = -1.0 / 121645100408832000.0;        /* -1/19! */
    s = fma(s, r2, 1.0 / 355687428096000.0);
    s = fma(s, r2, -1.0 / 1307674368000.0);
    s = fma(s, r2, 1.0 / 6227020800.0);
    s = fma(s, r2, -1.0 / 39916800.0);
    s = fma(s, r2, 1.0 / 362880.0);
    s = fma(s, r2, -1.0 / 5040.0);
    s = fma(s, r2, 1.0 / 120.0);
    s = fma(s, r2, -1.0 / 6.0);
    s = fma(s * r2, r, r);
    double c = 1.0 / 6402373705728000.0;           /* 1/18! */
    c = fma(c, r2, -1.0 / 20922789888000.0);
    c = fma(c, r2, 1.0 / 87178291200.0);
    c = fma(c, r2, -1.0 / 479001600.0);
    c = fma(c, r2, 1.0 / 3628800.0);
    c = fma(c, r2, -1.0 / 40320.0);
    c = fma(c, r2, 1.0 / 720.0);
    c = fma(c, r2, -1.0 / 24.0);
    c = fma(c, r2, 0.5);
    c = fma(-c, r2, 1.0);
    int64_t ki = (int64_t)k;
    return (ki & 1) ? -c / s : s / c;
}

/* exp in binary32 (the arithmetic contract's exp; the reference's exp under
 * -cl-fast-relaxed-math has implementation-defined accuracy): k =
 * rint(x log2 e), r = x - k ln2 in two Cody-Waite steps (k * ln2_hi exact for
 * |k| < 2^11), e^r = 1 + r p(r), p of degree 5 fitted to (e^r - 1) / r on
 * [-ln2/2, ln2/2] (relative error of e^r 3.4e-9), then ldexp.  Within one
 * ulp of the correctly rounded exp on [-104, 88.75]. */
float oracle_exp(float x) {
    if (!(x >= -104.0f)) return x != x ? x : 0.0f;
    if (x > 88.75f) return INFINITY;
    const float k = rintf(x * 1.44269502f);
    float r = fmaf(-k, 0.693145752f, x);
    r = fmaf(-k, 1.42860677e-06f, r);
    float p = 0.001399628003127873f;
    p = fmaf(p, r, 0.008374124765396118f);
    p = fmaf(p, r, 0.04166550189256668f);
    p = fmaf(p, r, PLANT(13) ? 0.17f : 0.16666433215141296f);
    p = fmaf(p, r, 0.5000000596046448f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, (int)k);
}
float oracle_tan(float x) { return (float)ocrx_tan((double)x); }
/* native_powr(x, 5.f) (cl2.cl:6236) and pow(x, 4.f) (cl2.cl:2480) */
float oracle_powr5(float x) { double d = x; double d2 = d * d; return (float)((d2 * d2) * d); }
float oracle_pow4(float x)  { double d = x; double d2 = d * d; return (float)(d2 * d2); }

/* --------------------------------------------------- rotations, cl2.cl */
/* rot, cl2.cl:220-236: extrinsic XYZ rotation of point about c_pos. */
static inline f3 rot(f3 point, f3 c_pos, const oracle_sincos* sc) {
    f3 c = V3(sc->c[0], sc->c[1], sc->c[2]);
    f3 s = V3(sc->s[0], sc->s[1], sc->s[2]);
    f3 rel = sub3(point, c_pos);
    f3 r;
    float inner = fmaf(s.z, rel.y, c.z * rel.x);
    r.x = fmaf(c.y, inner, -s.y * rel.z);
    r.y = fmaf(s.x, fmaf(c.y, rel.z, s.y * inner), c.x * fmaf(c.z, rel.y, -s.z * rel.x));
    r.z = fmaf(c.x, fmaf(c.y, rel.z, s.y * inner), -s.x * fmaf(c.z, rel.y, -s.z * rel.x));
    return r;
}

/* back_rot with c_pos == 0, cl2.cl:272-349 ("transposed, with mads"). */
static inline f3 back_rot(f3 rel, const oracle_sincos* sc) {
    f3 c = V3(sc->c[0], sc->c[1], sc->c[2]);
    f3 s = V3(sc->s[0], sc->s[1], sc->s[2]);
    f3 r;
    r.x = c.z * fmaf(c.y, rel.x, fmaf(s.x, s.y * rel.y, c.x * s.y * rel.z))
        + s.z * (s.x * rel.z - c.x * rel.y);
    r.y = fmaf(s.z, c.y * rel.x,
               fmaf(fmaf(c.x, c.z, s.x * s.y * s.z), rel.y,
                    fmaf(-s.x, c.z, c.x * s.y * s.z) * rel.z));
    r.z = fmaf(-s.y, rel.x, c.y * fmaf(s.x, rel.y, c.x * rel.z));
    return r;
}

/* rot_quat, cl2.cl:350-357 (fast_normalize of the quaternion first). */
static inline f3 rot_quat(f3 point, f4 q) {
    float d = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    float inv = 1.0f / sqrtf(d);
    q.x = q.x * inv; q.y = q.y * inv; q.z = q.z * inv; q.w = q.w * inv;
    f3 qv = V3(q.x, q.y, q.z);
    f3 t = mul3(cross3(qv, point), 2.f);
    return add3(add3(point, mul3(t, q.w)), cross3(qv, t));
}

/* back_rot_quat, cl2.cl:359-370. */
static inline f3 back_rot_quat(f3 point, f4 q) {
    f4 conj = {-q.x, -q.y, -q.z, q.w};
    float len_sq = conj.x * conj.x + conj.y * conj.y + conj.z * conj.z + conj.w * conj.w;
    f4 n = {conj.x / len_sq, conj.y / len_sq, conj.z / len_sq, conj.w / len_sq};
    return rot_quat(point, n);
}

/* -------------------------------------------- clip / project, cl2.cl */
/* depth_project, cl2.cl:535-545 */
static inline f3 depth_project1(f3 r, float width, float height, float fovc) {
    float f = fovc / r.z;
    return V3(fmaf(r.x, f, width / 2.f), fmaf(r.y, f, height / 2.f), r.z);
}

/* generate_new_triangles, cl2.cl:577-663 */
static int generate_new_triangles(const f3 pts[3], f3 ret[2][3]) {
    int id_valid = 0, ids_behind[3], n_behind = 0;
    for (int i = 0; i < 3; i++) {
        if (pts[i].z <= (float)ICUTOFF || pts[i].z > DEPTH_FAR) ids_behind[n_behind++] = i;
        else id_valid = i;
    }
    if (n_behind > 2) return 0;
    if (n_behind == 0) { ret[0][0] = pts[0]; ret[0][1] = pts[1]; ret[0][2] = pts[2]; return 1; }
    int g1, g2, g3;
    if (n_behind == 1) {
        int id = ids_behind[0];
        g1 = id;
        g2 = (id + 1) >= 3 ? id - 2 : id + 1;
        g3 = (id + 2) >= 3 ? id - 1 : id + 2;
    } else {
        g2 = ids_behind[0]; g3 = ids_behind[1]; g1 = id_valid;
    }
    f3 p1 = add3(pts[g2], div3(mul3(sub3(pts[g1], pts[g2]), (float)ICUTOFF - pts[g2].z),
                               pts[g1].z - pts[g2].z));
    f3 p2 = add3(pts[g3], div3(mul3(sub3(pts[g1], pts[g3]), (float)ICUTOFF - pts[g3].z),
                               pts[g1].z - pts[g3].z));
    if (n_behind == 1) {
        ret[0][0] = p1; ret[0][1] = pts[g2]; ret[0][2] = pts[g3];
        ret[1][0] = p1; ret[1][1] = pts[g3]; ret[1][2] = p2;
        return 2;
    }
    ret[0][ids_behind[0]] = p1;
    ret[0][ids_behind[1]] = p2;
    ret[0][id_valid] = pts[g1];
    return 1;
}

/* full_rotate_quat, cl2.cl:700-729 */
static int full_rotate_quat(f3 v1, f3 v2, f3 v3, f3 out[2][3], f3 c_pos, const oracle_sincos* sc,
                            f3 offset, f4 q, float scale, float fovc, float w, float h) {
    f3 pr[3];
    pr[0] = rot(add3(rot_quat(mul3(v1, scale), q), offset), c_pos, sc);
    pr[1] = rot(add3(rot_quat(mul3(v2, scale), q), offset), c_pos, sc);
    pr[2] = rot(add3(rot_quat(mul3(v3, scale), q), offset), c_pos, sc);
    f3 tris[2][3];
    int n = generate_new_triangles(pr, tris);
    for (int t = 0; t < n; t++)
        for (int i = 0; i < 3; i++) out[t][i] = depth_project1(tris[t][i], w, h, fovc);
    return n;
}

/* -------------------------------------------------- raster helpers */
/* calc_min_max, cl2.cl:420-441 */
static inline void calc_min_max(const f3 p[3], float width, float height, float ret[4]) {
    float x0 = roundf(p[0].x), x1 = roundf(p[1].x), x2 = roundf(p[2].x);
    float y0 = roundf(p[0].y), y1 = roundf(p[1].y), y2 = roundf(p[2].y);
    ret[0] = clampf_(min3f(x0, x1, x2) - 1, 0.0f, width - 1);
    ret[1] = clampf_(max3f(x0, x1, x2), 0.0f, width - 1);
    ret[2] = clampf_(min3f(y0, y1, y2) - 1, 0.0f, height - 1);
    ret[3] = clampf_(max3f(y0, y1, y2), 0.0f, height - 1);
}

/* calc_rconstant_v, cl2.cl:403-406 */
static inline float calc_rconstant_v(f3 x, f3 y) {
    return 1.0f / (x.y * y.z + x.x * (y.y - y.z) - x.z * y.y + (x.z - x.y) * y.x);
}

/* interpolate_get_const, cl2.cl:408-418 */
static inline void interpolate_get_const(f3 f, f3 x, f3 y, float rc, float* A, float* B, float* C) {
    *A = ((f.y * y.z + f.x * (y.y - y.z) - f.z * y.y + (f.z - f.y) * y.x) * rc);
    *B = (-(f.y * x.z + f.x * (x.y - x.z) - f.z * x.y + (f.z - f.y) * x.x) * rc);
    *C = f.x - (*A) * x.x - (*B) * y.x;
}

/* backface_cull_expanded, cl2.cl:488-494: 1 = kept (front facing) */
static inline int backface_keep(f3 p0, f3 p1, f3 p2) {
    return cross3(sub3(p1, p0), sub3(p2, p0)).z < 0.f;
}

/* point_in_tri, cl2.cl:4798-4807 */
static inline int point_in_tri(float px, float py, float p0x, float p0y, float p1x, float p1y,
                               float p2x, float p2y) {
    float A = 0.5f * (-p1y * p2x + p0y * (-p1x + p2x) + p0x * (p1y - p2y) + p1x * p2y);
    float sign = A < 0 ? -1.f : 1.f;
    float s = (p0y * p2x - p0x * p2y + (p2y - p0y) * px + (p0x - p2x) * py) * sign;
    float t = (p0x * p1y - p0y * p1x + (p0y - p1y) * px + (p1x - p0x) * py) * sign;
    return s > -0.0001f && t > -0.0001f && (s + t) < 2.0001f * A * sign;
}

/* Walk of one fragment, cl2.cl:5016-5126 (kernel1), :5419-5545 (kernel2),
 * :5152-5245 (kernel1_realtime_shadowing).  Calls visit(x, y, depth). */
typedef void (*visit_fn)(void* ctx, int x, int y, uint32_t d);

static inline void walk_fragment(const f3 tp[3], float ewidth, float eheight, uint32_t distance,
                                 int op_size, float rconst, visit_fn visit, void* ctx) {
    float mm[4];
    calc_min_max(tp, ewidth, eheight, mm);
    int width = (int)(mm[1] - mm[0]);
    if (width <= 0) return;   /* unreachable: prearrange emits no fragment for a 0-area bbox */
    int pixel_along = op_size * (int)distance;
    f3 xpv = V3(roundf(tp[0].x), roundf(tp[1].x), roundf(tp[2].x));
    f3 ypv = V3(roundf(tp[0].y), roundf(tp[1].y), roundf(tp[2].y));
    f3 depths = V3(1.0f / (tp[0].z / DEPTH_FAR), 1.0f / (tp[1].z / DEPTH_FAR),
                   1.0f / (tp[2].z / DEPTH_FAR));
    int pcount = -1;
    float x = (float)((pixel_along + 0) % width) + mm[0] - 1;
    float y = floorf((float)(pixel_along + pcount) / (float)width) + mm[2];
    float A, B, C;
    interpolate_get_const(depths, xpv, ypv, rconst, &A, &B, &C);
    float iwidth = 1.f / (float)width;
    float rwm = (float)((pixel_along + pcount) % width);
    while (pcount < op_size) {
        pcount++;
        x += 1;
        rwm += 1.f;
        if (rwm >= (float)width) rwm = 0;
        float ty = y;
        y = floorf(fmaf((float)(pixel_along + pcount), iwidth, mm[2]));
        x = y != ty ? rwm + mm[0] : x;
        if (y >= mm[3]) break;
        if (x >= mm[1] || x < mm[0] || y < mm[2]) continue;
        if (point_in_tri(x, y, xpv.x, ypv.x, xpv.y, ypv.y, xpv.z, ypv.z)) {
            float fmydepth = fmaf(A, x, fmaf(B, y, C));
            uint32_t mydepth = sat_u32(MULINT_F / fmydepth);
            visit(ctx, (int)x, (int)y, mydepth);
        }
    }
}

/* ------------------------------------------------------- prearrange */
typedef struct { uint32_t tri, chunk, cut; float rconst; uint32_t oid, sub; } frag_rec;
typedef struct { f3 p[3]; } cut_tri;

typedef struct {
    cut_tri*  cuts;   uint32_t ncut_reserved;   /* c_id space (includes culled) */
    frag_rec* frags;  uint32_t nfrag;
    uint32_t  nsurvive;
} prelist;

/* One triangle's clip + cull + fragment count, cl2.cl:4292-4408.  Returns
 * `num` (cut triangles reserved) and fills keep[], cuts and counts. */
typedef struct {
    int num;
    int keep[2];
    f3 tp[2][3];
    float rconst[2];
    uint32_t nfrag[2];
} tri_setup;

static void prearrange_one(const oracle_scene* sc, const ocr_view* v, const oracle_sincos* cam,
                           uint32_t id, tri_setup* ts) {
    ts->num = 0;
    const ocr_triangle* T = &sc->tris[id];
    uint32_t o_id = T->vertices[0].object_id;
    if (o_id >= sc->nobj) return;
    const ocr_obj_desc* G = &sc->objs[o_id];
    const float efov = v->fov_const;
    const float ewidth = (float)v->width, eheight = (float)v->height;
    f3 wp = ld3(G->world_pos);
    f3 cpos = ld3(v->c_pos);
    if (flen3(sub3(wp, cpos)) > DEPTH_FAR) return;
    f4 q = {G->world_rot_quat[0], G->world_rot_quat[1], G->world_rot_quat[2], G->world_rot_quat[3]};
    int num = full_rotate_quat(ld3(T->vertices[0].pos), ld3(T->vertices[1].pos),
                               ld3(T->vertices[2].pos), ts->tp, cpos, cam, wp, q, G->scale,
                               efov, ewidth, eheight);
    ts->num = num;
    int two_sided = (G->feature_flag & OCR_FEATURE_TWO_SIDED) > 0;
    for (int i = 0; i < num; i++) {
        f3* p = ts->tp[i];
        ts->keep[i] = 0;
        int valid = two_sided || backface_keep(p[0], p[1], p[2]);
        int cond = (p[0].x < 0 && p[1].x < 0 && p[2].x < 0) ||
                   (p[0].x >= ewidth && p[1].x >= ewidth && p[2].x >= ewidth) ||
                   (p[0].y < 0 && p[1].y < 0 && p[2].y < 0) ||
                   (p[0].y >= eheight && p[1].y >= eheight && p[2].y >= eheight);
        if (!valid || cond) continue;
        f3 xpv = V3(roundf(p[0].x), roundf(p[1].x), roundf(p[2].x));
        f3 ypv = V3(roundf(p[0].y), roundf(p[1].y), roundf(p[2].y));
        ts->rconst[i] = calc_rconstant_v(xpv, ypv);
        float mm[4];
        calc_min_max(p, ewidth, eheight, mm);
        float area = (mm[1] - mm[0]) * (mm[3] - mm[2]);
        int thread_num = (int)ceilf(area / (float)OP_SIZE);
        ts->nfrag[i] = (uint32_t)thread_num;
        ts->keep[i] = 1;
    }
}

static void prearrange_all(const oracle_scene* sc, const ocr_view* v, const oracle_sincos* cam,
                           int nthreads, prelist* pl) {
    uint32_t T = sc->ntri;
    tri_setup* ts = (tri_setup*)malloc(sizeof(tri_setup) * (T ? T : 1));
    #pragma omp parallel for schedule(static) num_threads(nthreads > 1 ? nthreads : 1)
    for (long i = 0; i < (long)T; i++) prearrange_one(sc, v, cam, (uint32_t)i, &ts[i]);
    /* Sequential work-item order: atomic_add(id_cutdown_tris, num) then
     * atomic_add(id_buffer_atomc, thread_num) per survivor (cl2.cl:4342, :4388). */
    uint32_t ncut = 0, nfrag = 0, nsurv = 0;
    uint32_t* cbase = (uint32_t*)malloc(sizeof(uint32_t) * (T ? T : 1));
    uint32_t* fbase = (uint32_t*)malloc(sizeof(uint32_t) * (T ? T : 1));
    for (uint32_t i = 0; i < T; i++) {
        cbase[i] = ncut; fbase[i] = nfrag;
        ncut += (uint32_t)ts[i].num;
        for (int k = 0; k < ts[i].num; k++)
            if (ts[i].keep[k]) { nfrag += ts[i].nfrag[k]; if (ts[i].nfrag[k]) nsurv++; }
    }
    pl->cuts = (cut_tri*)malloc(sizeof(cut_tri) * (ncut ? ncut : 1));
    pl->frags = (frag_rec*)malloc(sizeof(frag_rec) * (nfrag ? nfrag : 1));
    pl->ncut_reserved = ncut; pl->nfrag = nfrag; pl->nsurvive = nsurv;
    #pragma omp parallel for schedule(static) num_threads(nthreads > 1 ? nthreads : 1)
    for (long i = 0; i < (long)T; i++) {
        uint32_t f = fbase[i];
        for (int k = 0; k < ts[i].num; k++) {
            if (!ts[i].keep[k]) continue;
            uint32_t c_id = cbase[i] + (uint32_t)k;
            for (int j = 0; j < 3; j++) pl->cuts[c_id].p[j] = ts[i].tp[k][j];
            uint32_t oid = sc->tris[i].vertices[0].object_id;
            for (uint32_t a = 0; a < ts[i].nfrag[k]; a++) {
                frag_rec* r = &pl->frags[f++];
                r->tri = (uint32_t)i; r->chunk = a; r->cut = c_id; r->rconst = ts[i].rconst[k];
                r->oid = oid;
                r->sub = (uint32_t)k;
            }
        }
    }
    free(ts); free(cbase); free(fbase);
}

/* ------------------------------------------------- kernel1 / kernel2 */
typedef struct { uint32_t* depth; uint32_t W; uint64_t samples; int atomic; } k1ctx;
static void k1_visit(void* c, int x, int y, uint32_t d) {
    k1ctx* k = (k1ctx*)c;
    uint32_t* ft = &k->depth[(size_t)y * k->W + (size_t)x];   /* cl2.cl:5122 */
    k->samples++;
    if (k->atomic) {
        uint32_t old = __atomic_load_n(ft, __ATOMIC_RELAXED);
        while (d < old && !__atomic_compare_exchange_n(ft, &old, d, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
    } else if (d < *ft) {
        *ft = d;
    }
}

typedef struct { const uint32_t* depth; uint32_t* ids; uint32_t W; uint32_t frag; int atomic; } k2ctx;
static void k2_visit(void* c, int x, int y, uint32_t mydepth) {
    k2ctx* k = (k2ctx*)c;
    size_t p = (size_t)y * k->W + (size_t)x;
    uint32_t val = k->depth[p];
    /* cl2.cl:5534: u32 window with wrap-around */
    int c2 = mydepth > val - BUF_ERROR && mydepth < val + BUF_ERROR;
    if (!c2) return;
    if (k->atomic) {
        /* last writer in sequential order == largest fragment index */
        uint32_t old = __atomic_load_n(&k->ids[p], __ATOMIC_RELAXED);
        uint32_t nv = k->frag + 1;
        while ((old == 0 || nv > old) &&
               !__atomic_compare_exchange_n(&k->ids[p], &old, nv, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
    } else {
        k->ids[p] = k->frag + 1;        /* 0 = none */
    }
}

/* -------------------------------------------------- textures, cl2.cl */
/* read_tex_array, cl2.cl:785-818 */
static inline f4 read_tex_array(float cx, float cy, uint32_t tid, const uint32_t* num,
                                const uint32_t* size, const uint8_t* array) {
    int nv = (int)num[tid];
    int slice = nv >> 16;
    int which = nv & 0x0000FFFF;
    const float max_tex_size = 2048;
    float width = (float)size[slice];
    float hnum = floorf(max_tex_size / width);
    float tnumy = floorf((float)which / hnum);
    float tnumx = fmaf(-tnumy, hnum, (float)which);
    cx = clampf_(cx, 0.001f, width - 0.001f);
    cy = clampf_(cy, 0.001f, width - 0.001f);
    int ix = (int)fmaf(tnumx, width, cx);
    int iy = (int)fmaf(tnumy, width, cy);
    const uint8_t* t = &array[((size_t)slice * 2048 * 2048 + (size_t)iy * 2048 + (size_t)ix) * 4];
    f4 r = {(float)t[0], (float)t[1], (float)t[2], (float)t[3]};
    return r;
}

/* read_tex_array_all_precalculated, cl2.cl:820-851 */
static inline f4 read_tex_pre(float cx, float cy, int which, int slice, float width,
                              const uint8_t* array) {
    const float imax_tex_size = 1.f / 2048;
    float ihnum = width * imax_tex_size;
    float tnumy = floorf((float)which * ihnum);
    float tnumx = (float)which - tnumy / ihnum;
    cx = clampf_(cx, 0.001f, width - 0.001f);
    cy = clampf_(cy, 0.001f, width - 0.001f);
    int ix = (int)fmaf(tnumx, width, cx);
    int iy = (int)fmaf(tnumy, width, cy);
    const uint8_t* t = &array[((size_t)slice * 2048 * 2048 + (size_t)iy * 2048 + (size_t)ix) * 4];
    f4 r = {(float)t[0], (float)t[1], (float)t[2], (float)t[3]};
    return r;
}

/* write_tex_array, cl2.cl:856-889 (uchar4 store truncates each u32) */
static inline void write_tex_array(const uint32_t col[4], float cx, float cy, uint32_t tid,
                                   const uint32_t* num, const uint32_t* size, uint8_t* array) {
    int nv = (int)num[tid];
    int slice = nv >> 16;
    int which = nv & 0x0000FFFF;
    const float max_tex_size = 2048;
    float width = (float)size[slice];
    float hnum = floorf(max_tex_size / width);
    float tnumy = floorf((float)which / hnum);
    float tnumx = fmaf(-tnumy, hnum, (float)which);
    float tx = tnumx * width, ty = tnumy * width;
    cx = fmodf(cx, width); cy = fmodf(cy, width);
    cx = clampf_(cx, 0.001f, width - 0.001f);
    cy = clampf_(cy, 0.001f, width - 0.001f);
    int ix = (int)(tx + cx), iy = (int)(ty + cy);
    uint8_t* t = &array[((size_t)slice * 2048 * 2048 + (size_t)iy * 2048 + (size_t)ix) * 4];
    for (int i = 0; i < 4; i++) t[i] = (uint8_t)col[i];
}

/* return_bilinear_col_all_precalculated, cl2.cl:1426-1455 */
static inline f4 bilinear_pre(float mx, float my, int which, int slice, float width,
                              const uint8_t* array) {
    float px = floorf(mx), py = floorf(my);
    f4 c0 = read_tex_pre(px, py, which, slice, width, array);
    f4 c1 = read_tex_pre(px + 1, py, which, slice, width, array);
    f4 c2 = read_tex_pre(px, py + 1, which, slice, width, array);
    f4 c3 = read_tex_pre(px + 1, py + 1, which, slice, width, array);
    float ux = mx - px, uy = my - py;
    float bx = 1.f - ux, by = 1.f - uy;
    f4 r;
    r.x = fmaf(c0.x, bx, c1.x * ux) * by + fmaf(c2.x, bx, c3.x * ux) * uy;
    r.y = fmaf(c0.y, bx, c1.y * ux) * by + fmaf(c2.y, bx, c3.y * ux) * uy;
    r.z = fmaf(c0.z, bx, c1.z * ux) * by + fmaf(c2.z, bx, c3.z * ux) * uy;
    r.w = fmaf(c0.w, bx, c1.w * ux) * by + fmaf(c2.w, bx, c3.w * ux) * uy;
    return r;
}

/* texture_mod, cl2.cl:1457-1468 */
static inline float tmod1(float v) {
    v = v >= 1 ? 1.0f - (v - floorf(v)) : v;
    v = v < 0 ? 1.0f + fabsf(v) - fabsf(floorf(v)) : v;
    return v;
}

/* log2_approx, cl2.cl:1498-1505 */
float oracle_log2_approx(float val) {
    int32_t x; memcpy(&x, &val, 4);
    float log_2 = (float)(((x >> 23) & 255) - 128);
    x &= ~(255 << 23);
    x += 127 << 23;
    float uv; memcpy(&uv, &x, 4);
    log_2 += ((-0.3358287811f) * uv + 2.0f) * uv - 0.65871759316667f;
    return log_2;
}

/* texture_filter_diff, cl2.cl:1511-1573 */
static f4 texture_filter_diff(float vtx, float vty, float dx, float dy, int tid2, uint32_t mip_start,
                              const uint32_t* nums, const uint32_t* sizes, const uint8_t* array) {
    int nv = (int)nums[tid2];
    int slice = nv >> 16;
    int tsize = (int)sizes[slice];
    float vmx = tmod1(vtx), vmy = tmod1(vty);
    float tx = dx * (float)tsize, ty = dy * (float)tsize;
    float worst = sqrtf(tx * tx + ty * ty);
    float wf = oracle_log2_approx(worst);
    wf = fmaxf(wf, 0.f);
    float mip_lower = floorf(wf);
    mip_lower = clampf_(mip_lower, 0.f, (float)MIPL);
    float fmd = wf - mip_lower;
    if (PLANT(10)) fmd = 1.f - fmd;
    int tid_lower = mip_lower == 0 ? tid2
                  : (int)(mip_lower - 1 + (float)mip_start + (float)(tid2 * MIPL));
    int tid_higher = (int)(clampf_(mip_lower, 0.f, MIPL - 1.f) + (float)mip_start + (float)(tid2 * MIPL));
    int lower_nv = (int)nums[tid_lower], higher_nv = (int)nums[tid_higher];
    int slice_lower = lower_nv >> 16, slice_higher = higher_nv >> 16;
    int which_lower = lower_nv & 0xFFFF, which_higher = higher_nv & 0xFFFF;
    float size_lower = (float)sizes[slice_lower], size_higher = (float)sizes[slice_higher];
    f4 c1 = bilinear_pre(vmx * size_lower, vmy * size_lower, which_lower, slice_lower, size_lower, array);
    f4 c2 = bilinear_pre(vmx * size_higher, vmy * size_higher, which_higher, slice_higher, size_higher, array);
    /* mix(a, b, t) = a + (b - a) * t */
    f4 f = {c1.x + (c2.x - c1.x) * fmd, c1.y + (c2.y - c1.y) * fmd,
            c1.z + (c2.z - c1.z) * fmd, c1.w + (c2.w - c1.w) * fmd};
    const float i255 = 1.f / 255.f;
    f.x = f.x * i255; f.y = f.y * i255; f.z = f.z * i255; f.w = f.w * i255;
    return f;
}

/* ----------------------------------------------------- shading helpers */
uint32_t oracle_wang_hash(uint32_t seed) {     /* cl2.cl:1919-1927 */
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= PLANT(4) ? 0x27d4eb2fu : 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed;
}
uint32_t oracle_xorshift(uint32_t s) {          /* cl2.cl:1929-1936 */
    s ^= (s << 13);
    s ^= (s >> 17);
    s ^= (s << 5);
    return s;
}

/* rational_acos, cl2.cl:2469-2477 */
float oracle_rational_acos(float x) {
    float a = -0.939115566365855f;
    float b = 0.9217841528914573f;
    float c = -1.2845906244690837f;
    float d = 0.295624144969963174f;
    return M_PI_F / 2.f + (a * x + b * x * x * x) / (1.f + c * x * x + d * oracle_pow4(x));
}

/* gamma_transform_approx, cl2.cl:5763-5772 */
static inline float gtf(float c) {
    return 0.012522878f * c + 0.682171111f * c * c + 0.305306011f * c * c * c;
}
void oracle_gamma_transform(const float in[4], float out[4]) {
    out[0] = gtf(in[0]); out[1] = gtf(in[1]); out[2] = gtf(in[2]); out[3] = in[3];
}
/* gamma_inverse_approx, cl2.cl:5774-5782 */
static inline float gif(float c) {
    float s1 = sqrtf(c), s2 = sqrtf(s1), s3 = sqrtf(s2);
    return 0.585122381f * s1 + 0.783140355f * s2 - 0.368262736f * s3;
}
void oracle_gamma_inverse(const float in[4], float out[4]) {
    out[0] = gif(in[0]); out[1] = gif(in[1]); out[2] = gif(in[2]); out[3] = in[3];
}

/* ret_cubeface, cl2.cl:1745-1790 */
static inline int ret_cubeface(f3 point, f3 light) {
    f3 rel = sub3(point, light);
    f3 arel = V3(fabsf(rel.x), fabsf(rel.y), fabsf(rel.z));
    if (arel.x >= arel.y && arel.x >= arel.z) return (rel.x < 0) != PLANT(11) ? 4 : 5;
    if (arel.y > arel.x && arel.y >= arel.z) return rel.y < 0 ? 1 : 3;
    if (arel.z > arel.x && arel.z > arel.y) { if (rel.z < 0) return 2; }
    return 0;
}
int oracle_ret_cubeface(const float p[3], const float l[3]) { return ret_cubeface(ld3(p), ld3(l)); }

/* get_barycentric, cl2.cl:5372-5384 */
static inline void get_barycentric(f3 p, f3 a, f3 b, f3 c, float* u, float* v, float* w) {
    f3 v0 = sub3(b, a), v1 = sub3(c, a), v2 = sub3(p, a);
    float d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1);
    float d20 = dot3(v2, v0), d21 = dot3(v2, v1);
    float denom = d00 * d11 - d01 * d01;
    *v = (d11 * d20 - d01 * d21) / denom;
    *w = (d00 * d21 - d01 * d20) / denom;
    *u = 1.0f - *v - *w;
}

/* encode_normal / float_to_short, cl2.cl:5592-5630; POW2e16-1 expands
 * without parentheses: ((r + 1) / 2) * 65536 - 1, converted with saturation. */
static inline uint16_t f2short(float r) {
    float f = ((r + 1) / 2) * 65536 - 1;
    if (!(f > 0.f)) return 0;
    if (f >= 65535.f) return 65535;
    return (uint16_t)f;
}
static inline uint32_t encode_normal(f3 val) {
    float len_sq = val.x * val.x + val.y * val.y;
    if (len_sq < 0.0001f) val.x = 0.01f;
    float l2 = val.x * val.x + val.y * val.y;
    float inv = 1.0f / sqrtf(l2);
    float sq = sqrtf(fmaxf(val.z * 0.5f + 0.5f, 0.f));
    float rx = (val.x * inv) * sq, ry = (val.y * inv) * sq;
    return (uint32_t)f2short(rx) | ((uint32_t)f2short(ry) << 16);
}

/* generate_ssao, cl2.cl:2194-2260 */
static float generate_ssao(int sx, int sy, const uint32_t* depth_buffer, uint32_t W, uint32_t H,
                           float fov) {
    uint32_t seed1 = oracle_wang_hash((uint32_t)sx + W * H * (uint32_t)sy);
    uint32_t seed2 = oracle_xorshift(seed1);
    float foffset = (float)seed2 / MULINT_F;
    float depth = ((float)depth_buffer[(size_t)sy * W + (size_t)sx] / MULINT_F) * DEPTH_FAR;
    float rad = SSAO_RAD_F;
    rad += foffset / 2.f;
    float world_rad = rad * fov / depth;
    const int samples = 2;
    float fsx = (float)sx, fsy = (float)sy;
    float acc = 0.f;
    for (int y = -samples; y <= samples; y++) {
        for (int x = -samples; x <= samples; x++) {
            float ox = roundf((float)x * world_rad), oy = roundf((float)y * world_rad);
            float wx = clampf_(fsx + ox, 1.f, (float)W - 2.f);
            float wy = clampf_(fsy + oy, 1.f, (float)H - 2.f);
            float d2 = ((float)depth_buffer[(size_t)(int)wy * W + (size_t)(int)wx] / MULINT_F) * DEPTH_FAR;
            for (float z = (float)-samples; z <= (float)samples; z += 1.f)
                if (d2 > depth + z) acc += 1.f;
        }
    }
    acc /= 125.f;       /* pow(samples*2.f + 1.f, 3.f) */
    return 1.f - (1.f - acc) / (PLANT(8) ? 2.0f : SSAO_DIV_F);
}

/* bilinear_interpolate, cl2.cl:2286-2302 */
static inline float bilinear_interpolate(float cx, float cy, const float v[4]) {
    float mx = cx - 0.5f, my = cy - 0.5f;
    float ux = mx - floorf(mx), uy = my - floorf(my);
    float bx = 1.0f - ux, by = 1.0f - uy;
    return (v[0] * bx + v[1] * ux) * by + (v[2] * bx + v[3] * ux) * uy;
}

/* generate_hard_occlusion, cl2.cl:2536-2701 (SMOOTH_SHADOWS variant) */
static float generate_hard_occlusion(f3 lpos, f3 normal, f3 ptl, const uint32_t* light_depth_buffer,
                                     int face, f3 global_position, int shnum, uint32_t L,
                                     const oracle_sincos face_sc[6]) {
    ptl = fnorm3(ptl);
    f3 local_pos = rot(global_position, lpos, &face_sc[face]);
    float lf = (float)L;
    float lfov = PLANT(12) ? lf : lf / 2.0f;      /* LFOV_CONST, cl2.cl:11 */
    f3 post = depth_project1(local_pos, lf, lf, lfov);
    float dpth = post.z;
    const uint32_t* map = &light_depth_buffer[((size_t)face + (size_t)shnum * 6) * L * L];
    float px = clampf_(post.x, 3.f, lf - 4.f);
    float py = clampf_(post.y, 3.f, lf - 4.f);
    int ix = (int)px, iy = (int)py;
    float acos_res = oracle_rational_acos(clampf_(dot3(normal, ptl), 0.05f, 0.95f));
    float bias = SHADOWBIAS_I * oracle_tan(acos_res);
    bias = clampf_(bias, 0.1f * SHADOWBIAS_I, (float)SHADOWBIAS_I);   /* pow(50, SHADOWEXP=1) */
    if (PLANT(3)) bias = -bias;
    float shadow = 0.f;
    int cond[16];
    for (int y = -1; y <= 2; y++)
        for (int x = -1; x <= 2; x++) {
            uint32_t raw = map[(size_t)(iy + y + PLANT(9)) * L + (size_t)(ix + x)];
            float ldp1 = ((float)raw / MULINT_F) * DEPTH_FAR;
            cond[(y + 1) * 4 + x + 1] = dpth > ldp1 + bias ? 1 : 0;
        }
    for (int y = -1; y <= 1; y++)
        for (int x = -1; x <= 1; x++) {
            float vals[4] = {(float)cond[(y + 1) * 4 + x + 1], (float)cond[(y + 1) * 4 + x + 2],
                             (float)cond[(y + 2) * 4 + x + 1], (float)cond[(y + 2) * 4 + x + 2]};
            shadow += bilinear_interpolate(px + 0.5f + (float)x, py + 0.5f + (float)y, vals);
        }
    shadow /= 9.f;
    return shadow;
}

/* get_vertex_col, cl2.cl:5670-5680 */
static inline f4 vertex_col4(uint32_t c) {
    f4 r = {(float)(c >> 24) / 255.f, (float)((c >> 16) & 0xFF) / 255.f,
            (float)((c >> 8) & 0xFF) / 255.f, (float)(c & 0xFF) / 255.f};
    return r;
}

static inline float qz(float v) {      /* RGBA8 quantisation of the float output */
    float c = clampf_(v, 0.f, 1.f);
    return floorf(c * 255.f + 0.5f);
}

/* kernel3, cl2.cl:5792-6408, one pixel. */
static void shade_pixel(const oracle_scene* sc, const ocr_view* v, const oracle_sincos* cam,
                        const oracle_sincos face_sc[6], const uint32_t* depth_buffer,
                        const uint32_t* ids, const prelist* pl, int x, int y, uint8_t* rgba,
                        uint16_t* nrm) {
    const uint32_t W = v->width, H = v->height;
    const float fov = v->fov_const;
    size_t pix = (size_t)y * W + (size_t)x;
    uint32_t ft = depth_buffer[pix];
    uint32_t idv = ids[pix];
    if (ft == 0xFFFFFFFFu || idv == 0) {
        for (int c = 0; c < 4; c++) rgba[c] = (uint8_t)qz(v->clear_col[c]);
        if (nrm) { nrm[0] = 0; nrm[1] = 0; }
        return;
    }
    const frag_rec* fr = &pl->frags[idv - 1];
    uint32_t tri_global = fr->tri;
    const cut_tri* ct = &pl->cuts[fr->cut];
    float rconst = fr->rconst;
    uint32_t o_id = fr->oid;
    f3 camera_pos = ld3(v->c_pos);
    const ocr_triangle* T = &sc->tris[tri_global];
    const ocr_obj_desc* G = &sc->objs[o_id];
    f3 p1 = mul3(ld3(T->vertices[0].pos), G->scale);
    f3 p2 = mul3(ld3(T->vertices[1].pos), G->scale);
    f3 p3 = mul3(ld3(T->vertices[2].pos), G->scale);
    float vt1x = T->vertices[0].vt[0], vt1y = T->vertices[0].vt[1];
    float vt2x = T->vertices[1].vt[0], vt2y = T->vertices[1].vt[1];
    float vt3x = T->vertices[2].vt[0], vt3y = T->vertices[2].vt[1];
    f3 n1 = ld3(T->vertices[0].normal), n2 = ld3(T->vertices[1].normal), n3 = ld3(T->vertices[2].normal);
    f3 wpos = ld3(G->world_pos);
    f4 q = {G->world_rot_quat[0], G->world_rot_quat[1], G->world_rot_quat[2], G->world_rot_quat[3]};

    float ldepth = ((float)ft / MULINT_F) * DEPTH_FAR;
    float actual_depth = ldepth;
    f3 local_position = V3(((float)x - (float)W / 2.0f) * actual_depth / fov,
                           ((float)y - (float)H / 2.0f) * actual_depth / fov, actual_depth);
    f3 global_position = add3(back_rot(local_position, cam), camera_pos);
    f3 object_local = back_rot_quat(sub3(global_position, wpos), q);
    float l1, l2, l3;
    get_barycentric(object_local, p1, p2, p3, &l1, &l2, &l3);
    float vtx = fmaf(vt1x, l1, fmaf(vt2x, l2, vt3x * l3));
    float vty = fmaf(vt1y, l1, fmaf(vt2y, l2, vt3y * l3));
    f3 normal = V3(fmaf(n1.x, l1, fmaf(n2.x, l2, n3.x * l3)),
                   fmaf(n1.y, l1, fmaf(n2.y, l2, n3.y * l3)),
                   fmaf(n1.z, l1, fmaf(n2.z, l2, n3.z * l3)));
    normal = rot_quat(normal, q);
    int has_colour_already = T->vertices[0].vertex_col != 0;
    f4 vertex_col = {0, 0, 0, 0};
    if (has_colour_already) {
        f4 a = vertex_col4(T->vertices[0].vertex_col), b = vertex_col4(T->vertices[1].vertex_col),
           c = vertex_col4(T->vertices[2].vertex_col);
        vertex_col.x = fmaf(a.x, l1, fmaf(b.x, l2, c.x * l3));
        vertex_col.y = fmaf(a.y, l1, fmaf(b.y, l2, c.y * l3));
        vertex_col.z = fmaf(a.z, l1, fmaf(b.z, l2, c.z * l3));
        vertex_col.w = fmaf(a.w, l1, fmaf(b.w, l2, c.w * l3));
    }
    const f3* tp = ct->p;

    /* get_vtdiff, cl2.cl:5691-5760 */
    float vdiffx, vdiffy;
    {
        float fx = (float)x, fy = (float)y;
        f3 xpv = V3(roundf(tp[0].x), roundf(tp[1].x), roundf(tp[2].x));
        f3 ypv = V3(roundf(tp[0].y), roundf(tp[1].y), roundf(tp[2].y));
        f3 depths = V3(1.0f / tp[0].z, 1.0f / tp[1].z, 1.0f / tp[2].z);
        float DA, DB, DC;
        interpolate_get_const(depths, xpv, ypv, rconst, &DA, &DB, &DC);
        float dmx = fmaf(DA, fx + 1, fmaf(DB, fy, DC));
        float dmy = fmaf(DA, fx, fmaf(DB, fy + 1, DC));
        f3 lmx = V3((fx + 1 - (float)W / 2.f) / fov, (fy - (float)H / 2.f) / fov, 1);
        f3 lmy = V3((fx - (float)W / 2.f) / fov, (fy + 1 - (float)H / 2.f) / fov, 1);
        lmx = div3(lmx, dmx);
        lmy = div3(lmy, dmy);
        f3 gmx = back_rot_quat(sub3(add3(back_rot(lmx, cam), camera_pos), wpos), q);
        f3 gmy = back_rot_quat(sub3(add3(back_rot(lmy, cam), camera_pos), wpos), q);
        float lx1, lx2, lx3, ly1, ly2, ly3;
        get_barycentric(gmx, p1, p2, p3, &lx1, &lx2, &lx3);
        get_barycentric(gmy, p1, p2, p3, &ly1, &ly2, &ly3);
        float vtxx = fmaf(vt1x, lx1, fmaf(vt2x, lx2, vt3x * lx3));
        float vtxy = fmaf(vt1y, lx1, fmaf(vt2y, lx2, vt3y * lx3));
        float vtyx = fmaf(vt1x, ly1, fmaf(vt2x, ly2, vt3x * ly3));
        float vtyy = fmaf(vt1y, ly1, fmaf(vt2y, ly2, vt3y * ly3));
        float vdxx = fabsf(vtxx - vtx), vdxy = fabsf(vtxy - vty);
        float vdyx = fabsf(vtyx - vtx), vdyy = fabsf(vtyy - vty);
        const float mip_bias = PLANT(2) ? MIP_BIAS_F : 1.f / MIP_BIAS_F;
        vdiffx = (vdxx + vdyx) * mip_bias;
        vdiffy = (vdxy + vdyy) * mip_bias;
    }

    f4 col;
    if (!has_colour_already)
        col = texture_filter_diff(vtx, vty, vdiffx, vdiffy, (int)G->tid, sc->mip_start, sc->nums,
                                  sc->sizes, sc->atlas);
    else
        col = vertex_col;
    if (v->use_linear_rendering) { col.x = gtf(col.x); col.y = gtf(col.y); col.z = gtf(col.z); }

    uint32_t seed1 = oracle_wang_hash((uint32_t)x + (uint32_t)y * W * H);
    uint32_t seed2 = oracle_xorshift(seed1);
    uint32_t seed3 = oracle_xorshift(seed2);
    uint32_t seed4 = oracle_xorshift(seed3);
    f3 rseed = V3((float)seed2 / MULINT_F, (float)seed3 / MULINT_F, (float)seed4 / MULINT_F);
    rseed = mul3(V3(rseed.x - 0.5f, rseed.y - 0.5f, rseed.z - 0.5f), 2);

    int shnum = 0, static_num = 0;
    f3 diffuse_sum = V3(0, 0, 0), specular_sum = V3(0, 0, 0);
    f3 l2p = fnorm3(sub3(camera_pos, global_position));
    int feature_flag = G->feature_flag;
    int is_two_sided = (feature_flag & OCR_FEATURE_TWO_SIDED) > 0;
    int receives_dynamic_shadows = !((feature_flag & OCR_FEATURE_DOES_NOT_RECEIVE_DYNAMIC_SHADOWS) > 0);
    int is_front = backface_keep(tp[0], tp[1], tp[2]);
    int flip_normals = !is_front && is_two_sided == 1;
    if (flip_normals) normal = neg3(normal);
    float ssao = generate_ssao(x, y, depth_buffer, W, H, fov);
    normal = fnorm3(normal);
    f3 lighting_normal = fnorm3(add3(normal, div3(rseed, 100.f)));
    float ambient = AMBIENT_F;
    if (v->use_linear_rendering) ambient = gtf(AMBIENT_F);

    for (uint32_t i = 0; i < sc->nlights; i++) {
        const ocr_light* l = &sc->lights[i];
        f3 lpos = ld3(l->pos);
        f3 point_to_light = sub3(lpos, global_position);
        float occlusion = 1;
        if (l->shadow && l->is_static) {
            int face = ret_cubeface(global_position, lpos);
            occlusion = 1.f - generate_hard_occlusion(lpos, normal, point_to_light, sc->static_maps,
                                                      face, global_position, static_num,
                                                      sc->lbuf_dim, face_sc);
            static_num++;
        }
        float distance = flen3(point_to_light);
        float t = (distance / l->radius) + 1;
        float illumination = l->brightness / (t * t);       /* pow(t, 2) */
        const float cutoff = 0.1f;
        illumination -= cutoff;
        illumination *= 1.f / (1.f - cutoff);
        if (illumination <= 0) continue;
        f3 light_col = ld3(l->col);
        if (v->use_linear_rendering) light_col = V3(gtf(light_col.x), gtf(light_col.y), gtf(light_col.z));
        if (l->shadow && receives_dynamic_shadows) {
            int face = ret_cubeface(global_position, lpos);
            float dyn = 1.f - generate_hard_occlusion(lpos, normal, point_to_light, sc->shadow_maps,
                                                      face, global_position, shnum, sc->lbuf_dim,
                                                      face_sc);
            occlusion = fminf(occlusion, dyn);
            shnum++;
        }
        point_to_light = fnorm3(point_to_light);
        float light = dot3(point_to_light, lighting_normal);
        light *= occlusion;
        light = fmaxf(light, 0.f);
        float diffuse = (1.0f - ambient) * light;
        diffuse_sum = add3(diffuse_sum, mul3(light_col, (diffuse + ambient) * l->diffuse * G->diffuse * illumination));
        f3 Hv = fnorm3(add3(l2p, point_to_light));
        f3 N = normal;
        const float kS = 0.4f;
        float ndh = fmaxf(0.f, dot3(N, Hv));
        float ndv = fmaxf(0.f, dot3(N, l2p));
        float vdh = fmaxf(0.f, dot3(l2p, Hv));
        float ndl = fmaxf(0.f, dot3(N, point_to_light));
        const float F0 = 0.4f;
        float fresnel = F0 + (1 - F0) * oracle_powr5(1.f - vdh);
        float rough = clampf_(1.f - G->specular, 0.001f, 10.f);
        const float gauss_constant = 0.8346f;
        float alpha = oracle_rational_acos(ndh);
        float microfacet = gauss_constant * oracle_exp(-alpha * alpha / (rough * rough));
        float sv = 2 * ndh / vdh;
        float c1 = sv * ndv;
        float c2 = sv * ndl;
        float geometric = min3f(1.f, c1, c2);
        float spec = (fresnel * microfacet * geometric) / (M_PI_F * ndv);
        specular_sum = add3(specular_sum, mul3(mul3(light_col, spec * kS * illumination), G->spec_mult));
        specular_sum = V3(fmaxf(specular_sum.x, 0.f), fmaxf(specular_sum.y, 0.f), fmaxf(specular_sum.z, 0.f));
        specular_sum = mul3(specular_sum, occlusion);
    }
    specular_sum = mul3(specular_sum, ssao);
    diffuse_sum = mul3(diffuse_sum, ssao);
    const float reflected_surface_colour = 0.7f;
    f3 colclamp = add3(add3(V3(col.x, col.y, col.z), V3(0, 0, 0)), mul3(specular_sum, reflected_surface_colour));
    f3 final_col = V3(fmaf(colclamp.x, diffuse_sum.x, specular_sum.x * (1.f - reflected_surface_colour)),
                      fmaf(colclamp.y, diffuse_sum.y, specular_sum.y * (1.f - reflected_surface_colour)),
                      fmaf(colclamp.z, diffuse_sum.z, specular_sum.z * (1.f - reflected_surface_colour)));
    if (v->use_linear_rendering) final_col = V3(gif(final_col.x), gif(final_col.y), gif(final_col.z));
    rgba[0] = (uint8_t)qz(final_col.x);
    rgba[1] = (uint8_t)qz(final_col.y);
    rgba[2] = (uint8_t)qz(final_col.z);
    rgba[3] = (uint8_t)qz(col.w);
    if (nrm) {
        uint32_t e = encode_normal(normal);
        nrm[0] = (uint16_t)(e & 0xFFFF); nrm[1] = (uint16_t)(e >> 16);
    }
}

/* r_struct (cl2.cl:2549-2574 and :4487-4511; the reference builds with
 * -cl-single-precision-constant so both tables hold the same binary32
 * angles), sin/cos evaluated in binary64 and rounded once. */
void oracle_face_sincos(oracle_sincos out[6]) {
    const float ang[6][3] = {{0, 0, 0}, {M_PI_F / 2.0f, 0, 0}, {0, M_PI_F, 0},
                             {3.0f * M_PI_F / 2.0f, 0, 0}, {0, 3.0f * M_PI_F / 2.0f, 0},
                             {0, M_PI_F / 2.0f, 0}};
    for (int f = 0; f < 6; f++)
        for (int a = 0; a < 3; a++) {
            const int g = PLANT(1) ? (f == 1 ? 3 : f == 3 ? 1 : f) : f;
            out[f].s[a] = (float)sin((double)ang[g][a]);
            out[f].c[a] = (float)cos((double)ang[g][a]);
        }
}

/* ------------------------------------------------------------ frame */
int oracle_render(const oracle_scene* sc, const ocr_view* v, const oracle_sincos* cam, int nthreads,
                  oracle_out* out) {
    if (!sc || !v || !cam || !out || !out->depth || v->width == 0 || v->height == 0) return -1;
    const uint32_t W = v->width, H = v->height;
    const size_t npix = (size_t)W * H;
    int nt = nthreads > 1 ? nthreads : 1;
    double t0 = now_ms();
    prelist pl;
    prearrange_all(sc, v, cam, nt, &pl);
    double t1 = now_ms();
    for (size_t i = 0; i < npix; i++) out->depth[i] = 0xFFFFFFFFu;
    uint64_t samples = 0;
    /* kernel1: one work item per fragment */
    #pragma omp parallel for schedule(dynamic, 256) num_threads(nt) reduction(+:samples)
    for (long f = 0; f < (long)pl.nfrag; f++) {
        const frag_rec* r = &pl.frags[f];
        k1ctx k = {out->depth, W, 0, nt > 1};
        walk_fragment(pl.cuts[r->cut].p, (float)W, (float)H, r->chunk, OP_SIZE, r->rconst, k1_visit, &k);
        samples += k.samples;
    }
    double t2 = now_ms();
    uint32_t* ids = (uint32_t*)calloc(npix, sizeof(uint32_t));
    #pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
    for (long f = 0; f < (long)pl.nfrag; f++) {
        const frag_rec* r = &pl.frags[f];
        k2ctx k = {out->depth, ids, W, (uint32_t)f, nt > 1};
        walk_fragment(pl.cuts[r->cut].p, (float)W, (float)H, r->chunk, OP_SIZE, r->rconst, k2_visit, &k);
    }
    double t3 = now_ms();
    if (out->tri_id)
        for (size_t i = 0; i < npix; i++)
            out->tri_id[i] = (ids[i] && out->depth[i] != 0xFFFFFFFFu) ? pl.frags[ids[i] - 1].tri : OCR_NO_TRIANGLE;
    /* distinct visible triangles / objects (for the algorithmic byte count) */
    {
        uint8_t* seen = (uint8_t*)calloc(sc->ntri ? sc->ntri : 1, 1);
        uint8_t* seeno = (uint8_t*)calloc(sc->nobj ? sc->nobj : 1, 1);
        uint32_t nv = 0, no = 0;
        for (size_t i = 0; i < npix; i++) {
            if (!ids[i] || out->depth[i] == 0xFFFFFFFFu) continue;
            const frag_rec* r = &pl.frags[ids[i] - 1];
            if (!seen[r->tri]) { seen[r->tri] = 1; nv++; }
            if (r->oid < sc->nobj && !seeno[r->oid]) { seeno[r->oid] = 1; no++; }
        }
        out->n_visible_tris = nv; out->n_visible_objs = no;
        free(seen); free(seeno);
    }
    oracle_sincos face_sc[6];
    oracle_face_sincos(face_sc);
    if (out->rgba8) {
        #pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
        for (long yy = 0; yy < (long)H; yy++)
            for (uint32_t xx = 0; xx < W; xx++) {
                size_t p = (size_t)yy * W + xx;
                shade_pixel(sc, v, cam, face_sc, out->depth, ids, &pl, (int)xx, (int)yy,
                            &out->rgba8[p * 4], out->normals ? &out->normals[p * 2] : NULL);
            }
    }
    double t4 = now_ms();
    out->n_cut = pl.nsurvive;
    out->n_frag = pl.nfrag;
    out->n_samples = samples;
    out->ms_prearrange = t1 - t0; out->ms_depth = t2 - t1; out->ms_id = t3 - t2; out->ms_shade = t4 - t3;
    free(ids); free(pl.cuts); free(pl.frags);
    return 0;
}

/* --------------------------------------------------------- shadows */
/* prearrange_realtime_shadowing (cl2.cl:4419-4636) + kernel1_realtime_shadowing
 * (cl2.cl:5129-5246) for one light into one 6-face map. */
typedef struct { uint32_t* map; uint32_t L; uint32_t face; int atomic; uint64_t samples; } kshctx;
static void ksh_visit(void* c, int x, int y, uint32_t d) {
    kshctx* k = (kshctx*)c;
    k->samples++;
    uint32_t* ft = &k->map[(size_t)y * k->L + (size_t)x + (size_t)k->face * k->L * k->L];
    if (k->atomic) {
        uint32_t old = __atomic_load_n(ft, __ATOMIC_RELAXED);
        while (d < old && !__atomic_compare_exchange_n(ft, &old, d, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
    } else if (d < *ft) {
        *ft = d;
    }
}

static void render_light_map(const oracle_scene* sc, f3 lpos, int only_static,
                             const oracle_sincos face_sc[6], int nt, uint32_t* map, uint64_t* frags,
                             uint64_t* samples) {
    const uint32_t L = sc->lbuf_dim;
    const float efov = PLANT(5) ? (float)L : (float)L / 2.0f, ew = (float)L, eh = (float)L;
    const int op_l = PLANT(6) ? 500 : OP_SIZE_L;
    #pragma omp parallel for schedule(dynamic, 1024) num_threads(nt)
    for (long id = 0; id < (long)sc->ntri; id++) {
        const ocr_triangle* T = &sc->tris[id];
        uint32_t o_id = T->vertices[0].object_id;
        if (o_id >= sc->nobj) continue;
        const ocr_obj_desc* G = &sc->objs[o_id];
        int ff = G->feature_flag;
        int is_static = (ff & OCR_FEATURE_IS_STATIC) > 0;
        if (!only_static && is_static) continue;
        if (only_static && !is_static) continue;
        f3 wp = ld3(G->world_pos);
        if (flen3(sub3(wp, lpos)) > DEPTH_FAR) continue;
        f4 q = {G->world_rot_quat[0], G->world_rot_quat[1], G->world_rot_quat[2], G->world_rot_quat[3]};
        int skip[6] = {0, 0, 0, 0, 0, 0};
        for (int kk = 0; kk < 3; kk++) {
            f3 r = add3(rot_quat(mul3(ld3(T->vertices[kk].pos), G->scale), q), wp);
            skip[ret_cubeface(r, lpos)] = 1;
        }
        if (PLANT(7)) for (int kk = 0; kk < 6; kk++) skip[kk] = 1;
        int two_sided = (ff & OCR_FEATURE_TWO_SIDED) > 0;
        for (int kk = 0; kk < 6; kk++) {
            if (!skip[kk]) continue;
            f3 tp[2][3];
            int num = full_rotate_quat(ld3(T->vertices[0].pos), ld3(T->vertices[1].pos),
                                       ld3(T->vertices[2].pos), tp, lpos, &face_sc[kk], wp, q,
                                       G->scale, efov, ew, eh);
            for (int i = 0; i < num; i++) {
                f3* p = tp[i];
                int valid = two_sided || backface_keep(p[0], p[1], p[2]);
                int cond = (p[0].x < 0 && p[1].x < 0 && p[2].x < 0) ||
                           (p[0].x >= ew && p[1].x >= ew && p[2].x >= ew) ||
                           (p[0].y < 0 && p[1].y < 0 && p[2].y < 0) ||
                           (p[0].y >= eh && p[1].y >= eh && p[2].y >= eh);
                if (!valid || cond) continue;
                f3 xpv = V3(roundf(p[0].x), roundf(p[1].x), roundf(p[2].x));
                f3 ypv = V3(roundf(p[0].y), roundf(p[1].y), roundf(p[2].y));
                float rconst = calc_rconstant_v(xpv, ypv);
                float mm[4];
                calc_min_max(p, ew, eh, mm);
                float area = (mm[1] - mm[0]) * (mm[3] - mm[2]);
                float thread_num = ceilf(area / (float)op_l);
                kshctx k = {map, L, (uint32_t)kk, nt > 1, 0};
                if (frags) {
                    #pragma omp atomic
                    frags[kk] += (uint64_t)thread_num;
                }
                for (uint32_t a = 0; (float)a < thread_num; a++)
                    walk_fragment(p, ew, eh, a, op_l, rconst, ksh_visit, &k);
                if (samples) {
                    #pragma omp atomic
                    *samples += k.samples;
                }
            }
        }
    }
}

int oracle_shadow_maps(const oracle_scene* sc, int nthreads, uint32_t* dyn_maps,
                       uint32_t* static_maps) {
    return oracle_shadow_maps_ex(sc, nthreads, dyn_maps, static_maps, NULL);
}

int oracle_shadow_maps_ex(const oracle_scene* sc, int nthreads, uint32_t* dyn_maps,
                          uint32_t* static_maps, uint64_t* samples) {
    if (samples) *samples = 0;
    oracle_sincos face_sc[6];
    oracle_face_sincos(face_sc);
    const size_t per = (size_t)6 * sc->lbuf_dim * sc->lbuf_dim;
    int nt = nthreads > 1 ? nthreads : 1;
    uint32_t nn = 0, kk = 0;
    for (uint32_t i = 0; i < sc->nlights; i++) {
        const ocr_light* l = &sc->lights[i];
        f3 lpos = ld3(l->pos);
        if (l->shadow == 1) {
            uint32_t* m = &dyn_maps[per * nn];
            for (size_t j = 0; j < per; j++) m[j] = 0xFFFFFFFFu;
            render_light_map(sc, lpos, 0, face_sc, nt, m, NULL, samples);
            nn++;
        }
        if (l->shadow && l->is_static) {
            uint32_t* m = &static_maps[per * kk];
            for (size_t j = 0; j < per; j++) m[j] = 0xFFFFFFFFu;
            render_light_map(sc, lpos, 1, face_sc, nt, m, NULL, samples);
            kk++;
        }
    }
    return 0;
}

/* ----------------------------------------------------------- atlas */
static uint32_t pow2_up(uint32_t larg) {          /* texture.cpp:46-63 */
    uint32_t n = 1;
    while (n < larg) n <<= 1;
    return n;
}

/* texture_context.cpp:94-262: pages per size class, slot descriptors. */
int oracle_atlas_plan(const uint32_t* widths, const uint32_t* heights, uint32_t ntex,
                      uint32_t* npages, uint32_t* nums, uint32_t* sizes, uint32_t* mip_start) {
    /* size -> count, ascending sizes (std::map order) */
    uint32_t sz[64]; int cnt[64]; int nsz = 0;
    for (uint32_t i = 0; i < ntex; i++) {
        uint32_t base = pow2_up(widths[i] > heights[i] ? widths[i] : heights[i]);
        if (base < 16 || base > 2048) return -1;
        for (int j = 0; j <= MIPL; j++) {
            uint32_t s = j == 0 ? base : base / (1u << j);
            int k = 0;
            for (; k < nsz; k++) if (sz[k] == s) break;
            if (k == nsz) { sz[nsz] = s; cnt[nsz] = 0; nsz++; }
            cnt[k]++;
        }
    }
    for (int a = 0; a < nsz; a++)         /* sort ascending */
        for (int b = a + 1; b < nsz; b++)
            if (sz[b] < sz[a]) { uint32_t t = sz[a]; sz[a] = sz[b]; sz[b] = t; int c = cnt[a]; cnt[a] = cnt[b]; cnt[b] = c; }
    uint32_t psize[4096]; int pnum[4096]; int np = 0;
    for (int k = 0; k < nsz; k++) {
        int per = (int)((2048 / sz[k]) * (2048 / sz[k]));
        int rem = cnt[k];
        while (rem >= per) { psize[np] = sz[k]; pnum[np] = per; np++; rem -= per; }
        if (rem > 0) { psize[np] = sz[k]; pnum[np] = rem; np++; }
    }
    for (uint32_t i = 0; i < ntex * 5; i++) nums[i] = 0;
    for (uint32_t i = 0; i < ntex; i++) {
        uint32_t base = pow2_up(widths[i] > heights[i] ? widths[i] : heights[i]);
        int slice = 0, free_num = -1;
        for (; slice < np; slice++)
            if (psize[slice] == base && pnum[slice] > 0) { free_num = pnum[slice] - 1; pnum[slice]--; break; }
        if (free_num < 0) return -2;
        nums[i] = ((uint32_t)slice << 16) | (uint32_t)free_num;
    }
    for (uint32_t i = 0; i < ntex; i++) {
        uint32_t base = pow2_up(widths[i] > heights[i] ? widths[i] : heights[i]);
        for (int j = 0; j < MIPL; j++) {
            uint32_t ms = base / (1u << (j + 1));
            int slice = 0, free_num = -1;
            for (; slice < np; slice++)
                if (psize[slice] == ms && pnum[slice] > 0) { free_num = pnum[slice] - 1; pnum[slice]--; break; }
            if (free_num < 0) return -3;
            nums[ntex + i * MIPL + j] = ((uint32_t)slice << 16) | (uint32_t)free_num;
        }
    }
    for (int s = 0; s < np; s++) sizes[s] = psize[s];
    *npages = (uint32_t)np;
    *mip_start = ntex;
    return 0;
}

/* generate_mips (cl2.cl:1070-1129) for level 0 -> 1 when src_id is a base
 * texture, generate_mip_mips (cl2.cl:1131-1189) for level L -> L+1. */
static void mip_pass(uint32_t src_id, uint32_t dst_id, const uint32_t* nums, const uint32_t* sizes,
                     uint32_t gw, uint32_t gh, uint8_t* pages) {
    static const float gauss[3][3] = {{1, 2, 1}, {2, 4, 2}, {1, 2, 1}};
    int slice = (int)(nums[src_id] >> 16);
    float width = (float)sizes[slice];
    int w2 = (int)(nums[dst_id] >> 16);
    float nwidth = (float)sizes[w2];
    /* The launch covers the source image's own dimensions gw x gh
     * (texture.cpp:465-493) and each work item returns when x or y >= width.
     * Every work item reads level L and writes level L+1 of the same pages;
     * the two regions never overlap, so a sequential sweep is exact. */
    for (int y = 0; y < (int)gh; y++)
        for (int x = 0; x < (int)gw; x++) {
            if ((float)x >= width || (float)y >= width) continue;
            float ax = 0, ay = 0, az = 0, aw = 0, div = 0.f;
            for (int j = -1; j <= 1; j++)
                for (int i = -1; i <= 1; i++) {
                    f4 col = read_tex_array((float)(x * 2 + i), (float)(y * 2 + j), src_id, nums, sizes, pages);
                    col.w /= 255.f;
                    col.x *= col.w; col.y *= col.w; col.z *= col.w;
                    float g = gauss[j + 1][i + 1];
                    ax += col.x * g; ay += col.y * g; az += col.z * g; aw += col.w * g;
                    div += g;
                }
            ax /= div; ay /= div; az /= div; aw /= div;
            if (aw > 0.00000001f) { ax /= aw; ay /= aw; az /= aw; }
            aw *= 255.f;
            float yx = ((float)(x * 2) / width) * nwidth;
            float yy = ((float)(y * 2) / width) * nwidth;
            if (yx >= nwidth || yy >= nwidth) continue;
            uint32_t c[4] = {sat_u32(ax), sat_u32(ay), sat_u32(az), sat_u32(aw)};
            write_tex_array(c, yx, yy, dst_id, nums, sizes, pages);
        }
}

int oracle_atlas_fill(const ocr_tex_src* texs, uint32_t ntex, const uint32_t* nums,
                      const uint32_t* sizes, uint32_t mip_start, uint8_t* pages) {
    for (uint32_t t = 0; t < ntex; t++) {
        /* update_gpu_tex with flip = 1 (cl2.cl:923-953, texture.cpp:340-350):
         * UNORM8 -> float -> *255 -> convert_uint4 */
        int slice = (int)(nums[t] >> 16);
        float width = (float)sizes[slice];
        for (uint32_t y = 0; y < texs[t].height; y++)
            for (uint32_t x = 0; x < texs[t].width; x++) {
                const uint8_t* s = &texs[t].rgba[((size_t)y * texs[t].width + x) * 4];
                uint32_t c[4];
                for (int k = 0; k < 4; k++) c[k] = sat_u32(((float)s[k] / 255.0f) * 255.f);
                float fy = width - (float)(int)y;
                write_tex_array(c, (float)x, fy, t, nums, sizes, pages);
            }
        mip_pass(t, t * MIPL + mip_start, nums, sizes, texs[t].width, texs[t].height, pages);
        for (uint32_t lvl = 0; lvl < MIPL - 1; lvl++) {
            uint32_t proper = t * MIPL + mip_start + lvl;
            mip_pass(proper, proper + 1, nums, sizes, texs[t].width, texs[t].height, pages);
        }
    }
    return 0;
}

/* ---------------------------------------------------- test exports */
int oracle_cut_list(const oracle_scene* sc, const ocr_view* v, const oracle_sincos* cam,
                    float* cuts, uint32_t* keys, int max_out) {
    prelist pl;
    prearrange_all(sc, v, cam, 1, &pl);
    /* survivors (>= 1 fragment) in fragment order, key = 2*tri + sub */
    int n = 0;
    uint32_t last_cut = 0xFFFFFFFFu;
    for (uint32_t f = 0; f < pl.nfrag; f++) {
        const frag_rec* r = &pl.frags[f];
        if (r->cut == last_cut) continue;
        last_cut = r->cut;
        if (n < max_out) {
            for (int j = 0; j < 3; j++) {
                cuts[n * 9 + j * 3 + 0] = pl.cuts[r->cut].p[j].x;
                cuts[n * 9 + j * 3 + 1] = pl.cuts[r->cut].p[j].y;
                cuts[n * 9 + j * 3 + 2] = pl.cuts[r->cut].p[j].z;
            }
            keys[n] = 2u * r->tri + r->sub;
        }
        n++;
    }
    free(pl.cuts); free(pl.frags);
    return n;
}

void oracle_rot(const float p[3], const float cpos[3], const oracle_sincos* sc, float out[3]) {
    f3 r = rot(ld3(p), ld3(cpos), sc); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void oracle_back_rot(const float p[3], const oracle_sincos* sc, float out[3]) {
    f3 r = back_rot(ld3(p), sc); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void oracle_rot_quat(const float p[3], const float q[4], float out[3]) {
    f4 qq = {q[0], q[1], q[2], q[3]};
    f3 r = rot_quat(ld3(p), qq); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void oracle_back_rot_quat(const float p[3], const float q[4], float out[3]) {
    f4 qq = {q[0], q[1], q[2], q[3]};
    f3 r = back_rot_quat(ld3(p), qq); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
int oracle_clip(const float pts[3][3], float out[2][3][3]) {
    f3 p[3] = {ld3(pts[0]), ld3(pts[1]), ld3(pts[2])};
    f3 r[2][3];
    int n = generate_new_triangles(p, r);
    for (int t = 0; t < n; t++)
        for (int i = 0; i < 3; i++) { out[t][i][0] = r[t][i].x; out[t][i][1] = r[t][i].y; out[t][i][2] = r[t][i].z; }
    return n;
}
void oracle_calc_min_max(const float pts[3][3], float w, float h, float mm[4]) {
    f3 p[3] = {ld3(pts[0]), ld3(pts[1]), ld3(pts[2])};
    calc_min_max(p, w, h, mm);
}
int oracle_point_in_tri(float x, float y, const float xs[3], const float ys[3]) {
    return point_in_tri(x, y, xs[0], ys[0], xs[1], ys[1], xs[2], ys[2]);
}
typedef struct { int32_t* xs; int32_t* ys; uint32_t* ds; int n, max; } walkcap;
static void cap_visit(void* c, int x, int y, uint32_t d) {
    walkcap* w = (walkcap*)c;
    if (w->n < w->max) { w->xs[w->n] = x; w->ys[w->n] = y; w->ds[w->n] = d; }
    w->n++;
}
int oracle_walk(const float cut[3][3], float w, float h, uint32_t chunk, int op_size,
                int32_t* xs, int32_t* ys, uint32_t* ds, int max_out) {
    f3 p[3] = {ld3(cut[0]), ld3(cut[1]), ld3(cut[2])};
    f3 xpv = V3(roundf(p[0].x), roundf(p[1].x), roundf(p[2].x));
    f3 ypv = V3(roundf(p[0].y), roundf(p[1].y), roundf(p[2].y));
    float rc = calc_rconstant_v(xpv, ypv);
    walkcap c = {xs, ys, ds, 0, max_out};
    walk_fragment(p, w, h, chunk, op_size, rc, cap_visit, &c);
    return c.n;
}
void oracle_get_barycentric(const float p[3], const float a[3], const float b[3], const float c[3],
                            float* u, float* v, float* w) {
    get_barycentric(ld3(p), ld3(a), ld3(b), ld3(c), u, v, w);
}
uint32_t oracle_encode_normal(const float n[3]) { return encode_normal(ld3(n)); }

float oracle_ssao(int x, int y, const uint32_t* depth, uint32_t W, uint32_t H, float fov) {
    return generate_ssao(x, y, depth, W, H, fov);
}

/* The shadow term of one light as kernel3 forms it (cl2.cl:6115-6165): the
 * cube face of the point, then generate_hard_occlusion on map `shnum`. */
float oracle_shadow_term(const float gpos[3], const float normal[3], const float lpos[3],
                         const uint32_t* maps, uint32_t shnum, uint32_t L) {
    oracle_sincos face_sc[6];
    oracle_face_sincos(face_sc);
    f3 g = ld3(gpos), lp = ld3(lpos);
    int face = ret_cubeface(g, lp);
    return generate_hard_occlusion(lp, ld3(normal), sub3(lp, g), maps, face, g, (int)shnum, L, face_sc);
}

void oracle_texture_filter(float vtx, float vty, float dx, float dy, int tid, uint32_t mip_start,
                           const uint32_t* nums, const uint32_t* sizes, const uint8_t* atlas, float out[4]) {
    f4 c = texture_filter_diff(vtx, vty, dx, dy, tid, mip_start, nums, sizes, atlas);
    out[0] = c.x; out[1] = c.y; out[2] = c.z; out[3] = c.w;
}

int oracle_shadow_map_one(const oracle_scene* sc, uint32_t light, int only_static, uint32_t* map,
                          uint64_t frags[6]) {
    if (light >= sc->nlights) return -1;
    oracle_sincos face_sc[6];
    oracle_face_sincos(face_sc);
    const size_t per = (size_t)6 * sc->lbuf_dim * sc->lbuf_dim;
    for (size_t j = 0; j < per; j++) map[j] = 0xFFFFFFFFu;
    if (frags) for (int k = 0; k < 6; k++) frags[k] = 0;
    render_light_map(sc, ld3(sc->lights[light].pos), only_static, face_sc, 1, map, frags, NULL);
    return 0;
}
void oracle_texture_mod(const float vt[2], float out[2]) { out[0] = tmod1(vt[0]); out[1] = tmod1(vt[1]); }

/* blend_screens_with_depth, cl2.cl:7842-7875: the pixel is left alone when
 * the destination is nearer (d2 < d1), else alpha_blend (cl2.cl:1721-1728):
 * rgb = mad(src.rgb, src.a, dst.rgb * (1 - src.a)), a = src.a.  The images
 * are RGBA8 (read_imagef gives c / 255); the result is quantised like the
 * shading pass's output. */
void oracle_blend_with_depth(uint32_t npix, const uint32_t* src_depth, const uint32_t* dst_depth,
                             const uint8_t* src_rgba, uint8_t* dst_rgba) {
    for (uint32_t p = 0; p < npix; p++) {
        const uint32_t d1 = src_depth[p], d2 = dst_depth[p];
        if (d2 < d1) continue;
        float s[4], d[4];
        for (int c = 0; c < 4; c++) {
            s[c] = (float)src_rgba[4 * (size_t)p + c] / 255.f;
            d[c] = (float)dst_rgba[4 * (size_t)p + c] / 255.f;
        }
        const float inv = 1.f - s[3];
        for (int c = 0; c < 3; c++) dst_rgba[4 * (size_t)p + c] = (uint8_t)qz(fmaf(s[c], s[3], d[c] * inv));
        dst_rgba[4 * (size_t)p + 3] = (uint8_t)qz(s[3]);
    }
}

/* decode_normal(ushort2), cl2.cl:5611-5646: short_to_float is
 * (s / 65535) * 2 - 1 (the POW2e16-1 macro binds as 65536 - 1). */
static f3 decode_normal_u16(uint16_t sx, uint16_t sy) {
    float vx = (float)sx / 65535.f * 2.f - 1.f;
    float vy = (float)sy / 65535.f * 2.f - 1.f;
    float z = (vx * vx + vy * vy) * 2.f - 1.f;
    float r = 1.0f / sqrtf(vx * vx + vy * vy);
    float k = sqrtf(fmaxf(1.f - z * z, 0.f));
    return V3((vx * r) * k, (vy * r) * k, z);
}

/* do_pseudo_aa, cl2.cl:6438-6657.  Per interior pixel: the eight
 * neighbours are split by two tests, depth difference > 100 and a normal
 * turned by more than 20 degrees; for the first test (then the second) with
 * exactly one differing x neighbour, one differing y neighbour and at least
 * one differing corner, the pixel becomes 0.65 * mean(same side) + 0.35 *
 * mean(other side), alpha 1.  The reference's avg_depth / avg_normal feed
 * only the disabled REDUCED_AA exit and are not formed. */
void oracle_pseudo_aa(uint32_t W, uint32_t H, const uint32_t* depth, const uint16_t* normals, uint8_t* rgba) {
    const size_t npix = (size_t)W * H;
    uint8_t* in = (uint8_t*)malloc(npix * 4);
    if (!in) return;
    memcpy(in, rgba, npix * 4);
    /* cos(20 * 2 * M_PI / 360) in single precision: the argument rounds to
     * 0.34906587f and its cosine to 0.9396926f (0x1.e11f64p-1) */
    const float cosrad = (float)cos((double)(20.f * 2.f * 3.14159265f / 360.f));
    for (uint32_t y = 1; y + 1 < H; y++)
        for (uint32_t x = 1; x + 1 < W; x++) {
            const size_t p = (size_t)y * W + x;
            const uint32_t my_raw = depth[p];
            if ((float)my_raw == MULINT_F) continue;
            f3 mn = decode_normal_u16(normals[2 * p], normals[2 * p + 1]);
            mn = fnorm3(mn);
            const float my_depth = ((float)my_raw / MULINT_F) * DEPTH_FAR;
            int nx[2] = {0, 0}, ny[2] = {0, 0}, nc[2] = {0, 0};
            f3 mine[2] = {{0, 0, 0}, {0, 0, 0}}, theirs[2] = {{0, 0, 0}, {0, 0, 0}};
            float ms[2] = {0, 0}, ts[2] = {0, 0};
            for (int j = -1; j < 2; j++)
                for (int i = -1; i < 2; i++) {
                    if (i == 0 && j == 0) continue;
                    const size_t q = (size_t)(y + j) * W + (x + i);
                    const float d = ((float)depth[q] / MULINT_F) * DEPTH_FAR;
                    const f3 fn = decode_normal_u16(normals[2 * q], normals[2 * q + 1]);
                    const f3 val = V3((float)in[4 * q] / 255.f, (float)in[4 * q + 1] / 255.f,
                                      (float)in[4 * q + 2] / 255.f);
                    const int tests[2] = {fabsf(d - my_depth) > 100.f, dot3(mn, fn) < cosrad};
                    for (int k = 0; k < 2; k++) {
                        if (tests[k]) {
                            if (i == j || i == -j) nc[k]++;
                            else if (i == 1 || i == -1) nx[k]++;
                            else if (j == 1 || j == -1) ny[k]++;
                            theirs[k] = add3(theirs[k], val);
                            ts[k] += 1.f;
                        } else {
                            mine[k] = add3(mine[k], val);
                            ms[k] += 1.f;
                        }
                    }
                }
            for (int k = 0; k < 2; k++) {
                if (nx[k] == 1 && ny[k] == 1 && nc[k] >= 1) {
                    const f3 a = div3(mine[k], ms[k]), b = div3(theirs[k], ts[k]);
                    const f3 acc = add3(mul3(a, 0.65f), mul3(b, 0.35f));
                    rgba[4 * p] = (uint8_t)qz(acc.x);
                    rgba[4 * p + 1] = (uint8_t)qz(acc.y);
                    rgba[4 * p + 2] = (uint8_t)qz(acc.z);
                    rgba[4 * p + 3] = 255;
                    break;
                }
            }
        }
    free(in);
}

/* read_imagef with CLK_FILTER_LINEAR on unnormalised coordinates (u, v),
 * OpenCL 1.2 section 8.2: i0 = floor(u - 0.5), a = frac(u - 0.5) (and the
 * same for v), the four texels weighted (1-a)(1-b), a(1-b), (1-a)b, ab and
 * summed in that order; RGBA8 texels read as c / 255; indices beyond the
 * image edge read the edge. */
static void sample_linear(const uint8_t* img, uint32_t W, uint32_t H, float u, float v, float out[4]) {
    const float tx = u - 0.5f, ty = v - 0.5f;
    const float fx = floorf(tx), fy = floorf(ty);
    const float a = tx - fx, b = ty - fy;
    int i0 = (int)fx, j0 = (int)fy, i1 = i0 + 1, j1 = j0 + 1;
    i0 = i0 < 0 ? 0 : (i0 > (int)W - 1 ? (int)W - 1 : i0);
    i1 = i1 < 0 ? 0 : (i1 > (int)W - 1 ? (int)W - 1 : i1);
    j0 = j0 < 0 ? 0 : (j0 > (int)H - 1 ? (int)H - 1 : j0);
    j1 = j1 < 0 ? 0 : (j1 > (int)H - 1 ? (int)H - 1 : j1);
    const uint8_t* t00 = img + 4 * ((size_t)j0 * W + i0);
    const uint8_t* t10 = img + 4 * ((size_t)j0 * W + i1);
    const uint8_t* t01 = img + 4 * ((size_t)j1 * W + i0);
    const uint8_t* t11 = img + 4 * ((size_t)j1 * W + i1);
    const float w00 = (1.f - a) * (1.f - b), w10 = a * (1.f - b), w01 = (1.f - a) * b, w11 = a * b;
    for (int c = 0; c < 4; c++)
        out[c] = ((w00 * ((float)t00[c] / 255.f) + w10 * ((float)t10[c] / 255.f)) + w01 * ((float)t01[c] / 255.f))
               + w11 * ((float)t11[c] / 255.f);
}

void oracle_sample_linear(const uint8_t* rgba, uint32_t W, uint32_t H, float u, float v, float out[4]) {
    sample_linear(rgba, W, H, u, v, out);
}

/* (int) of a float as the GPU converts it: NaN -> 0, saturating (the
 * reference's implicit conversion is undefined out of range). */
static int to_int_sat(float m) {
    if (m != m) return 0;
    if (m >= 2147483648.f) return 0x7FFFFFFF;
    if (m <= -2147483648.f) return (int)0x80000000u;
    return (int)m;
}

/* do_motion_blur, cl2.cl:6715-6858, as engine::do_motion_blur binds it
 * (engine.cpp:1518-1536), over the frame's host buffers.  Pixels are visited
 * in row order; the pose writes (cl2.cl:6770-6785) go to the slot that no
 * pixel of the pass reads, so the order does not matter. */
void oracle_motion_blur(const ocr_view* view, const oracle_sincos* cam, const float c_pos_old[3],
                        const oracle_sincos* cam_old, const uint32_t* depth, const uint32_t* tri_id,
                        const ocr_triangle* tris, uint32_t ntri, ocr_obj_desc* objs, uint32_t nobj,
                        uint32_t frame_id, float strength, float camera_contribution, uint8_t* rgba) {
    const uint32_t W = view->width, H = view->height;
    const float fov = view->fov_const;
    const size_t npix = (size_t)W * H;
    uint8_t* in = (uint8_t*)malloc(npix * 4);
    if (!in) return;
    memcpy(in, rgba, npix * 4);
    const f3 c_pos = ld3(view->c_pos), c_old = ld3(c_pos_old);
    const int even = (frame_id & 1) == 0;
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            const size_t p = (size_t)y * W + x;
            const uint32_t dbuf_val = depth[p], t = tri_id[p];
            if (dbuf_val == 0xFFFFFFFFu || t >= ntri) continue;   /* background: nothing written */
            const uint32_t o_id = tris[t].vertices[0].object_id;
            if (o_id >= nobj) continue;
            ocr_obj_desc* G = &objs[o_id];
            const float ldepth = ((float)dbuf_val / MULINT_F) * DEPTH_FAR;   /* idcalc */
            const float actual_depth = ldepth;
            const f3 local_position = V3(((float)x - (float)W / 2.0f) * actual_depth / fov,
                                         ((float)y - (float)H / 2.0f) * actual_depth / fov, actual_depth);
            const f3 global_position = add3(back_rot(local_position, cam), c_pos);
            const f4 q = {G->world_rot_quat[0], G->world_rot_quat[1], G->world_rot_quat[2], G->world_rot_quat[3]};
            const f3 object_local = back_rot_quat(sub3(global_position, ld3(G->world_pos)), q);
            f3 old_world_pos;
            f4 old_world_quat;
            if (even) {
                old_world_pos = ld3(G->old_world_pos_1);
                old_world_quat = (f4){G->old_world_rot_quat_1[0], G->old_world_rot_quat_1[1],
                                      G->old_world_rot_quat_1[2], G->old_world_rot_quat_1[3]};
                memcpy(G->old_world_pos_2, G->world_pos, 12);
                memcpy(G->old_world_rot_quat_2, G->world_rot_quat, 16);
            } else {
                old_world_pos = ld3(G->old_world_pos_2);
                old_world_quat = (f4){G->old_world_rot_quat_2[0], G->old_world_rot_quat_2[1],
                                      G->old_world_rot_quat_2[2], G->old_world_rot_quat_2[3]};
                memcpy(G->old_world_pos_1, G->world_pos, 12);
                memcpy(G->old_world_rot_quat_1, G->world_rot_quat, 16);
            }
            f3 last_frame_pos = add3(rot_quat(object_local, old_world_quat), old_world_pos);
            f3 last_frame_no_camera = rot(last_frame_pos, c_pos, cam);
            last_frame_pos = rot(last_frame_pos, c_old, cam_old);
            last_frame_no_camera = depth_project1(last_frame_no_camera, (float)W, (float)H, fov);
            last_frame_pos = depth_project1(last_frame_pos, (float)W, (float)H, fov);
            if (last_frame_pos.z < 20.f) {      /* depth_icutoff: the pixel's own colour */
                float c[4];
                sample_linear(in, W, H, (float)x + 0.5f, (float)y + 0.5f, c);
                for (int k = 0; k < 4; k++) rgba[4 * p + k] = (uint8_t)qz(c[k]);
                continue;
            }
            const float csx = (float)x, csy = (float)y;
            float tmx = csx - last_frame_pos.x, tmy = csy - last_frame_pos.y;
            const float ncx = csx - last_frame_no_camera.x, ncy = csy - last_frame_no_camera.y;
            tmx = camera_contribution * tmx + (1.f - camera_contribution) * ncx;
            tmy = camera_contribution * tmy + (1.f - camera_contribution) * ncy;
            tmx *= strength;
            tmy *= strength;
            int n = to_int_sat(fmaxf(fabsf(tmx), fabsf(tmy)) + 1.f);
            const int bound = 50;
            if (n > bound) {
                tmx /= (float)n;
                tmy /= (float)n;
                tmx *= (float)bound;
                tmy *= (float)bound;
                n = bound;
            }
            float dx = 0.f, dy = 0.f;
            if (n != 0) {
                dx = tmx / (float)n;
                dy = tmy / (float)n;
            }
            float cx = csx - tmx / 2.f, cy = csy - tmy / 2.f;
            float acc[4] = {0.f, 0.f, 0.f, 0.f}, fcount = 0.f;
            for (int i = 0; i < n; i++, cx += dx, cy += dy) {
                if (cx < 0.f || cx >= (float)W || cy < 0.f || cy >= (float)H) continue;
                float c[4];
                sample_linear(in, W, H, cx + 0.5f, cy + 0.5f, c);
                for (int k = 0; k < 4; k++) acc[k] += c[k] * 1.f;
                fcount += 1.f;
            }
            if (fcount != 0.f)
                for (int k = 0; k < 4; k++) acc[k] /= fcount;
            for (int k = 0; k < 4; k++) rgba[4 * p + k] = (uint8_t)qz(acc[k]);
        }
    free(in);
}
