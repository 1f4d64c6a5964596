/*
 * ocr_oracle.h — CPU restatement of OpenCLRenderer's rasterisation +
 * deferred-shading path (cl2.cl prearrange/kernel1/kernel2/kernel3, the cube
 * shadow passes and the texture atlas/mip kernels).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the reference result the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * Arithmetic contract (the reference builds with -cl-fast-relaxed-math and
 * FP_CONTRACT ON, ocl.h:227 / cl2.cl:3, which leaves these implementation
 * defined; this restatement fixes them):
 *   - IEEE-754 binary32, round-to-nearest-even, no contraction of a*b+c;
 *   - mad(a,b,c) = fmaf(a,b,c); native_divide / native_recip / `/` =
 *     correctly rounded division; sqrt correctly rounded;
 *   - fast_normalize(v) = v * (1/sqrt(dot(v,v))); fast_length = sqrt(dot);
 *     dot is summed left to right;
 *   - sin/cos of the camera and cube-face angles are binary64 sin/cos
 *     rounded once (the camera's are supplied by the caller);
 *   - pow(x,2) = x*x; pow(x,4), powr(x,5) are formed in binary64 by exact
 *     squaring and rounded once to binary32;
 *   - tan is evaluated in binary64 (ocrx_tan) and rounded once to binary32;
 *     exp by a fixed binary32 sequence (oracle_exp);
 *   - float -> u32 conversions of depths and texels saturate (NaN -> 0);
 *   - the id buffer tie rule is the sequential one: fragments are produced
 *     in triangle order and the last fragment that passes the depth window
 *     writes the pixel.
 */
#ifndef OCR_ORACLE_H
#define OCR_ORACLE_H

#include <stdint.h>
#include "../include/ocr.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene {
    const ocr_triangle* tris;  uint32_t ntri;
    const ocr_obj_desc* objs;  uint32_t nobj;
    const ocr_light*    lights; uint32_t nlights;
    /* atlas: npages x 2048 x 2048 x RGBA8, descriptors as the reference's
     * g_texture_nums / g_texture_sizes / mipmap_start */
    const uint8_t*  atlas;     uint32_t npages;
    const uint32_t* nums;      uint32_t nnums;
    const uint32_t* sizes;     uint32_t nsizes;
    uint32_t        mip_start;
    /* cube maps: n x 6 x dim x dim u32 (NULL when the scene has none) */
    const uint32_t* shadow_maps;
    const uint32_t* static_maps;
    uint32_t        lbuf_dim;
} oracle_scene;

typedef struct oracle_out {
    uint32_t* depth;      /* W*H, required */
    uint32_t* tri_id;     /* W*H, may be NULL */
    uint8_t*  rgba8;      /* W*H*4, may be NULL (skips shading) */
    uint16_t* normals;    /* W*H*2, may be NULL */
    uint32_t  n_cut;      /* cut triangles that survived cull with >= 1 fragment */
    uint32_t  n_frag;     /* fragment records */
    uint64_t  n_samples;  /* covered samples (point_in_tri passes) in the depth pass */
    uint32_t  n_visible_tris;  /* distinct triangles in the id buffer */
    uint32_t  n_visible_objs;  /* distinct objects in the id buffer */
    double    ms_prearrange, ms_depth, ms_id, ms_shade;
} oracle_out;

/* Sine/cosine table of a rotation, computed by the caller. */
typedef struct oracle_sincos { float s[3]; float c[3]; } oracle_sincos;

/* Full frame: prearrange -> kernel1 -> kernel2 -> kernel3.
 * nthreads <= 1 runs every stage sequentially in work-item order (this is the
 * golden run).  nthreads > 1 runs the per-triangle, per-fragment and
 * per-pixel loops on OpenMP threads with the same results (prefix-summed
 * fragment order, atomic min for depth, atomic max of the fragment index for
 * the id buffer). */
int oracle_render(const oracle_scene* sc, const ocr_view* view,
                  const oracle_sincos* cam, int nthreads, oracle_out* out);

/* engine::generate_realtime_shadowing: fills dyn_maps (one 6-face map per
 * light with shadow == 1, dynamic geometry) and static_maps (one per light
 * with shadow && is_static, static geometry).  The cube-face rotations'
 * sin/cos come from oracle_face_sincos. */
int oracle_shadow_maps(const oracle_scene* sc, int nthreads, uint32_t* dyn_maps,
                       uint32_t* static_maps);
/* The same, also returning the covered samples (depth tests) of all maps. */
int oracle_shadow_maps_ex(const oracle_scene* sc, int nthreads, uint32_t* dyn_maps,
                          uint32_t* static_maps, uint64_t* samples);
/* sin/cos of the six cube-face rotations r_struct (cl2.cl:4487-4511). */
void oracle_face_sincos(oracle_sincos out[6]);

/* texture_context::alloc_gpu page plan: given each texture's width/height,
 * returns page count, descriptor count (ntex*5) and fills nums/sizes. */
int oracle_atlas_plan(const uint32_t* widths, const uint32_t* heights, uint32_t ntex,
                      uint32_t* npages, uint32_t* nums, uint32_t* sizes,
                      uint32_t* mip_start);
/* update_gpu_tex (flip = 1) + generate_mips + generate_mip_mips for every
 * texture, into `pages` (npages x 2048^2 x 4, zero-initialised by caller). */
int oracle_atlas_fill(const ocr_tex_src* texs, uint32_t ntex, const uint32_t* nums,
                      const uint32_t* sizes, uint32_t mip_start, uint8_t* pages);

/* ---- single functions, exported for closed-form / brute-force tests ---- */
/* Surviving cut triangles of a view in prearrange order (n x 9 floats:
 * x,y,z per vertex) with key 2*tri + sub; returns the count. */
int   oracle_cut_list(const oracle_scene* sc, const ocr_view* v, const oracle_sincos* cam,
                      float* cuts, uint32_t* keys, int max_out);
void  oracle_rot(const float p[3], const float cpos[3], const oracle_sincos* sc, float out[3]);
void  oracle_back_rot(const float p[3], const oracle_sincos* sc, float out[3]);
void  oracle_rot_quat(const float p[3], const float q[4], float out[3]);
void  oracle_back_rot_quat(const float p[3], const float q[4], float out[3]);
/* generate_new_triangles: returns count (0..2) and fills out[2][3][3]. */
int   oracle_clip(const float pts[3][3], float out[2][3][3]);
void  oracle_calc_min_max(const float pts[3][3], float w, float h, float mm[4]);
int   oracle_point_in_tri(float x, float y, const float xs[3], const float ys[3]);
/* The kernel1 walk of one fragment of one cut triangle: writes up to
 * max_out (x, y, depth) triples of covered samples; returns the count. */
int   oracle_walk(const float cut[3][3], float w, float h, uint32_t chunk, int op_size,
                  int32_t* xs, int32_t* ys, uint32_t* ds, int max_out);
uint32_t oracle_wang_hash(uint32_t s);
uint32_t oracle_xorshift(uint32_t s);
float oracle_rational_acos(float x);
float oracle_log2_approx(float x);
void  oracle_gamma_transform(const float in[4], float out[4]);
void  oracle_gamma_inverse(const float in[4], float out[4]);
int   oracle_ret_cubeface(const float p[3], const float l[3]);
float oracle_exp(float x);
void  oracle_exp_n(const float* x, float* y, uint32_t n);
float oracle_tan(float x);
float oracle_powr5(float x);
float oracle_pow4(float x);
void  oracle_get_barycentric(const float p[3], const float a[3], const float b[3],
                             const float c[3], float* u, float* v, float* w);
uint32_t oracle_encode_normal(const float n[3]);
/* generate_ssao (cl2.cl:2194-2260) at pixel (x, y) of a depth buffer. */
float oracle_ssao(int x, int y, const uint32_t* depth, uint32_t W, uint32_t H, float fov);
/* The shadow term kernel3 forms for one light: ret_cubeface of the point,
 * then generate_hard_occlusion (cl2.cl:2536-2701) on cube map `shnum` of
 * `maps` (n x 6 x L x L u32).  0 = lit, 1 = fully shadowed. */
float oracle_shadow_term(const float gpos[3], const float normal[3], const float lpos[3],
                         const uint32_t* maps, uint32_t shnum, uint32_t L);
/* texture_filter_diff (cl2.cl:1511-1573) on an atlas: RGBA in [0, 1]. */
void  oracle_texture_filter(float vtx, float vty, float dx, float dy, int tid, uint32_t mip_start,
                            const uint32_t* nums, const uint32_t* sizes, const uint8_t* atlas,
                            float out[4]);
/* One light's cube map (6 x L x L, from the dynamic geometry, or the static
 * one with only_static), sequentially, with the fragment records
 * prearrange_realtime_shadowing emits per face (may be NULL). */
int   oracle_shadow_map_one(const oracle_scene* sc, uint32_t light, int only_static, uint32_t* map,
                            uint64_t frags[6]);
void  oracle_texture_mod(const float vt[2], float out[2]);

/* blend_screens_with_depth (cl2.cl:7842-7875) with alpha_blend
 * (cl2.cl:1721-1728), engine::blend_with_depth (engine.cpp:2550-2570): for
 * every pixel where dst_depth >= src_depth, dst_rgba = src over dst (RGBA8
 * read as c / 255, written with the shading pass's quantisation); dst depth
 * is left as it is. */
void  oracle_blend_with_depth(uint32_t npix, const uint32_t* src_depth, const uint32_t* dst_depth,
                              const uint8_t* src_rgba, uint8_t* dst_rgba);

/* do_pseudo_aa (cl2.cl:6438-6657), engine::do_pseudo_aa (engine.cpp:1513-1516)
 * on one frame: depth (u32), encoded normals (ushort2 per pixel) and RGBA8
 * of a render.  Reads the colours as they were before the pass and rewrites
 * rgba in place where an edge qualifies. */
void  oracle_pseudo_aa(uint32_t width, uint32_t height, const uint32_t* depth, const uint16_t* normals,
                       uint8_t* rgba);

/* do_motion_blur (cl2.cl:6715-6858), engine::do_motion_blur
 * (engine.cpp:1518-1536) on one frame: its depth, triangle ids
 * (OCR_NO_TRIANGLE where nothing is visible) and RGBA8 (rewritten in place,
 * from the colours as they were before the pass), the scene's triangles (for
 * each visible triangle's object id) and descriptors (old_world_* written
 * for every object with a covered pixel), the view it was drawn with and its
 * camera sin/cos, and the previous camera. */
void  oracle_motion_blur(const ocr_view* view, const oracle_sincos* cam, const float c_pos_old[3],
                         const oracle_sincos* cam_old, const uint32_t* depth, const uint32_t* tri_id,
                         const ocr_triangle* tris, uint32_t ntri, ocr_obj_desc* objs, uint32_t nobj,
                         uint32_t frame_id, float strength, float camera_contribution, uint8_t* rgba);
/* The bilinear sampler it reads the colours with (unnormalised u, v). */
void  oracle_sample_linear(const uint8_t* rgba, uint32_t width, uint32_t height, float u, float v,
                           float out[4]);

#ifdef __cplusplus
}
#endif
#endif
