/*
 * ocr.h — C ABI of the MI355X-native rasterisation + deferred-shading path.
 *
 * This is the drop-in boundary for OpenCLRenderer's draw path.  Every entry
 * point takes plain pointers and sizes; device work is ordered on a caller
 * supplied HIP stream passed as `void*` (a hipStream_t, NULL = the legacy
 * default stream).  All entry points return OCR_OK (0) or a negative status;
 * ocr_last_error() returns a thread-local message for the last failure.
 *
 * Record layouts below are byte-compatible with the reference's host structs,
 * so buffers the reference's object_context / light code build can be handed
 * over unchanged.
 */
#ifndef OCR_H_INCLUDED
#define OCR_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define OCR_OK               0
#define OCR_ERR_INVALID     -1  /* bad handle, size or argument            */
#define OCR_ERR_HIP         -2  /* a HIP runtime call failed               */
#define OCR_ERR_OOM         -3  /* device allocation failed                */
#define OCR_ERR_OVERFLOW    -4  /* a reserved cut list was too small; it
                                   was grown and the frame must be re-rendered */
#define OCR_ERR_UNSUPPORTED -5  /* input outside what the path supports     */
#define OCR_ERR_BUSY        -6  /* admission refused (serving)              */

/* ---- fixed constants of the path (cl2.cl:17, :21-23, :4247-4249) ------- */
#define OCR_DEPTH_FAR        350000.0f
#define OCR_DEPTH_ICUTOFF    20
#define OCR_OP_SIZE          500
#define OCR_OP_SIZE_LIGHT    300
#define OCR_BUF_ERROR        20
#define OCR_SHADOWBIAS       50
#define OCR_MIP_LEVELS       4
#define OCR_ATLAS_DIM        2048
#define OCR_NO_TRIANGLE      0xFFFFFFFFu

/* object feature flags (cl2.cl:90-97) */
#define OCR_FEATURE_SS_REFLECTIVE                    1
#define OCR_FEATURE_TWO_SIDED                        2
#define OCR_FEATURE_OUTLINE                          4
#define OCR_FEATURE_IS_STATIC                        8
#define OCR_FEATURE_DOES_NOT_RECEIVE_DYNAMIC_SHADOWS 16

/* ---- records ----------------------------------------------------------- */

/* struct vertex, cl2.cl:131-138 / vertex.hpp:8-38 (48 bytes).  object_id is
 * only meaningful in vertices[0] (object_context.cpp:421-446, cl2.cl:4240);
 * vertex_col == 0 in vertices[0] means "use the texture" (cl2.cl:5933). */
typedef struct ocr_vertex {
    float    pos[4];
    float    normal[4];
    float    vt[2];
    uint32_t object_id;
    uint32_t vertex_col;        /* 0xRRGGBBAA */
} ocr_vertex;

/* struct triangle, cl2.cl:148-151 / triangle.hpp:8-15 (144 bytes).  The
 * triangle id of the id buffer is the index into the array handed to
 * ocr_scene_create (object_context.cpp:398-453 concatenation order). */
typedef struct ocr_triangle {
    ocr_vertex vertices[3];
} ocr_triangle;

/* struct obj_g_descriptor, cl2.cl:99-124 / obj_g_descriptor.hpp:9-36
 * (136 bytes of fields, 144 with cl_float4 alignment). */
typedef struct ocr_obj_desc {
    float    world_pos[4];
    float    world_rot_quat[4];
    float    old_world_pos_1[4];
    float    old_world_pos_2[4];
    float    old_world_rot_quat_1[4];
    float    old_world_rot_quat_2[4];
    float    scale;
    uint32_t tid;               /* texture position id (texture_context::get_gpu_position_id) */
    uint32_t rid;
    uint32_t ssid;
    uint32_t has_bump;
    float    specular;
    float    spec_mult;
    float    diffuse;
    int32_t  buffer_offset;
    int32_t  feature_flag;
    uint32_t _pad[2];
} ocr_obj_desc;

/* struct light, cl2.cl:77-87 / light.hpp:25-69 (56 bytes, 64 aligned). */
typedef struct ocr_light {
    float    pos[4];
    float    col[4];
    uint32_t shadow;
    float    brightness;
    float    radius;
    float    diffuse;
    float    godray_intensity;
    int32_t  is_static;
    uint32_t _pad[2];
} ocr_light;

/* One view.  Replaces the compile-time -D SCREENWIDTH/SCREENHEIGHT/FOV_CONST
 * options (ocl.h:227-240) and the c_pos/c_rot kernel arguments
 * (engine.cpp:1868-1877) with runtime values, so views of different sizes can
 * share one launch.  c_rot is extrinsic-XYZ Euler radians (cl2.cl:220-236). */
typedef struct ocr_view {
    uint32_t width;
    uint32_t height;
    float    fov_const;         /* (width/2)/tan(hfov/2), engine.cpp:119-133; in [2^-60, 2^60] */
    uint32_t use_linear_rendering;  /* object_context.hpp:87, TEST_LINEAR path */
    float    c_pos[4];
    float    c_rot[4];
    float    clear_col[4];      /* object_context::set_clear_colour */
} ocr_view;

/* A source texture: tightly packed RGBA8 rows, row 0 first (sf::Image order). */
typedef struct ocr_tex_src {
    uint32_t       width;
    uint32_t       height;
    const uint8_t* rgba;
} ocr_tex_src;

/* Counters of the last render of a frame, and per-pass device times when
 * profiling is on: hipEvent pairs on the render stream, averaged over every
 * render since profiling was (re)enabled. */
typedef struct ocr_frame_stats {
    uint32_t n_cut;             /* cut (post clip/cull) triangles            */
    uint32_t n_frag;            /* fragment records                          */
    uint32_t overflow;          /* 1 if a work list hit its capacity         */
    uint32_t n_profiled;        /* renders the pass times are averaged over  */
    float    ms_clear;
    float    ms_prearrange;
    float    ms_depth;
    float    ms_id;
    float    ms_shade;
    float    ms_total;
} ocr_frame_stats;

typedef struct ocr_ctx   ocr_ctx;
typedef struct ocr_scene ocr_scene;
typedef struct ocr_frame ocr_frame;

/* ---- context ------------------------------------------------------------ */

/* Replaces oclstuff()/build() (ocl.h:328-537): selects HIP device `device`. */
int  ocr_create(int device, ocr_ctx** out);
void ocr_destroy(ocr_ctx* ctx);
/* Thread-local description of the last non-OK status. */
const char* ocr_last_error(void);
/* Free / total device memory, the serving admission input. */
int  ocr_query_free_hbm(ocr_ctx* ctx, size_t* free_bytes, size_t* total_bytes);

/* ---- scene: device-resident geometry, descriptors, atlas, lights -------- */

/* object_context::build (object_context.cpp:646) + alloc_gpu (:346) +
 * alloc_object_descriptors (:460): uploads `ntri` triangles and `nobj`
 * descriptors.  Host buffers may be freed on return. */
int  ocr_scene_create(ocr_ctx* ctx, const ocr_triangle* tris, uint32_t ntri,
                      const ocr_obj_desc* objs, uint32_t nobj, ocr_scene** out);
/* object::g_flush (object.cpp:652-857): rewrite descriptors [first, first+n). */
int  ocr_scene_update_objs(ocr_scene* s, uint32_t first, uint32_t n,
                           const ocr_obj_desc* objs);
/* The same, stream-ordered like the reference's enqueued descriptor writes
 * (object::g_flush, object.cpp:652-857, on the render queue): renders,
 * shadow passes, batches and post passes issued afterwards, on any stream,
 * see the update, and work enqueued on `stream` afterwards follows it; work
 * issued before it keeps seeing the state before it.  The scene holds two
 * versions of its descriptors and world-space triangles: the update builds
 * the one no pass issued since the previous update reads (a copy of the
 * current version plus the new records, on a stream of the scene's own,
 * after the passes that read that version have finished), so it overlaps
 * the renders in flight.  The descriptors are staged in pinned memory (the
 * host array may be reused on return); the host blocks only when four
 * earlier updates of the scene are still in flight.  The second version is
 * allocated on the first call (40 bytes per triangle + 64 per object). */
int  ocr_scene_update_objs_async(ocr_scene* s, uint32_t first, uint32_t n,
                                 const ocr_obj_desc* objs, void* stream);
/* texture_context::alloc_gpu (texture_context.cpp:350-517) +
 * texture::update_me_to_gpu / update_gpu_mipmaps (texture.cpp:323-493):
 * packs the textures into 2048x2048 RGBA8 pages per power-of-two size class,
 * uploads them (update_gpu_tex, cl2.cl:923-953) and builds 4 mip levels
 * (generate_mips / generate_mip_mips, cl2.cl:1070-1189).  Texture i of the
 * array is descriptor tid == i. */
int  ocr_atlas_upload(ocr_scene* s, const ocr_tex_src* texs, uint32_t ntex);
/* light::build (light.cpp:145-276): uploads the light list and sizes the
 * cube shadow maps (n_shadow x 6 x dim^2 u32, dim = LIGHTBUFFERDIM, 8..4000). */
int  ocr_lights_upload(ocr_scene* s, const ocr_light* lights, uint32_t n,
                       uint32_t light_buffer_dim);
/* engine::generate_realtime_shadowing (engine.cpp:1601-1790): clears and
 * renders every shadow-casting light's cube map from dynamic geometry, and
 * the static-light maps from static geometry when lights were re-uploaded.
 * Stream-ordered like ocr_scene_update_objs_async: renders, batches and work
 * on `stream` issued afterwards see the new maps, renders issued before keep
 * the maps they were issued with.  The scene holds up to two sets of dynamic
 * maps (the second allocated when a pass follows a render): a pass renders
 * the set no render issued since the previous pass reads, on a stream of the
 * scene's own, so it overlaps the shading of the frame before it. */
int  ocr_shadow_pass(ocr_scene* s, void* stream);
void ocr_scene_destroy(ocr_scene* s);

/* Inspection of device state (tests, tools). */
int  ocr_scene_atlas_info(ocr_scene* s, uint32_t* npages, uint32_t* nnums,
                          uint32_t* nsizes, uint32_t* mip_start);
int  ocr_scene_read_atlas(ocr_scene* s, uint8_t* pages, uint32_t* nums,
                          uint32_t* sizes);
int  ocr_scene_read_shadow_maps(ocr_scene* s, uint32_t* dynamic_maps,
                                uint32_t* static_maps);
/* The motion-blur poses of objects [first, first+n): 16 floats each,
 * old_world_pos_1, old_world_pos_2, old_world_rot_quat_1,
 * old_world_rot_quat_2 (obj_g_descriptor.hpp:15-18).  They start as the
 * descriptors given to ocr_scene_create; ocr_scene_update_objs leaves them
 * alone (as object::g_flush does) and ocr_frame_motion_blur records them. */
int  ocr_scene_read_motion(ocr_scene* s, uint32_t first, uint32_t n, float* poses);

/* ---- frames: offscreen framebuffer (replaces cl_gl_interop_texture) ----- */

/* Frames are 4x4 up to 16,000,000 pixels (3840 x 2160, 5120 x 2880 ...). */
int  ocr_frame_create(ocr_ctx* ctx, uint32_t width, uint32_t height, ocr_frame** out);
void ocr_frame_destroy(ocr_frame* f);
/* Fixes the capacity of the frame's cut-triangle list (the reference's
 * fixed g_tid_buf_max_len, engine.cpp:600-618).  By default the list holds
 * 2 pieces per scene triangle and never overflows; a reserved list that is
 * too small sets the overflow word, ocr_frame_read then reports
 * OCR_ERR_OVERFLOW and the next render grows the list to fit. */
int  ocr_frame_reserve(ocr_frame* f, uint32_t max_cut_triangles);
/* Raster path of the frame's renders (ocr_render).  Both give the same
 * depth and triangle-id buffers, bit for bit.
 *   OCR_RASTER_SAMPLES  kernel1 / kernel2 (cl2.cl:4985-5127, :5390-5546):
 *       every cut triangle's fragment walk in one global sample space, a
 *       depth pass with u32 atomicMin and an id pass with u64 atomicMax on
 *       the frame's buffers (the default).
 *   OCR_RASTER_TILES    engine::draw_tiled_deferred (engine.cpp:2447-2535:
 *       split_into_tiled_chunks, tile_render, cl2.cl:3850-4210): cut
 *       triangles binned to screen tiles of 2^tile_w_log2 x 2^tile_h_log2
 *       pixels (256 to 2048 pixels, sides 8 to 64), each tile's depth and
 *       ids resolved in on-chip memory by one workgroup and written once.
 *       The bins hold bin_capacity entries (0: two per cut-triangle slot);
 *       a render that needs more reports OCR_ERR_OVERFLOW at ocr_frame_read
 *       and the next render grows them.
 * ocr_frame_stats then reports the bin count + scatter as ms_clear, the
 * tile pass as ms_depth and 0 as ms_id. */
#define OCR_RASTER_SAMPLES 0u
#define OCR_RASTER_TILES   1u
typedef struct ocr_raster_config {
    uint32_t mode;              /* OCR_RASTER_SAMPLES or OCR_RASTER_TILES */
    uint32_t tile_w_log2;       /* 3..6 */
    uint32_t tile_h_log2;       /* 3..6, tile_w_log2 + tile_h_log2 in 8..11 */
    uint32_t bin_capacity;      /* bin entries (0 = automatic) */
} ocr_raster_config;
int  ocr_frame_set_raster(ocr_frame* f, const ocr_raster_config* cfg);
int  ocr_frame_get_raster(ocr_frame* f, ocr_raster_config* cfg);
/* Record hipEvent timings on every render of this frame (up to 1024
 * renders are averaged; enabling again restarts the average): enable 1
 * times every pass, 2 only the shading pass (two events per render, for
 * timing a stream of renders with as little added as possible), 0 stops.
 * The shading pass is timed from the end of its wait for the scene's
 * shadow maps. */
int  ocr_frame_set_profiling(ocr_frame* f, int enable);

/* engine::draw_bulk_objs_n -> render_tris (engine.cpp:1794-2161):
 * prearrange -> depth (kernel1) -> id (kernel2) -> shade (kernel3), with no
 * host round trip.  The view size must equal the frame's.  A frame holds
 * three output sets and each render writes the one written longest ago, on
 * a stream of the frame's own, so render N+1 can start while renders N and
 * N-1 are still running; `stream` waits for the render's end.  Render N+1
 * waits for the work on `stream` enqueued before render N-1 was issued:
 * enqueue what reads a render's outputs (ocr_frame_device_buffers changes
 * with every render) on `stream` before the next ocr_render on the frame,
 * and use one stream for the renders of one frame.  A frame serves either ocr_render or batches:
 * a frame already set on a batch is refused (OCR_ERR_INVALID). */
int  ocr_render(ocr_scene* s, const ocr_view* view, ocr_frame* f, void* stream);
/* No counterpart in the reference: ocr_render for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Renders `view` of `s` on `stream` into
 * the caller's device buffers of view->width * view->height elements: depth
 * (u32), id_key (u64: ((2*tri + piece + 1) << 32) | record, 0 where nothing
 * is drawn; ocr_frame_read's triangle id is (key >> 32) - 1 >> 1), rgba8
 * (u32) and normals (u32, two u16).  The work lists live in a 64 MiB device
 * scratch array the library holds per GPU (524,286 cut-triangle records, two
 * per scene triangle), so the scene may have at most 262,111 triangles
 * (OCR_ERR_UNSUPPORTED beyond); the scene's lights must already be uploaded.  Calls on one GPU share the
 * scratch: calls that overlap on one GPU race (order them on one stream).
 * A view needing more than 2^31 - 1 raster samples is not drawn (outputs
 * hold the clear); ocr_render reports that case. */
int  ocr_render_noalloc(ocr_scene* s, const ocr_view* view, uint32_t* depth, unsigned long long* id_key,
                        uint32_t* rgba8, uint32_t* normals, void* stream);

/* Replaces the CL/GL blit and its host read-back path
 * (cl_gl_interop_texture.hpp:189-252, :312-331).  Synchronises `stream`,
 * then copies any non-NULL output: depth u32 (UINT_MAX = background),
 * triangle id u32 (OCR_NO_TRIANGLE where nothing is visible), RGBA8 and the
 * encoded screen normals (ushort2 per pixel, cl2.cl:5617-5630).  Returns
 * OCR_ERR_OVERFLOW if the last render overflowed a reserved cut list, and
 * OCR_ERR_UNSUPPORTED if it needed more than 2^31 - 1 raster samples. */
int  ocr_frame_read(ocr_frame* f, uint32_t* depth, uint32_t* tri_id,
                    uint8_t* rgba8, uint16_t* normals, void* stream);
/* engine::blend_with_depth -> blend_screens_with_depth (engine.cpp:2550-2570,
 * cl2.cl:7842-7875): composites src onto dst (same size) on `stream`:
 * wherever dst's depth is not nearer than src's, dst's RGBA8 becomes src
 * over dst (alpha_blend, cl2.cl:1721-1728); dst's depth is unchanged.  Uses
 * each frame's last render; enqueue it on the stream those renders used. */
int  ocr_frame_blend_with_depth(ocr_frame* src, ocr_frame* dst, void* stream);
/* engine::do_pseudo_aa -> do_pseudo_aa (engine.cpp:1513-1516,
 * cl2.cl:6438-6657): edge-aware smoothing of the frame's last render on
 * `stream`, from its depth, encoded normals and colours (read as they were
 * before the pass); RGBA8 is rewritten in place where an edge qualifies.
 * Allocates a colour snapshot (4 bytes per pixel) on first use. */
int  ocr_frame_pseudo_aa(ocr_frame* f, void* stream);
/* engine::do_motion_blur -> do_motion_blur (engine.cpp:1518-1536,
 * cl2.cl:6715-6858): blurs the frame's last render, drawn from scene `s`
 * with `view`, on `stream`.  Each covered pixel is carried back to where it
 * was in the previous frame (camera c_pos_old / c_rot_old, its object's pose
 * recorded by the last pass of the other frame_id parity) and replaced by
 * the mean of up to 50 bilinear samples of the colours (as they were before
 * the pass) along the line from there, scaled by `strength`;
 * camera_contribution weighs camera motion against object motion.  Pixels
 * whose earlier position lies behind the previous camera's near plane keep
 * their colour; background pixels are left alone (the reference writes
 * nothing there).  Records every drawn object's current pose in the slot
 * the other parity reads (old_world_*_2 when frame_id is even, _1 when odd).
 * Passes on one scene race unless ordered on one stream.  Allocates a
 * colour snapshot (4 bytes per pixel) on first use. */
int  ocr_frame_motion_blur(ocr_frame* f, ocr_scene* s, const ocr_view* view,
                           const float c_pos_old[4], const float c_rot_old[4],
                           uint32_t frame_id, float strength,
                           float camera_contribution, void* stream);
/* Device pointers of the frame's buffers (zero-copy consumers). */
int  ocr_frame_device_buffers(ocr_frame* f, void** depth, void** idkey,
                              void** rgba8, void** normals);
int  ocr_frame_get_stats(ocr_frame* f, void* stream, ocr_frame_stats* out);
/* Instrumented re-run of the shading pass of the last render: counts the
 * distinct atlas texels and cube-map texels it reads (the X_tex / X_shadow
 * terms of the shading pass's algorithmic byte count).  Allocates scratch. */
int  ocr_frame_count_texels(ocr_scene* s, const ocr_view* view, ocr_frame* f,
                            uint64_t* atlas_texels, uint64_t* shadow_texels,
                            void* stream);

/* ---- batched form: many views of several resident scenes ------------- */

/* One render of a batch covers every view with one launch per pass; per-view
 * screen size, FOV and camera are runtime values (the reference bakes them
 * into its program, ocl.h:227-240, and draws one view per call,
 * engine.cpp:1794-2161). */
typedef struct ocr_batch ocr_batch;
typedef struct ocr_batch_tuning {
    uint32_t raster_blocks_per_cu;   /* persistent depth / id grid: blocks per CU (1..32) */
    uint32_t shade_tile_w_log2;      /* shading tile: 2^k wide, 256 >> k high (0..8) */
    uint32_t groups;                 /* 0 or 1: one launch sequence for all views; 2..8:
                                        the views split in that many pixel-balanced groups,
                                        each on a stream of the batch's own, so one group's
                                        shading overlaps another's raster passes */
    uint32_t raster_tile;            /* 0: the sample-space depth and id passes; otherwise
                                        the tiled raster path (see ocr_frame_set_raster) with
                                        tiles of 2^(raster_tile & 0xFF) x 2^(raster_tile >> 8)
                                        pixels */
} ocr_batch_tuning;

int  ocr_batch_create(ocr_ctx* ctx, uint32_t max_views, ocr_batch** out);
void ocr_batch_destroy(ocr_batch* b);
/* View i renders scenes[i] into frames[i] (distinct frames, frame size =
 * view size).  Uploads the view and scene tables and sizes the shared work
 * lists; enqueues no rendering.  Frames rendered by ocr_render are refused:
 * a batch frame has no cut list or counters of its own, so the passes that
 * need them (ocr_frame_motion_blur, ocr_frame_count_texels,
 * ocr_frame_get_stats) refuse it, and ocr_batch_get_stats reports the
 * batch's status. */
int  ocr_batch_set_views(ocr_batch* b, ocr_scene* const* scenes, const ocr_view* views,
                         ocr_frame* const* frames, uint32_t n);
int  ocr_batch_set_tuning(ocr_batch* b, const ocr_batch_tuning* t);
int  ocr_batch_get_tuning(ocr_batch* b, ocr_batch_tuning* t);
/* hipEvent timing around every render / graph launch (average in
 * ocr_batch_get_stats ms_total). */
int  ocr_batch_set_profiling(ocr_batch* b, int enable);
/* clear -> prearrange -> depth -> id -> shade for all views on `stream`. */
int  ocr_batch_render(ocr_batch* b, void* stream);
/* The same five launches captured once into a hipGraph and replayed. */
int  ocr_batch_launch_graph(ocr_batch* b, void* stream);
/* Synchronises `stream`; counters of the last render (OCR_ERR_UNSUPPORTED if
 * the views needed more than 2^31 - 1 raster samples together). */
int  ocr_batch_get_stats(ocr_batch* b, void* stream, ocr_frame_stats* out);
/* Times raster grid sizes x shading tile shapes x groups (1, 2, 4) on the
 * current views and keeps the fastest (reps graph launches each). */
int  ocr_batch_autotune(ocr_batch* b, void* stream, uint32_t reps, ocr_batch_tuning* best);

/* ---- served form ----------------------------------------------------------
 * No counterpart in the reference, whose main loop renders one view per draw
 * call (main.cpp:254-342 -> engine::draw_bulk_objs_n, engine.cpp:2356-2445).
 * A server keeps scenes and their shadow maps resident, admits requests
 * against a budget of free HBM and renders whatever is pending as one batch
 * (ocr_batch) each time one of max_inflight batch slots frees. */

typedef struct ocr_server ocr_server;

#define OCR_SERVE_READ_RGBA  1u   /* completion host buffer holds RGBA8      */
#define OCR_SERVE_READ_DEPTH 2u   /* ... then the u32 depth buffer           */
#define OCR_SERVE_READ_IDKEY 4u   /* ... then the u64 id keys (see ocr_frame_read) */

typedef struct ocr_server_config {
    uint32_t max_batch_views;   /* requests per batch, 1..4096             */
    uint32_t max_inflight;      /* batch slots (streams) in flight, 1..8   */
    uint64_t max_batch_tris;    /* scene triangles per batch (a larger single
                                   request still runs alone); sizes each
                                   slot's cut list                         */
    float    hbm_fraction;      /* of the free HBM at creation, held by
                                   admitted requests' frames at most       */
    uint32_t read_mask;         /* OCR_SERVE_READ_* copied to host per request */
} ocr_server_config;            /* 24 B */

typedef struct ocr_completion {
    uint64_t request_id;
    uint64_t t_submit_ns, t_start_ns, t_done_ns;   /* ocr_server_now_ns clock */
    void*    host;              /* pinned outputs per read_mask; valid until
                                   ocr_server_release                      */
    int32_t  status;            /* OCR_OK or the batch's error             */
    uint32_t scene_slot, width, height;
    uint32_t batch_views;       /* requests rendered in the same batch     */
    uint32_t _pad;
} ocr_completion;               /* 64 B */

typedef struct ocr_server_info {
    uint64_t hbm_budget, hbm_held;
    uint64_t n_batches, n_done, n_refused;
    uint64_t n_splits;          /* batches re-queued in halves because their
                                   views needed more raster samples than one
                                   render holds                            */
    uint32_t n_pending, n_frames;
    uint64_t frame_bytes;       /* HBM of the pooled frames, idle and in use
                                   (kept within hbm_budget)                */
    uint64_t n_evicted;         /* idle frames freed to make room for a
                                   frame of another size                   */
} ocr_server_info;              /* 72 B */

int  ocr_server_create(ocr_ctx* ctx, const ocr_server_config* cfg, ocr_server** out);
/* Registers a scene (atlas uploaded) and renders its shadow maps once. */
int  ocr_server_add_scene(ocr_server* s, ocr_scene* scene, uint32_t* slot);
/* OCR_ERR_BUSY when the request's frame (20 B per pixel) does not fit the
 * HBM budget beside the admitted, unreleased requests.  Frames are pooled per
 * size; idle frames of other sizes are freed when a new frame would take the
 * pool past the budget, so the frames' HBM stays within it. */
int  ocr_server_submit(ocr_server* s, uint32_t scene_slot, const ocr_view* view,
                       uint64_t* request_id);
/* Up to `max` completions; waits up to wait_ms for the first one. */
int  ocr_server_poll(ocr_server* s, ocr_completion* out, uint32_t max, uint32_t* n, int wait_ms);
/* Returns a completed request's host buffer and HBM budget to the server. */
int  ocr_server_release(ocr_server* s, uint64_t request_id);
int  ocr_server_get_info(ocr_server* s, ocr_server_info* info);
uint64_t ocr_server_now_ns(void);
/* Waits for in-flight batches, drops requests that never started.  A server
 * still alive at process exit has its threads stopped by an exit handler. */
void ocr_server_destroy(ocr_server* s);

#ifdef __cplusplus
}
#endif

#endif /* OCR_H_INCLUDED */
