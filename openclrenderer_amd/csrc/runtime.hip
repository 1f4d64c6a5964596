// runtime.hip — the C ABI of include/ocr.h: context, device-resident scene
// (geometry, descriptors, atlas, lights, cube maps), offscreen frames and the
// frame orchestration (clear -> prearrange -> depth -> id -> shade).
//
// Replaces the reference's OpenCL state (ocl.h:328-537, clstate.h:36-120),
// object_context_data's device buffers (object_context.hpp:82-165),
// texture_context::alloc_gpu (texture_context.cpp:350-517), light::build
// (light.cpp:145-276), engine::render_tris (engine.cpp:1794-2161) and
// engine::generate_realtime_shadowing (engine.cpp:1601-1790).
#pragma clang fp contract(off)

#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <new>

#include "ocr_internal.h"

namespace ocr {

static thread_local std::string g_last_error;

// Scratch of ocr_render_noalloc: the counters and the cut list (CutTri +
// RasterTri records) of one render, in a file-scope device array that every
// GPU the library runs on holds until the process ends.  Host code reaches
// it only through the address hipGetSymbolAddress returns for the device
// (Ctx::noalloc_scratch); calls on one GPU share it and must not overlap.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr size_t kNoAllocCounterBytes = 256;
constexpr uint32_t kNoAllocCapCut =
    (uint32_t)((kNoAllocScratchBytes - kNoAllocCounterBytes) / (sizeof(CutTri) + sizeof(RasterTri)));
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

template <typename T>
static int dev_alloc(T** p, size_t count, const char* what) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(OCR_ERR_OOM, std::string("hipMalloc failed for ") + what + ": " + hipGetErrorString(e));
  }
  return OCR_OK;
}

template <typename T>
static void dev_free(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// sin/cos of an angle: binary64 evaluation rounded once to binary32.
void sincos_of(const float ang[3], SinCos& sc) {
  for (int i = 0; i < 3; i++) {
    sc.s[i] = (float)sin((double)ang[i]);
    sc.c[i] = (float)cos((double)ang[i]);
  }
}

// The six cube-face rotations r_struct (cl2.cl:2549-2574, :4487-4511); the
// reference builds with -cl-single-precision-constant, so these are binary32.
static void face_table(SinCos out[6]) {
  const float pi = kPi;
  const float ang[6][3] = {{0, 0, 0},          {pi / 2.0f, 0, 0},         {0, pi, 0},
                           {3.0f * pi / 2.0f, 0, 0}, {0, 3.0f * pi / 2.0f, 0}, {0, pi / 2.0f, 0}};
  for (int i = 0; i < 6; i++) sincos_of(ang[i], out[i]);
}

// Depth and id pass blocks per CU for a single view (each block loops over
// sample chunks): 16 keeps the passes supplied while they share the GPU with
// the previous render's shading (8: -2.3% frames/s, 4: -5% on the 1080p,
// 1 M-triangle frame).
constexpr int kRasterBlocksPerCu = 16;

static ViewParams make_view_params(const ocr_view* v) {
  ViewParams p;
  memset(&p, 0, sizeof(p));
  p.width = v->width;
  p.height = v->height;
  p.fov = v->fov_const;
  p.ifov = 1.0f / v->fov_const;
  p.linear = v->use_linear_rendering;
  for (int i = 0; i < 3; i++) p.cpos[i] = v->c_pos[i];
  sincos_of(v->c_rot, p.cam);
  for (int i = 0; i < 4; i++) p.clear[i] = v->clear_col[i];
  return p;
}

static int g_cu_count = 0;

int device_cu_count() { return g_cu_count > 0 ? g_cu_count : 256; }

DevScene dev_scene(const Scene* sc) {
  DevScene d;
  memset(&d, 0, sizeof(d));
  d.tri_pos = sc->tri_pos;
  d.world = sc->world;
  d.tri_attr = sc->tri_attr;
  d.objs = sc->objs;
  d.pages = sc->atlas.pages;
  d.nums = sc->atlas.nums;
  d.sizes = sc->atlas.sizes;
  d.lights = sc->lights;
  d.light_lin = sc->light_lin;
  d.shadow_maps = sc->shadow_maps;
  d.static_maps = sc->static_maps;
  d.ntri = sc->ntri;
  d.nobj = sc->nobj;
  d.mip_start = sc->atlas.mip_start;
  d.nlights = sc->nlights;
  d.L = sc->lbuf_dim;
  return d;
}

DevView dev_view(const ocr_view* v, uint32_t scene, Frame* f) {
  DevView d;
  memset(&d, 0, sizeof(d));
  d.v = make_view_params(v);
  d.scene = scene;
  d.depth = f->depth;
  d.idkey = f->idkey;
  d.rgba = f->rgba;
  d.normals = f->normals;
  return d;
}

static uint32_t round256(uint64_t n) { return (uint32_t)((n + 255) / 256 * 256); }

// Pass arguments of a single view rendered into its own frame.
static PassArgs single_pass(const Scene* s, const ocr_view* view, Frame* f) {
  PassArgs a;
  memset(&a, 0, sizeof(a));
  a.view0 = dev_view(view, 0, f);
  a.scene0 = dev_scene(s);
  a.nviews = 1;
  a.tri_slots = round256(s->ntri);
  a.pix_slots = round256((uint64_t)view->width * view->height);
  a.tile_w_log2 = 4;
  a.unshadowed = s->n_shadow == 0 ? 1u : 0u;
  a.tiles = view_tiles(view->width, view->height, a.tile_w_log2);
  const Frame::OutBuf& ob = f->out[f->cur];
  a.lists = make_lists(ob.cuts, ob.counters, ob.cap_cut);
  a.lists.ent = ob.ent;
  if (f->raster_mode == OCR_RASTER_TILES && ob.tb.bin) {
    a.lists.tb = ob.tb;
    a.view0.bin_begin = 0;
    a.view0.bin_tiles_x = bin_tiles_x(view->width, ob.tb.tw_log2);
  }
  return a;
}

void free_bins(TileBins& tb, uint32_t& box_cap) {
  dev_free(tb.cnt);
  dev_free(tb.off);
  dev_free(tb.cur);
  dev_free(tb.sums);
  dev_free(tb.bin);
  dev_free(tb.box);
  tb.ntiles = tb.cap_bin = 0;
  box_cap = 0;
}

int ensure_bins(TileBins& tb, uint32_t& box_cap, uint32_t cap_cut, uint32_t ntiles, uint32_t twl, uint32_t thl,
                uint32_t cap_bin, bool exact) {
  const bool tiles_ok = tb.cnt && tb.ntiles == ntiles;
  const bool bins_ok = tb.bin && (exact ? tb.cap_bin == cap_bin : tb.cap_bin >= cap_bin);
  const bool box_ok = tb.box && box_cap >= cap_cut;
  tb.tw_log2 = twl;
  tb.th_log2 = thl;
  if (tiles_ok && bins_ok && box_ok) return OCR_OK;
  OCR_HIP_CHECK(hipDeviceSynchronize());
  int rc;
  if (!tiles_ok) {
    dev_free(tb.cnt);
    dev_free(tb.off);
    dev_free(tb.cur);
    dev_free(tb.sums);
      tb.ntiles = 0;
    if ((rc = dev_alloc(&tb.cnt, ntiles, "tile counts")) || (rc = dev_alloc(&tb.off, (size_t)ntiles + 1, "tile offsets")) ||
        (rc = dev_alloc(&tb.cur, ntiles, "tile cursors")) ||
        (rc = dev_alloc(&tb.sums, 2 * (size_t)bin_scan_chunks(ntiles), "tile scan sums")))
      return rc;
    OCR_HIP_CHECK(hipMemset(tb.cnt, 0, 4 * (size_t)ntiles));
    tb.ntiles = ntiles;
  }
  if (!bins_ok) {
    dev_free(tb.bin);
    tb.cap_bin = 0;
    if ((rc = dev_alloc(&tb.bin, cap_bin, "tile bins"))) return rc;
    tb.cap_bin = cap_bin;
  }
  if (!box_ok) {
    dev_free(tb.box);
    box_cap = 0;
    if ((rc = dev_alloc(&tb.box, cap_cut, "tile boxes"))) return rc;
    box_cap = cap_cut;
  }
  return OCR_OK;
}

// Bins of output set b for the frame's tile size: by default two entries
// per cut-list slot; an overflow raises want_bin for the renders after it.
static int ensure_frame_bins(Frame* f, int b) {
  Frame::OutBuf& o = f->out[b];
  const uint32_t ntiles = bin_tiles(f->width, f->height, f->tw_log2, f->th_log2);
  const uint64_t dflt = std::max<uint64_t>(2ull * o.cap_cut, 4096ull);
  const uint32_t cap = (uint32_t)std::min<uint64_t>(f->want_bin ? (uint64_t)f->want_bin : dflt, 0xFFFFFFF0ull);
  return ensure_bins(o.tb, o.box_cap, o.cap_cut, ntiles, f->tw_log2, f->th_log2, cap, f->want_bin != 0);
}

// Covered-sample entries of a frame's output set: room for the chunks of
// 10 raster samples per pixel (the 1080p, 1 M-triangle frame has 8.2; a
// render with more sweeps its later chunks in the id pass), 40 bytes per
// pixel.
constexpr uint32_t kEntSamplesPerPixel = 10;

int frame_alloc_set(Frame* f, int b) {
  Frame::OutBuf& o = f->out[b];
  if (o.depth) return OCR_OK;
  const size_t npix = (size_t)f->width * f->height;
  const uint32_t chunks = (uint32_t)((npix * kEntSamplesPerPixel + kChunk - 1) / kChunk);
  int rc;
  if ((rc = dev_alloc(&o.depth, npix, "depth")) || (rc = dev_alloc(&o.idkey, npix, "id buffer")) ||
      (rc = dev_alloc(&o.rgba, npix, "rgba")) || (rc = dev_alloc(&o.normals, npix, "normals")) ||
      (rc = dev_alloc(&o.counters, kCounterWords, "counters")) ||
      (rc = dev_alloc(&o.ent.ent, (size_t)chunks * kEntCap, "covered-sample entries")) ||
      (rc = dev_alloc(&o.ent.cnt, chunks, "covered-sample counts")))
    return rc;
  o.ent.chunks = chunks;
  OCR_HIP_CHECK(hipMemset(o.counters, 0, 4 * kCounterWords));
  return OCR_OK;
}

}  // namespace ocr

using namespace ocr;


extern "C" {

const char* ocr_last_error(void) { return g_last_error.c_str(); }

int ocr_create(int device, ocr_ctx** out) {
  if (!out) return fail(OCR_ERR_INVALID, "ocr_create: null out");
  *out = nullptr;
  int n = 0;
  OCR_HIP_CHECK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(OCR_ERR_INVALID, "ocr_create: no such device");
  OCR_HIP_CHECK(hipSetDevice(device));
  ocr_ctx* c = new (std::nothrow) ocr_ctx();
  if (!c) return fail(OCR_ERR_OOM, "ocr_create: host allocation");
  c->device = device;
  hipError_t e = hipGetDeviceProperties(&c->prop, device);
  if (e != hipSuccess) {
    delete c;
    return fail(OCR_ERR_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
  }
  face_table(c->face_sc);
  g_cu_count = c->prop.multiProcessorCount;
  e = hipGetSymbolAddress((void**)&c->noalloc_scratch, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    delete c;
    return fail(OCR_ERR_HIP, std::string("ocr_create: scratch address: ") + hipGetErrorString(e));
  }
  *out = c;
  return OCR_OK;
}

void ocr_destroy(ocr_ctx* ctx) { delete ctx; }

int ocr_query_free_hbm(ocr_ctx* ctx, size_t* free_bytes, size_t* total_bytes) {
  if (!ctx) return fail(OCR_ERR_INVALID, "ocr_query_free_hbm: null ctx");
  OCR_HIP_CHECK(hipSetDevice(ctx->device));
  size_t f = 0, t = 0;
  OCR_HIP_CHECK(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return OCR_OK;
}

// ----------------------------------------------------------------- scene
// The quaternion rot_quat rotates by (cl2.cl:350-357): q scaled by the
// reciprocal of its correctly rounded length, products and sums left to
// right in binary32 (host IEEE arithmetic, no contraction), exactly what the
// kernels' rot_quat forms from q.
static void normalised_quat(const float q[4], float out[4]) {
  const float d = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  const float inv = 1.0f / sqrtf(d);
  for (int i = 0; i < 4; i++) out[i] = q[i] * inv;
}

static void pack_obj(const ocr_obj_desc& s, ObjDesc& d) {
  memset(&d, 0, sizeof(d));
  for (int i = 0; i < 3; i++) d.wp[i] = s.world_pos[i];
  d.scale = s.scale;
  for (int i = 0; i < 4; i++) d.q[i] = s.world_rot_quat[i];
  d.tid = s.tid;
  d.specular = s.specular;
  d.spec_mult = s.spec_mult;
  d.diffuse = s.diffuse;
  d.flags = s.feature_flag;
  // per-object terms of kernel3 (cl2.cl:350-370, :6232-6240), formed once
  normalised_quat(d.q, d.qn);
  const float c[4] = {-d.q[0], -d.q[1], -d.q[2], d.q[3]};
  const float len_sq = c[0] * c[0] + c[1] * c[1] + c[2] * c[2] + c[3] * c[3];
  const float cb[4] = {c[0] / len_sq, c[1] / len_sq, c[2] / len_sq, c[3] / len_sq};
  normalised_quat(cb, d.qb);
  const float rough = fminf(fmaxf(1.f - d.specular, 0.001f), 10.f);
  d.rough2 = rough * rough;            // in [1e-6, 100]
  d.inv_rough2 = 1.0f / d.rough2;
}

int ocr_scene_create(ocr_ctx* ctx, const ocr_triangle* tris, uint32_t ntri, const ocr_obj_desc* objs,
                     uint32_t nobj, ocr_scene** out) {
  if (!ctx || !out || (ntri && !tris) || (nobj && !objs))
    return fail(OCR_ERR_INVALID, "ocr_scene_create: bad argument");
  if (ntri >= 0x7FFFFFFFu) return fail(OCR_ERR_UNSUPPORTED, "ocr_scene_create: too many triangles");
  *out = nullptr;
  OCR_HIP_CHECK(hipSetDevice(ctx->device));
  ocr_scene* s = new (std::nothrow) ocr_scene();
  if (!s) return fail(OCR_ERR_OOM, "ocr_scene_create: host allocation");
  s->ctx = ctx;
  s->ntri = ntri;
  s->nobj = nobj;
  // Split the 144-byte AoS triangle into the 40-byte position record read by
  // every raster pass and the 72-byte attribute record read only by shading.
  std::vector<TriPos> pos(ntri);
  std::vector<TriAttr> attr(ntri);
  for (uint32_t t = 0; t < ntri; t++) {
    const ocr_triangle& T = tris[t];
    for (int v = 0; v < 3; v++) {
      for (int k = 0; k < 3; k++) {
        pos[t].p[3 * v + k] = T.vertices[v].pos[k];
        attr[t].n[3 * v + k] = T.vertices[v].normal[k];
      }
      attr[t].vt[2 * v] = T.vertices[v].vt[0];
      attr[t].vt[2 * v + 1] = T.vertices[v].vt[1];
      attr[t].col[v] = T.vertices[v].vertex_col;
    }
    pos[t].oid = T.vertices[0].object_id;
  }
  std::vector<ObjDesc> od(nobj);
  std::vector<float> motion((size_t)nobj * 16);
  for (uint32_t o = 0; o < nobj; o++) {
    pack_obj(objs[o], od[o]);
    s->max_tid = std::max(s->max_tid, objs[o].tid);
    memcpy(&motion[16 * (size_t)o], objs[o].old_world_pos_1, 16);
    memcpy(&motion[16 * (size_t)o + 4], objs[o].old_world_pos_2, 16);
    memcpy(&motion[16 * (size_t)o + 8], objs[o].old_world_rot_quat_1, 16);
    memcpy(&motion[16 * (size_t)o + 12], objs[o].old_world_rot_quat_2, 16);
  }
  int rc;
  if ((rc = dev_alloc(&s->tri_pos, ntri, "triangles")) || (rc = dev_alloc(&s->world, ntri, "world-space triangles")) ||
      (rc = dev_alloc(&s->tri_attr, ntri, "attributes")) ||
      (rc = dev_alloc(&s->objs, nobj, "descriptors")) || (rc = dev_alloc(&s->motion, (size_t)nobj * 16, "motion poses")) ||
      (rc = dev_alloc(&s->sh_counters, kShadowCounterWords, "counters"))) {
    ocr_scene_destroy(s);
    return rc;
  }
  s->objs_v[0] = s->objs;
  s->world_v[0] = s->world;
  hipError_t e = hipSuccess;
  if (ntri) e = hipMemcpy(s->tri_pos, pos.data(), sizeof(TriPos) * ntri, hipMemcpyHostToDevice);
  if (e == hipSuccess && ntri) e = hipMemcpy(s->tri_attr, attr.data(), sizeof(TriAttr) * ntri, hipMemcpyHostToDevice);
  if (e == hipSuccess && nobj) e = hipMemcpy(s->objs, od.data(), sizeof(ObjDesc) * nobj, hipMemcpyHostToDevice);
  if (e == hipSuccess && nobj) e = hipMemcpy(s->motion, motion.data(), 4 * motion.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_world(s, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipEventCreateWithFlags(&s->written, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&s->geom_written, hipEventDisableTiming);
  if (e != hipSuccess) {
    ocr_scene_destroy(s);
    return fail(OCR_ERR_HIP, std::string("ocr_scene_create: ") + hipGetErrorString(e));
  }
  *out = s;
  return OCR_OK;
}

int ocr_scene_update_objs(ocr_scene* s, uint32_t first, uint32_t n, const ocr_obj_desc* objs) {
  if (!s || (n && !objs) || first > s->nobj || n > s->nobj - first)
    return fail(OCR_ERR_INVALID, "ocr_scene_update_objs: range outside the scene");
  if (!n) return OCR_OK;
  std::vector<ObjDesc> od(n);
  for (uint32_t o = 0; o < n; o++) {
    pack_obj(objs[o], od[o]);
    s->max_tid = std::max(s->max_tid, objs[o].tid);
  }
  OCR_HIP_CHECK(hipSetDevice(s->ctx->device));
  std::lock_guard<std::mutex> lk(s->ver_mu);
  // renders run on streams of their frames: let those reading the scene end
  OCR_HIP_CHECK(hipDeviceSynchronize());
  OCR_HIP_CHECK(hipMemcpy(s->objs + first, od.data(), sizeof(ObjDesc) * n, hipMemcpyHostToDevice));
  // the world-space copy follows the objects before any later render
  OCR_HIP_CHECK(launch_world(s, nullptr));
  OCR_HIP_CHECK(hipStreamSynchronize(nullptr));
  return OCR_OK;
}

int ocr_scene_update_objs_async(ocr_scene* s, uint32_t first, uint32_t n, const ocr_obj_desc* objs, void* stream) {
  if (!s || (n && !objs) || first > s->nobj || n > s->nobj - first)
    return fail(OCR_ERR_INVALID, "ocr_scene_update_objs_async: range outside the scene");
  if (!n) return OCR_OK;
  OCR_HIP_CHECK(hipSetDevice(s->ctx->device));
  std::lock_guard<std::mutex> lk(s->ver_mu);
  const int nv = 1 - s->ver;   // the version this update builds
  if (!s->objs_v[nv]) {
    int rc;
    if ((rc = dev_alloc(&s->objs_v[nv], s->nobj, "descriptors (second version)")) ||
        (rc = dev_alloc(&s->world_v[nv], s->ntri, "world-space triangles (second version)"))) {
      dev_free(s->objs_v[nv]);
      dev_free(s->world_v[nv]);
      return rc;
    }
  }
  if (!s->upd_stream) OCR_HIP_CHECK(hipStreamCreateWithFlags(&s->upd_stream, hipStreamNonBlocking));
  hipStream_t us = s->upd_stream;
  // the passes that read version nv when it was current end first
  OCR_HIP_CHECK(s->wait_readers(s->readers[nv], us));
  const int k = s->upd_next;
  s->upd_next = (k + 1) % Scene::kUpdRing;
  if (!s->upd_host[k]) {
    hipError_t e = hipHostMalloc((void**)&s->upd_host[k], sizeof(ObjDesc) * (size_t)s->nobj);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->upd_done[k], hipEventDisableTiming);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(OCR_ERR_OOM, std::string("ocr_scene_update_objs_async: pinned staging: ") + hipGetErrorString(e));
    }
  } else if (s->upd_used[k]) {
    OCR_HIP_CHECK(hipEventSynchronize(s->upd_done[k]));   // the slot's last copy has read it
  }
  for (uint32_t o = 0; o < n; o++) {
    pack_obj(objs[o], s->upd_host[k][o]);
    s->max_tid = std::max(s->max_tid, objs[o].tid);
  }
  hipStream_t st = (hipStream_t)stream;
  // Work enqueued on `stream` before this call that writes what the update
  // reads (nothing of the scene's own) is not waited for: the update reads
  // only the current version and the staged records.  Version nv starts as
  // a copy of the current one (unless every record is replaced), then takes
  // the new records.
  if (first > 0 || n < s->nobj)
    OCR_HIP_CHECK(hipMemcpyAsync(s->objs_v[nv], s->objs_v[s->ver], sizeof(ObjDesc) * s->nobj,
                                 hipMemcpyDeviceToDevice, us));
  OCR_HIP_CHECK(hipMemcpyAsync(s->objs_v[nv] + first, s->upd_host[k], sizeof(ObjDesc) * n, hipMemcpyHostToDevice, us));
  OCR_HIP_CHECK(hipEventRecord(s->upd_done[k], us));
  s->upd_used[k] = true;
  s->ver = nv;
  s->objs = s->objs_v[nv];
  s->world = s->world_v[nv];
  s->gen++;
  OCR_HIP_CHECK(launch_world(s, us));
  OCR_HIP_CHECK(hipEventRecord(s->geom_written, us));
  s->geom_valid = true;
  // later work on the caller's stream sees the update
  OCR_HIP_CHECK(hipStreamWaitEvent(st, s->geom_written, 0));
  return OCR_OK;
}

static void free_atlas(Atlas& a) {
  dev_free(a.pages);
  dev_free(a.nums);
  dev_free(a.sizes);
  a.npages = a.nnums = a.nsizes = a.mip_start = 0;
}

void ocr_scene_destroy(ocr_scene* s) {
  if (!s) return;
  (void)hipSetDevice(s->ctx->device);
  dev_free(s->tri_pos);
  if (!s->world_v[0]) s->world_v[0] = s->world;   // creation failed before the versions were set
  if (!s->objs_v[0]) s->objs_v[0] = s->objs;
  for (int v = 0; v < 2; v++) {
    dev_free(s->world_v[v]);
    dev_free(s->objs_v[v]);
  }
  s->world = nullptr;
  s->objs = nullptr;
  dev_free(s->tri_attr);
  dev_free(s->motion);
  free_atlas(s->atlas);
  dev_free(s->lights);
  dev_free(s->light_lin);
  if (!s->maps_v[0] && !s->maps_v[1]) s->maps_v[0] = s->shadow_maps;
  for (int v = 0; v < 2; v++) dev_free(s->maps_v[v]);
  s->shadow_maps = nullptr;
  dev_free(s->static_maps);
  dev_free(s->sh_rast);
  dev_free(s->sh_counters);
  if (s->written) (void)hipEventDestroy(s->written);
  if (s->geom_written) (void)hipEventDestroy(s->geom_written);
  for (int k = 0; k < Scene::kUpdRing; k++) {
    if (s->upd_done[k]) {
      (void)hipEventSynchronize(s->upd_done[k]);
      (void)hipEventDestroy(s->upd_done[k]);
    }
    if (s->upd_host[k]) (void)hipHostFree(s->upd_host[k]);
  }
  for (int v = 0; v < 2; v++) {
    for (const Scene::Reader& r : s->readers[v]) (void)hipEventDestroy(r.ev);
    for (const Scene::Reader& r : s->map_readers[v]) (void)hipEventDestroy(r.ev);
  }
  for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
  if (s->upd_stream) (void)hipStreamDestroy(s->upd_stream);
  if (s->sh_stream) (void)hipStreamDestroy(s->sh_stream);
  delete s;
}

// --------------------------------------------------------------- atlas
static uint32_t pow2_ceil(uint32_t v) {
  uint32_t n = 1;
  while (n < v) n <<= 1;
  return n;
}

int ocr_atlas_upload(ocr_scene* s, const ocr_tex_src* texs, uint32_t ntex) {
  if (!s || (ntex && !texs)) return fail(OCR_ERR_INVALID, "ocr_atlas_upload: bad argument");
  OCR_HIP_CHECK(hipSetDevice(s->ctx->device));
  OCR_HIP_CHECK(hipDeviceSynchronize());   // renders still reading the atlas end first
  {
    std::lock_guard<std::mutex> lk(s->ver_mu);
    s->gen++;   // the atlas buffers are replaced
  }
  // texture_context.cpp:94-165: count slots per power-of-two size class
  // (base + 4 mips per texture) and cut them into 2048^2 pages, smallest
  // class first.
  std::vector<uint32_t> base(ntex);
  std::map<uint32_t, uint32_t> per_size;
  for (uint32_t t = 0; t < ntex; t++) {
    if (!texs[t].rgba || texs[t].width == 0 || texs[t].height == 0)
      return fail(OCR_ERR_INVALID, "ocr_atlas_upload: empty texture");
    base[t] = pow2_ceil(std::max(texs[t].width, texs[t].height));
    if (base[t] < 16 || base[t] > OCR_ATLAS_DIM)
      return fail(OCR_ERR_UNSUPPORTED, "ocr_atlas_upload: texture size must round to 16..2048");
    for (int j = 0; j <= OCR_MIP_LEVELS; j++) per_size[base[t] >> j]++;
  }
  struct Page { uint32_t size; uint32_t free; };
  std::vector<Page> pages;
  for (auto& kv : per_size) {
    uint32_t per_page = (OCR_ATLAS_DIM / kv.first) * (OCR_ATLAS_DIM / kv.first);
    uint32_t rem = kv.second;
    while (rem >= per_page) { pages.push_back({kv.first, per_page}); rem -= per_page; }
    if (rem) pages.push_back({kv.first, rem});
  }
  // texture_context.cpp:167-250: each texture, then each texture's mips,
  // takes the last free slot of the first page of its size.
  std::vector<uint32_t> nums(std::max<uint32_t>(ntex * (1 + OCR_MIP_LEVELS), 1 + OCR_MIP_LEVELS), 0);
  auto take = [&](uint32_t size, uint32_t& desc) -> bool {
    for (size_t p = 0; p < pages.size(); p++)
      if (pages[p].size == size && pages[p].free > 0) {
        pages[p].free--;
        desc = ((uint32_t)p << 16) | pages[p].free;
        return true;
      }
    return false;
  };
  for (uint32_t t = 0; t < ntex; t++)
    if (!take(base[t], nums[t])) return fail(OCR_ERR_INVALID, "ocr_atlas_upload: page plan");
  for (uint32_t t = 0; t < ntex; t++)
    for (int j = 0; j < OCR_MIP_LEVELS; j++)
      if (!take(base[t] >> (j + 1), nums[ntex + t * OCR_MIP_LEVELS + j]))
        return fail(OCR_ERR_INVALID, "ocr_atlas_upload: mip page plan");
  std::vector<uint32_t> sizes(std::max<size_t>(pages.size(), 2), 16);
  for (size_t p = 0; p < pages.size(); p++) sizes[p] = pages[p].size;

  free_atlas(s->atlas);
  Atlas& a = s->atlas;
  a.npages = (uint32_t)std::max<size_t>(pages.size(), 1);
  a.nnums = (uint32_t)nums.size();
  a.nsizes = (uint32_t)sizes.size();
  a.mip_start = ntex;
  int rc;
  size_t page_bytes = (size_t)OCR_ATLAS_DIM * OCR_ATLAS_DIM * 4;
  if ((rc = dev_alloc(&a.pages, page_bytes * a.npages, "atlas")) || (rc = dev_alloc(&a.nums, a.nnums, "nums")) ||
      (rc = dev_alloc(&a.sizes, a.nsizes, "sizes"))) {
    free_atlas(a);
    return rc;
  }
  OCR_HIP_CHECK(hipMemset(a.pages, 0, page_bytes * a.npages));
  OCR_HIP_CHECK(hipMemcpy(a.nums, nums.data(), sizeof(uint32_t) * a.nnums, hipMemcpyHostToDevice));
  OCR_HIP_CHECK(hipMemcpy(a.sizes, sizes.data(), sizeof(uint32_t) * a.nsizes, hipMemcpyHostToDevice));
  // texture::update_me_to_gpu + update_gpu_mipmaps (texture.cpp:323-493).
  size_t max_bytes = 0;
  for (uint32_t t = 0; t < ntex; t++) max_bytes = std::max(max_bytes, (size_t)texs[t].width * texs[t].height * 4);
  uint8_t* staging = nullptr;
  if (ntex && (rc = dev_alloc(&staging, max_bytes, "texture staging"))) return rc;
  hipStream_t st = nullptr;
  int status = OCR_OK;
  for (uint32_t t = 0; t < ntex && status == OCR_OK; t++) {
    const ocr_tex_src& src = texs[t];
    hipError_t e = hipMemcpy(staging, src.rgba, (size_t)src.width * src.height * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_atlas_upload(a, t, staging, src.width, src.height, st);
    if (e == hipSuccess) e = launch_mip(a, t, t * OCR_MIP_LEVELS + ntex, src.width, src.height, st);
    for (uint32_t lvl = 0; lvl + 1 < OCR_MIP_LEVELS && e == hipSuccess; lvl++) {
      uint32_t proper = t * OCR_MIP_LEVELS + ntex + lvl;
      e = launch_mip(a, proper, proper + 1, src.width, src.height, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) status = fail(OCR_ERR_HIP, std::string("ocr_atlas_upload: ") + hipGetErrorString(e));
  }
  dev_free(staging);
  return status;
}

int ocr_scene_atlas_info(ocr_scene* s, uint32_t* npages, uint32_t* nnums, uint32_t* nsizes, uint32_t* mip_start) {
  if (!s) return fail(OCR_ERR_INVALID, "ocr_scene_atlas_info: null scene");
  if (npages) *npages = s->atlas.npages;
  if (nnums) *nnums = s->atlas.nnums;
  if (nsizes) *nsizes = s->atlas.nsizes;
  if (mip_start) *mip_start = s->atlas.mip_start;
  return OCR_OK;
}

int ocr_scene_read_atlas(ocr_scene* s, uint8_t* pages, uint32_t* nums, uint32_t* sizes) {
  if (!s || !s->atlas.pages) return fail(OCR_ERR_INVALID, "ocr_scene_read_atlas: no atlas");
  OCR_HIP_CHECK(hipSetDevice(s->ctx->device));
  OCR_HIP_CHECK(hipDeviceSynchronize());
  size_t page_bytes = (size_t)OCR_ATLAS_DIM * OCR_ATLAS_DIM * 4;
  if (pages) OCR_HIP_CHECK(hipMemcpy(pages, s->atlas.pages, page_bytes * s->atlas.npages, hipMemcpyDeviceToHost));
  if (nums) OCR_HIP_CHECK(hipMemcpy(nums, s->atlas.nums, 4 * s->atlas.nnums, hipMemcpyDeviceToHost));
  if (sizes) OCR_HIP_CHECK(hipMemcpy(sizes, s->atlas.sizes, 4 * s->atlas.nsizes, hipMemcpyDeviceToHost));
  return OCR_OK;
}

// ---------------------------------------------------------------- lights
// A light upload that failed after the old buffers were freed: the scene is
// left with no lights and no light buffers (passes then refuse it until a
// successful upload), and the error code is passed through.
static int lights_failed(ocr_scene* s, int rc) {
  dev_free(s->lights);
  dev_free(s->light_lin);
  for (int v = 0; v < 2; v++) dev_free(s->maps_v[v]);
  s->shadow_maps = nullptr;
  dev_free(s->static_maps);
  s->nlights = s->n_shadow = s->n_static = 0;
  s->host_lights.clear();
  s->static_dirty = false;
  return rc;
}

int ocr_lights_upload(ocr_scene* s, const ocr_light* lights, uint32_t n, uint32_t light_buffer_dim) {
  if (!s || (n && !lights)) return fail(OCR_ERR_INVALID, "ocr_lights_upload: bad argument");
  if (light_buffer_dim < 8 || (uint64_t)light_buffer_dim * light_buffer_dim > kMaxTargetPixels)
    return fail(OCR_ERR_UNSUPPORTED, "ocr_lights_upload: light buffer dimension must be 8..4000");
  // every light is checked before anything of the scene is replaced
  for (uint32_t i = 0; i < n; i++)
    if (lights[i].shadow > 1)
      return fail(OCR_ERR_UNSUPPORTED, "ocr_lights_upload: light.shadow must be 0 or 1");
  OCR_HIP_CHECK(hipSetDevice(s->ctx->device));
  OCR_HIP_CHECK(hipDeviceSynchronize());   // renders still reading the lights end first
  {
    std::lock_guard<std::mutex> lk(s->ver_mu);
    s->gen++;   // the light and shadow-map buffers are replaced
  }
  // From here on a failure leaves the scene with no lights (nlights = 0, no
  // buffers), which renders and batches refuse or render unlit, never with
  // counts that point past freed buffers.
  s->nlights = s->n_shadow = s->n_static = 0;
  s->host_lights.clear();
  dev_free(s->lights);
  dev_free(s->light_lin);
  for (int v = 0; v < 2; v++) {
    dev_free(s->maps_v[v]);
    for (const Scene::Reader& r : s->map_readers[v]) s->ev_pool.push_back(r.ev);   // all complete
    s->map_readers[v].clear();
  }
  s->shadow_maps = nullptr;
  s->mver = 0;
  dev_free(s->static_maps);
  uint32_t n_shadow = 0, n_static = 0;
  for (uint32_t i = 0; i < n; i++) {
    if (lights[i].shadow == 1) n_shadow++;
    if (lights[i].shadow && lights[i].is_static) n_static++;
  }
  s->lbuf_dim = light_buffer_dim;
  const size_t per = (size_t)6 * light_buffer_dim * light_buffer_dim;
  int rc;
  if ((rc = dev_alloc(&s->lights, n, "lights")) || (rc = dev_alloc(&s->light_lin, (size_t)4 * n, "light colours")) ||
      (rc = dev_alloc(&s->maps_v[0], per * std::max<uint32_t>(n_shadow, 1), "shadow maps")) ||
      (rc = dev_alloc(&s->static_maps, per * std::max<uint32_t>(n_static, 1), "static shadow maps")))
    return lights_failed(s, rc);
  s->shadow_maps = s->maps_v[0];
  hipError_t e = hipSuccess;
  if (n) e = hipMemcpy(s->lights, lights, sizeof(ocr_light) * n, hipMemcpyHostToDevice);
  if (n && e == hipSuccess) e = launch_light_prep(s->lights, n, s->light_lin, nullptr);
  if (e == hipSuccess) e = launch_fill_u32(s->shadow_maps, 0xFFFFFFFFu, per * std::max<uint32_t>(n_shadow, 1), nullptr);
  if (e == hipSuccess) e = launch_fill_u32(s->static_maps, 0xFFFFFFFFu, per * std::max<uint32_t>(n_static, 1), nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess)
    return lights_failed(s, fail(OCR_ERR_HIP, std::string("ocr_lights_upload: ") + hipGetErrorString(e)));
  // the counts become visible only with every buffer in place
  s->host_lights.assign(lights, lights + n);
  s->nlights = n;
  s->n_shadow = n_shadow;
  s->n_static = n_static;
  s->static_dirty = true;
  return OCR_OK;
}

static int ensure_shadow_lists(ocr_scene* s, uint32_t cap_cut) {
  if (s->sh_rast && s->sh_cap_cut >= cap_cut) return OCR_OK;
  dev_free(s->sh_rast);
  s->sh_cap_cut = 0;
  int rc = dev_alloc(&s->sh_rast, (size_t)cap_cut, "shadow work list");
  if (rc) return rc;
  s->sh_cap_cut = cap_cut;
  return OCR_OK;
}

// Lights per shadow prearrange launch.  The launch rasterises small pieces
// itself and lists the large ones while the list has room (then rasterises
// the rest too), so any list size is correct; the list holds up to 6
// records per triangle per light, at most kShadowMaxRecords (896 MB).
constexpr uint32_t kShadowGroup = 16;
static uint32_t shadow_list_cap(uint32_t ntri) {
  return (uint32_t)std::min<uint64_t>(((uint64_t)ntri * 6 + 64) * kShadowGroup, kShadowMaxRecords);
}

// Renders the maps of `n` lights (positions, map slots) in groups.
static int shadow_groups(ocr_scene* s, const std::vector<const ocr_light*>& ls, const std::vector<uint32_t>& slots,
                         int only_static, uint32_t* maps, hipStream_t st) {
  const uint32_t g = kShadowGroup;
  for (size_t b = 0; b < ls.size(); b += g) {
    const uint32_t n = (uint32_t)std::min<size_t>(g, ls.size() - b);
    float pos[16][3];
    for (uint32_t i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) pos[i][k] = ls[b + i]->pos[k];
    OCR_HIP_CHECK(launch_shadow_group(s, pos, &slots[b], n, only_static, maps, st));
  }
  return OCR_OK;
}

int ocr_shadow_pass(ocr_scene* s, void* stream) {
  if (!s) return fail(OCR_ERR_INVALID, "ocr_shadow_pass: null scene");
  if (s->n_shadow == 0 && !(s->static_dirty && s->n_static)) return OCR_OK;
  OCR_HIP_CHECK(hipSetDevice(s->ctx->device));
  std::lock_guard<std::mutex> lk(s->ver_mu);   // reads the current geometry version
  const size_t per = (size_t)6 * s->lbuf_dim * s->lbuf_dim;
  // the map version this pass renders: the current one when no render has
  // been issued against it (renders register under this lock), else the
  // other one, on the scene's own stream once its readers have finished
  const int nm = s->map_readers[s->mver].empty() ? s->mver : 1 - s->mver;
  if (!s->maps_v[nm]) {
    int rc = dev_alloc(&s->maps_v[nm], per * std::max<uint32_t>(s->n_shadow, 1), "shadow maps (second version)");
    if (rc) return rc;
  }
  if (!s->sh_stream) OCR_HIP_CHECK(hipStreamCreateWithFlags(&s->sh_stream, hipStreamNonBlocking));
  const hipStream_t caller = (hipStream_t)stream;
  const hipStream_t st = s->sh_stream;
  OCR_HIP_CHECK(s->wait_readers(s->map_readers[nm], st));
  // static maps have one version: when they are rendered, every pass that
  // may read them ends first
  if (s->static_dirty && s->n_static) OCR_HIP_CHECK(s->wait_readers(s->map_readers[s->mver], st));
  OCR_HIP_CHECK(s->wait_geometry(st));   // descriptor updates enqueued on another stream
  {
    int rc = ensure_shadow_lists(s, shadow_list_cap(s->ntri));
    if (rc) return rc;
  }
  uint32_t* const maps = s->maps_v[nm];
  // engine.cpp:1613-1625: clear every dynamic map, then render each
  // shadow-casting light from the dynamic geometry; static lights' maps are
  // rendered from static geometry when the lights were (re)uploaded.  Lights
  // are rendered in groups, one prearrange and one raster launch per group.
  OCR_HIP_CHECK(launch_fill_u32(maps, 0xFFFFFFFFu, per * std::max<uint32_t>(s->n_shadow, 1), st));
  std::vector<const ocr_light*> dyn, stat;
  std::vector<uint32_t> dyn_slot, stat_slot;
  uint32_t nn = 0, kk = 0;
  for (uint32_t i = 0; i < s->nlights; i++) {
    const ocr_light& l = s->host_lights[i];
    if (l.shadow == 1) {
      dyn.push_back(&l);
      dyn_slot.push_back(nn++);
    }
    if (l.shadow && l.is_static) {
      if (s->static_dirty) {
        stat.push_back(&l);
        stat_slot.push_back(kk);
      }
      kk++;
    }
  }
  int rc = shadow_groups(s, dyn, dyn_slot, 0, maps, st);
  if (!rc && !stat.empty()) {
    OCR_HIP_CHECK(launch_fill_u32(s->static_maps, 0xFFFFFFFFu, per * s->n_static, st));
    rc = shadow_groups(s, stat, stat_slot, 1, s->static_maps, st);
  }
  if (rc) return rc;
  s->static_dirty = false;
  OCR_HIP_CHECK(s->note_read(st));
  OCR_HIP_CHECK(hipEventRecord(s->written, st));
  s->written_valid = true;
  s->mver = nm;
  s->shadow_maps = maps;
  s->gen++;   // batches' scene tables point at the new maps
  // later work on the caller's stream sees the maps
  OCR_HIP_CHECK(hipStreamWaitEvent(caller, s->written, 0));
  return OCR_OK;
}

int ocr_scene_read_shadow_maps(ocr_scene* s, uint32_t* dynamic_maps, uint32_t* static_maps) {
  if (!s) return fail(OCR_ERR_INVALID, "ocr_scene_read_shadow_maps: null scene");
  OCR_HIP_CHECK(hipSetDevice(s->ctx->device));
  OCR_HIP_CHECK(hipDeviceSynchronize());
  uint32_t ov = 0;
  OCR_HIP_CHECK(hipMemcpy(&ov, s->sh_counters + kOverflow, 4, hipMemcpyDeviceToHost));
  const size_t per = (size_t)6 * s->lbuf_dim * s->lbuf_dim * 4;
  if (dynamic_maps && s->n_shadow)
    OCR_HIP_CHECK(hipMemcpy(dynamic_maps, s->shadow_maps, per * s->n_shadow, hipMemcpyDeviceToHost));
  if (static_maps && s->n_static)
    OCR_HIP_CHECK(hipMemcpy(static_maps, s->static_maps, per * s->n_static, hipMemcpyDeviceToHost));
  if (ov) return fail(OCR_ERR_UNSUPPORTED, "shadow pass: more raster samples than one pass can hold");
  return OCR_OK;
}

int ocr_scene_read_motion(ocr_scene* s, uint32_t first, uint32_t n, float* poses) {
  if (!s || (n && !poses) || first > s->nobj || n > s->nobj - first)
    return fail(OCR_ERR_INVALID, "ocr_scene_read_motion: range outside the scene");
  if (!n) return OCR_OK;
  OCR_HIP_CHECK(hipSetDevice(s->ctx->device));
  OCR_HIP_CHECK(hipDeviceSynchronize());
  OCR_HIP_CHECK(hipMemcpy(poses, s->motion + 16 * (size_t)first, 64 * (size_t)n, hipMemcpyDeviceToHost));
  return OCR_OK;
}

// ---------------------------------------------------------------- frames
int ocr_frame_create(ocr_ctx* ctx, uint32_t width, uint32_t height, ocr_frame** out) {
  if (!ctx || !out || width < 4 || height < 4) return fail(OCR_ERR_INVALID, "ocr_frame_create: bad argument (frames are at least 4x4)");
  if ((uint64_t)width * height > kMaxTargetPixels || width > 32768 || height > 32768)
    return fail(OCR_ERR_UNSUPPORTED, "ocr_frame_create: frame too large (at most 16,000,000 pixels)");
  *out = nullptr;
  OCR_HIP_CHECK(hipSetDevice(ctx->device));
  ocr_frame* f = new (std::nothrow) ocr_frame();
  if (!f) return fail(OCR_ERR_OOM, "ocr_frame_create: host allocation");
  f->ctx = ctx;
  f->width = width;
  f->height = height;
  int rc = frame_alloc_set(f, 0);
  if (rc) {
    ocr_frame_destroy(f);
    return rc;
  }
  f->select(0);
  hipError_t e = hipSuccess;
  for (int b = 0; b < Frame::kSets && e == hipSuccess; b++) {
    e = hipStreamCreateWithFlags(&f->ss[b], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&f->entry[b], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&f->done[b], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    ocr_frame_destroy(f);
    return fail(OCR_ERR_HIP, std::string("ocr_frame_create: ") + hipGetErrorString(e));
  }
  *out = f;
  return OCR_OK;
}

void ocr_frame_destroy(ocr_frame* f) {
  if (!f) return;
  (void)hipSetDevice(f->ctx->device);
  for (Frame::OutBuf& o : f->out) {
    dev_free(o.depth);
    dev_free(o.idkey);
    dev_free(o.rgba);
    dev_free(o.normals);
    dev_free(o.cuts);
    dev_free(o.counters);
    dev_free(o.ent.ent);
    dev_free(o.ent.cnt);
    free_bins(o.tb, o.box_cap);
  }
  dev_free(f->post_in);
  for (int b = 0; b < Frame::kSets; b++) {
    if (f->ss[b]) (void)hipStreamDestroy(f->ss[b]);
    if (f->entry[b]) (void)hipEventDestroy(f->entry[b]);
    if (f->done[b]) (void)hipEventDestroy(f->done[b]);
  }
  for (hipEvent_t e : f->ev)
    if (e) (void)hipEventDestroy(e);
  delete f;
}

int ocr_frame_set_profiling(ocr_frame* f, int enable) {
  if (!f) return fail(OCR_ERR_INVALID, "ocr_frame_set_profiling: null frame");
  if (enable < 0 || enable > 2) return fail(OCR_ERR_INVALID, "ocr_frame_set_profiling: 0, 1 or 2");
  f->profiling = enable;
  f->n_profiled = 0;
  if (f->profiling && f->ev.empty()) {
    OCR_HIP_CHECK(hipSetDevice(f->ctx->device));
    f->ev.assign((size_t)Frame::kRing * Frame::kEvents, nullptr);
    for (auto& e : f->ev) OCR_HIP_CHECK(hipEventCreate(&e));
  }
  return OCR_OK;
}

int ocr_frame_set_raster(ocr_frame* f, const ocr_raster_config* cfg) {
  if (!f || !cfg) return fail(OCR_ERR_INVALID, "ocr_frame_set_raster: null argument");
  if (cfg->mode != OCR_RASTER_SAMPLES && cfg->mode != OCR_RASTER_TILES)
    return fail(OCR_ERR_INVALID, "ocr_frame_set_raster: unknown mode");
  if (cfg->mode == OCR_RASTER_TILES) {
    const uint32_t a = cfg->tile_w_log2, b = cfg->tile_h_log2;
    if (a < 3 || a > 6 || b < 3 || b > 6 || a + b < 8 || a + b > 11)
      return fail(OCR_ERR_INVALID, "ocr_frame_set_raster: tiles are 2^3..2^6 on a side and 256..2048 pixels");
    f->tw_log2 = a;
    f->th_log2 = b;
    f->want_bin = cfg->bin_capacity;
  }
  f->raster_mode = cfg->mode;
  return OCR_OK;
}

int ocr_frame_get_raster(ocr_frame* f, ocr_raster_config* cfg) {
  if (!f || !cfg) return fail(OCR_ERR_INVALID, "ocr_frame_get_raster: null argument");
  cfg->mode = f->raster_mode;
  cfg->tile_w_log2 = f->tw_log2;
  cfg->tile_h_log2 = f->th_log2;
  cfg->bin_capacity = f->want_bin;
  return OCR_OK;
}

int ocr_frame_reserve(ocr_frame* f, uint32_t max_cut_triangles) {
  if (!f || max_cut_triangles == 0) return fail(OCR_ERR_INVALID, "ocr_frame_reserve: bad argument");
  OCR_HIP_CHECK(hipSetDevice(f->ctx->device));
  OCR_HIP_CHECK(hipDeviceSynchronize());
  for (Frame::OutBuf& o : f->out) {
    dev_free(o.cuts);
    o.cap_cut = 0;
  }
  f->want_cut = max_cut_triangles;
  return OCR_OK;
}

// A cut list holds at most 2 pieces per scene triangle; unless the caller
// reserved a capacity it is sized for that, so it never overflows.
static int ensure_frame_list(ocr_frame* f, int b, uint32_t ntri) {
  Frame::OutBuf& lb = f->out[b];
  uint64_t want = f->want_cut ? (uint64_t)f->want_cut : (uint64_t)ntri * 2 + 64;
  want = std::min<uint64_t>(want, 0xFFFFFFF0ull);
  if (lb.cuts && lb.cap_cut == want) return OCR_OK;
  if (lb.cuts && !f->want_cut && lb.cap_cut >= want) return OCR_OK;
  OCR_HIP_CHECK(hipDeviceSynchronize());
  dev_free(lb.cuts);
  lb.cap_cut = 0;
  int rc = dev_alloc(&lb.cuts, (size_t)want * 2, "cut list");
  if (rc) return rc;
  lb.cap_cut = (uint32_t)want;
  return OCR_OK;
}

int ocr_render(ocr_scene* s, const ocr_view* view, ocr_frame* f, void* stream) {
  if (!s || !view || !f) return fail(OCR_ERR_INVALID, "ocr_render: null argument");
  if (view->width != f->width || view->height != f->height)
    return fail(OCR_ERR_INVALID, "ocr_render: view size differs from the frame");
  if (!(view->fov_const >= 0x1p-60f && view->fov_const <= 0x1p60f))
    return fail(OCR_ERR_INVALID, "ocr_render: fov_const must be in [2^-60, 2^60]");
  if (!s->atlas.pages) return fail(OCR_ERR_INVALID, "ocr_render: scene has no atlas (ocr_atlas_upload)");
  if (s->nobj && s->max_tid >= std::max<uint32_t>(s->atlas.mip_start, 1))
    return fail(OCR_ERR_INVALID, "ocr_render: an object's texture id is outside the atlas");
  if (!s->lights && s->nlights) return fail(OCR_ERR_INVALID, "ocr_render: lights missing");
  if (f->use == Frame::kBatch) return fail(OCR_ERR_INVALID, "ocr_render: the frame is rendered by batches");
  OCR_HIP_CHECK(hipSetDevice(s->ctx->device));
  // this render's buffer set: the one the longest ago
  const int b = (f->cur + 1) % Frame::kSets;
  const bool tiled = f->raster_mode == OCR_RASTER_TILES;
  int rc = frame_alloc_set(f, b);
  if (!rc) rc = ensure_frame_list(f, b, s->ntri);
  if (!rc && tiled) rc = ensure_frame_bins(f, b);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (!s->lights) {
    // a scene with no lights still needs valid (empty) light and map buffers
    rc = ocr_lights_upload(s, nullptr, 0, s->lbuf_dim);
    if (rc) return rc;
  }
  // profiling: events around every pass (mode 1) or the shading pass only
  // (mode 2: two events per render)
  const bool prof_shade = f->profiling && f->n_profiled < Frame::kRing;
  const bool prof = prof_shade && f->profiling == 1;
  hipEvent_t* ev = prof_shade ? &f->ev[(size_t)f->n_profiled * Frame::kEvents] : nullptr;
  hipStream_t sb = f->ss[b];
  // Render N runs on the stream of its set once render N-kSets (same
  // buffers, same stream) and everything enqueued on the caller's stream
  // before render N-kSets+1 was issued (the event recorded at that point,
  // which follows render N-kSets) have completed, so it can overlap the
  // kSets-1 renders before it; the caller's stream then waits for it.  The
  // first renders follow the caller's stream.
  const int oldest = (b + 1) % Frame::kSets;
  OCR_HIP_CHECK(hipEventRecord(f->entry[b], st));
  OCR_HIP_CHECK(hipStreamWaitEvent(sb, f->entry_valid[oldest] ? f->entry[oldest] : f->entry[b], 0));
  // and the scene's last stream-ordered descriptor update; the render reads
  // the current geometry version, registered as its reader below
  std::lock_guard<std::mutex> lk(s->ver_mu);
  OCR_HIP_CHECK(s->wait_geometry(sb));
  f->entry_valid[b] = true;
  f->use = Frame::kRender;
  f->select(b);
  const Frame::OutBuf& o = f->out[b];
  const PassArgs pa = single_pass(s, view, f);
  PassArgs pc = pa;
  pc.lists.counters = nullptr;   // zeroed before prearrange below
  if (tiled) {
    // counters and tile counts, prearrange (+ boxes and counts), bin scan
    // and scatter, then depth and ids per tile (no clear: tiles write
    // every pixel)
    OCR_HIP_CHECK(launch_bin_reset(o.tb, o.counters, sb));
    if (prof) OCR_HIP_CHECK(hipEventRecord(ev[0], sb));
    OCR_HIP_CHECK(launch_prearrange(pa, sb));
    if (prof) OCR_HIP_CHECK(hipEventRecord(ev[1], sb));
    OCR_HIP_CHECK(launch_bins(pa, sb));
    if (prof) OCR_HIP_CHECK(hipEventRecord(ev[2], sb));
    OCR_HIP_CHECK(launch_tile_raster(pa, sb));
    if (prof) OCR_HIP_CHECK(hipEventRecord(ev[3], sb));
    if (prof) OCR_HIP_CHECK(hipEventRecord(ev[4], sb));
  } else {
    OCR_HIP_CHECK(launch_zero_counters(o.counters, sb));
    if (prof) OCR_HIP_CHECK(hipEventRecord(ev[0], sb));
    OCR_HIP_CHECK(launch_prearrange(pa, sb));
    if (prof) OCR_HIP_CHECK(hipEventRecord(ev[1], sb));
    OCR_HIP_CHECK(launch_clear(pc, sb));
    if (prof) OCR_HIP_CHECK(hipEventRecord(ev[2], sb));
    OCR_HIP_CHECK(launch_depth(pa, kRasterBlocksPerCu, sb));
    if (prof) OCR_HIP_CHECK(hipEventRecord(ev[3], sb));
    OCR_HIP_CHECK(launch_id(pa, kRasterBlocksPerCu, sb));
    if (prof) OCR_HIP_CHECK(hipEventRecord(ev[4], sb));
  }
  // Shading reads the shadow maps: it waits for the scene's last shadow
  // pass, wherever that was enqueued (the raster passes do not read them).
  // The shading pass is timed from after that wait.
  OCR_HIP_CHECK(s->wait_shadow_maps(sb));
  if (prof_shade) OCR_HIP_CHECK(hipEventRecord(ev[6], sb));
  OCR_HIP_CHECK(launch_shade(pa, s->ctx->face_sc, sb));
  OCR_HIP_CHECK(s->note_read(sb, true));
  if (prof_shade) OCR_HIP_CHECK(hipEventRecord(ev[5], sb));
  OCR_HIP_CHECK(hipEventRecord(f->done[b], sb));
  OCR_HIP_CHECK(hipStreamWaitEvent(st, f->done[b], 0));
  if (prof_shade) f->n_profiled++;
  f->rendered = true;
  return OCR_OK;
}

int ocr_render_noalloc(ocr_scene* s, const ocr_view* view, uint32_t* depth, unsigned long long* id_key,
                       uint32_t* rgba8, uint32_t* normals, void* stream) {
  if (!s || !view || !depth || !id_key || !rgba8 || !normals)
    return fail(OCR_ERR_INVALID, "ocr_render_noalloc: null argument");
  if (view->width < 4 || view->height < 4 || (uint64_t)view->width * view->height > kMaxTargetPixels ||
      view->width > 32768 || view->height > 32768)
    return fail(OCR_ERR_INVALID, "ocr_render_noalloc: view size outside 4x4 .. 16,000,000 pixels");
  if (!(view->fov_const >= 0x1p-60f && view->fov_const <= 0x1p60f))
    return fail(OCR_ERR_INVALID, "ocr_render_noalloc: fov_const must be in [2^-60, 2^60]");
  if (!s->atlas.pages) return fail(OCR_ERR_INVALID, "ocr_render_noalloc: scene has no atlas (ocr_atlas_upload)");
  if (s->nobj && s->max_tid >= std::max<uint32_t>(s->atlas.mip_start, 1))
    return fail(OCR_ERR_INVALID, "ocr_render_noalloc: an object's texture id is outside the atlas");
  // ocr_render gives a scene without lights empty light buffers, which
  // allocates: here the caller uploads them first
  if (!s->lights) return fail(OCR_ERR_INVALID, "ocr_render_noalloc: upload the scene's lights first (ocr_lights_upload)");
  // at most two cut pieces per triangle: the list cannot overflow
  if ((uint64_t)s->ntri * 2 + 64 > kNoAllocCapCut)
    return fail(OCR_ERR_UNSUPPORTED, "ocr_render_noalloc: scene larger than the 64 MiB scratch holds (at most " +
                                         std::to_string((kNoAllocCapCut - 64) / 2) + " triangles)");
  if (!s->ctx->noalloc_scratch) return fail(OCR_ERR_INVALID, "ocr_render_noalloc: context has no scratch");
  OCR_HIP_CHECK(hipSetDevice(s->ctx->device));
  hipStream_t st = (hipStream_t)stream;
  uint32_t* counters = reinterpret_cast<uint32_t*>(s->ctx->noalloc_scratch);
  CutTri* cuts = reinterpret_cast<CutTri*>(s->ctx->noalloc_scratch + kNoAllocCounterBytes);
  PassArgs pa;
  memset(&pa, 0, sizeof(pa));
  pa.view0.v = make_view_params(view);
  pa.view0.depth = depth;
  pa.view0.idkey = id_key;
  pa.view0.rgba = rgba8;
  pa.view0.normals = normals;
  pa.nviews = 1;
  pa.tri_slots = round256(s->ntri);
  pa.pix_slots = round256((uint64_t)view->width * view->height);
  pa.tile_w_log2 = 4;
  pa.tiles = view_tiles(view->width, view->height, pa.tile_w_log2);
  // reads the current geometry version, registered as its reader below
  std::lock_guard<std::mutex> lk(s->ver_mu);
  pa.scene0 = dev_scene(s);
  pa.unshadowed = s->n_shadow == 0 ? 1u : 0u;
  pa.lists = make_lists(cuts, counters, kNoAllocCapCut);
  PassArgs pc = pa;
  pc.lists.counters = nullptr;   // zeroed by their own kernel below
  OCR_HIP_CHECK(s->wait_geometry(st));
  OCR_HIP_CHECK(launch_zero_counters(counters, st));
  OCR_HIP_CHECK(launch_prearrange(pa, st));
  OCR_HIP_CHECK(launch_clear(pc, st));
  OCR_HIP_CHECK(launch_depth(pa, kRasterBlocksPerCu, st));
  OCR_HIP_CHECK(launch_id(pa, kRasterBlocksPerCu, st));
  OCR_HIP_CHECK(s->wait_shadow_maps(st));
  OCR_HIP_CHECK(launch_shade(pa, s->ctx->face_sc, st));
  OCR_HIP_CHECK(s->note_read(st, true));
  return OCR_OK;
}

int ocr_frame_get_stats(ocr_frame* f, void* stream, ocr_frame_stats* out) {
  if (!f || !out) return fail(OCR_ERR_INVALID, "ocr_frame_get_stats: null argument");
  if (f->use == Frame::kBatch) return fail(OCR_ERR_INVALID, "ocr_frame_get_stats: the frame is rendered by batches (ocr_batch_get_stats)");
  OCR_HIP_CHECK(hipSetDevice(f->ctx->device));
  OCR_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  uint32_t c[kCounterWords];
  OCR_HIP_CHECK(hipMemcpy(c, f->out[f->cur].counters, sizeof(c), hipMemcpyDeviceToHost));
  memset(out, 0, sizeof(*out));
  out->n_cut = c[kPacked + 1];
  out->n_frag = c[kFragCount];
  out->overflow = c[kOverflow] ? 1u : 0u;
  out->n_profiled = (uint32_t)f->n_profiled;
  if (f->n_profiled > 0) {
    // clear, prearrange, depth, id, shade (from the end of its wait for the
    // shadow maps), and the render's whole span
    double sum[6] = {0, 0, 0, 0, 0, 0};
    const int from[6] = {1, 0, 2, 3, 6, 0}, to[6] = {2, 1, 3, 4, 5, 5};
    for (int r = 0; r < f->n_profiled; r++) {
      hipEvent_t* ev = &f->ev[(size_t)r * Frame::kEvents];
      float ms = 0.f;
      for (int i = 0; i < 6; i++) {
        if (f->profiling == 2 && i != 4) continue;   // the shading pass alone
        OCR_HIP_CHECK(hipEventElapsedTime(&ms, ev[from[i]], ev[to[i]]));
        sum[i] += ms;
      }
    }
    const double n = (double)f->n_profiled;
    out->ms_clear = (float)(sum[0] / n);
    out->ms_prearrange = (float)(sum[1] / n);
    out->ms_depth = (float)(sum[2] / n);
    out->ms_id = (float)(sum[3] / n);
    out->ms_shade = (float)(sum[4] / n);
    out->ms_total = (float)(sum[5] / n);
  }
  return OCR_OK;
}

static uint64_t popcount_bits(const std::vector<uint32_t>& v) {
  uint64_t n = 0;
  for (uint32_t w : v) n += (uint64_t)__builtin_popcount(w);
  return n;
}

int ocr_frame_count_texels(ocr_scene* s, const ocr_view* view, ocr_frame* f, uint64_t* atlas_texels,
                           uint64_t* shadow_texels, void* stream) {
  if (!s || !view || !f || !f->rendered)
    return fail(OCR_ERR_INVALID, "ocr_frame_count_texels: render the frame with ocr_render first");
  if (view->width != f->width || view->height != f->height)
    return fail(OCR_ERR_INVALID, "ocr_frame_count_texels: view size differs from the frame");
  OCR_HIP_CHECK(hipSetDevice(s->ctx->device));
  hipStream_t st = (hipStream_t)stream;
  // reads the current geometry version until the synchronisation below
  std::lock_guard<std::mutex> lk(s->ver_mu);
  OCR_HIP_CHECK(s->wait_geometry(st));
  const size_t tex_words = ((size_t)s->atlas.npages * OCR_ATLAS_DIM * OCR_ATLAS_DIM + 31) / 32;
  const size_t per = (size_t)6 * s->lbuf_dim * s->lbuf_dim;
  const size_t sh_words = (per * std::max<uint32_t>(s->n_shadow, 1) + 31) / 32;
  const size_t st_words = (per * std::max<uint32_t>(s->n_static, 1) + 31) / 32;
  uint32_t *tb = nullptr, *sb = nullptr, *tt = nullptr;
  int rc;
  if ((rc = dev_alloc(&tb, tex_words, "texel bitmap")) || (rc = dev_alloc(&sb, sh_words, "shadow bitmap")) ||
      (rc = dev_alloc(&tt, st_words, "static bitmap"))) {
    dev_free(tb); dev_free(sb); dev_free(tt);
    return rc;
  }
  hipError_t e = launch_fill_u32(tb, 0u, tex_words, st);
  if (e == hipSuccess) e = launch_fill_u32(sb, 0u, sh_words, st);
  if (e == hipSuccess) e = launch_fill_u32(tt, 0u, st_words, st);
  const PassArgs pa = single_pass(s, view, f);
  if (e == hipSuccess) e = launch_shade_count(pa, s->ctx->face_sc, tb, sb, tt, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  std::vector<uint32_t> h1, h2, h3;
  if (e == hipSuccess) {
    h1.resize(tex_words); h2.resize(sh_words); h3.resize(st_words);
    e = hipMemcpy(h1.data(), tb, 4 * tex_words, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(h2.data(), sb, 4 * sh_words, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(h3.data(), tt, 4 * st_words, hipMemcpyDeviceToHost);
  }
  dev_free(tb); dev_free(sb); dev_free(tt);
  if (e != hipSuccess) return fail(OCR_ERR_HIP, std::string("ocr_frame_count_texels: ") + hipGetErrorString(e));
  if (atlas_texels) *atlas_texels = popcount_bits(h1);
  if (shadow_texels) *shadow_texels = popcount_bits(h2) + popcount_bits(h3);
  return OCR_OK;
}

int ocr_frame_read(ocr_frame* f, uint32_t* depth, uint32_t* tri_id, uint8_t* rgba8, uint16_t* normals,
                   void* stream) {
  if (!f) return fail(OCR_ERR_INVALID, "ocr_frame_read: null frame");
  OCR_HIP_CHECK(hipSetDevice(f->ctx->device));
  hipStream_t st = (hipStream_t)stream;
  OCR_HIP_CHECK(hipStreamSynchronize(st));
  uint32_t c[kCounterWords];
  // a frame rendered by batches has no counters of its own (the batch's
  // status comes from ocr_batch_get_stats); its counters stay zero
  OCR_HIP_CHECK(hipMemcpy(c, f->out[f->cur].counters, sizeof(c), hipMemcpyDeviceToHost));
  if (c[kOverflow] & kOverflowSamples)
    return fail(OCR_ERR_UNSUPPORTED, "ocr_frame_read: the render needs more than 2^31 - 1 raster samples");
  if (c[kOverflow]) {
    // A reserved cut list (or the tile bins) was too small: size the next
    // render's for this one's count (high word of the packed counter, the
    // bin total).
    if (c[kOverflow] & kOverflowCuts)
      f->want_cut = (uint32_t)std::min<uint64_t>((uint64_t)c[kPacked + 1] + c[kPacked + 1] / 4 + 64, 0xFFFFFFF0ull);
    if (c[kOverflow] & kOverflowBins)
      f->want_bin = (uint32_t)std::min<uint64_t>((uint64_t)c[kBinTotal] + c[kBinTotal] / 4 + 64, 0xFFFFFFF0ull);
    return fail(OCR_ERR_OVERFLOW, (c[kOverflow] & kOverflowCuts)
                                      ? "ocr_frame_read: cut list overflowed; re-render the frame"
                                      : "ocr_frame_read: tile bins overflowed; re-render the frame");
  }
  size_t npix = (size_t)f->width * f->height;
  if (depth) OCR_HIP_CHECK(hipMemcpy(depth, f->depth, 4 * npix, hipMemcpyDeviceToHost));
  if (rgba8) OCR_HIP_CHECK(hipMemcpy(rgba8, f->rgba, 4 * npix, hipMemcpyDeviceToHost));
  if (normals) OCR_HIP_CHECK(hipMemcpy(normals, f->normals, 4 * npix, hipMemcpyDeviceToHost));
  if (tri_id) {
    std::vector<unsigned long long> k(npix);
    std::vector<uint32_t> d(npix);
    OCR_HIP_CHECK(hipMemcpy(k.data(), f->idkey, 8 * npix, hipMemcpyDeviceToHost));
    OCR_HIP_CHECK(hipMemcpy(d.data(), f->depth, 4 * npix, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < npix; i++)
      tri_id[i] = (k[i] && d[i] != 0xFFFFFFFFu) ? (uint32_t)(((k[i] >> 32) - 1ull) >> 1) : OCR_NO_TRIANGLE;
  }
  return OCR_OK;
}

int ocr_frame_device_buffers(ocr_frame* f, void** depth, void** idkey, void** rgba8, void** normals) {
  if (!f) return fail(OCR_ERR_INVALID, "ocr_frame_device_buffers: null frame");
  if (depth) *depth = f->depth;
  if (idkey) *idkey = f->idkey;
  if (rgba8) *rgba8 = f->rgba;
  if (normals) *normals = f->normals;
  return OCR_OK;
}

}  // extern "C"
