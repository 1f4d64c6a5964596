// A host-only stand-in for libocr's C ABI (include/ocr.h) and the five HIP
// event calls facade.cpp declares, so that the facade (facade.cpp), the OBJ
// and PNG reader (obj_load.cpp) and the main.cpp-style demo can run under
// AddressSanitizer and UndefinedBehaviorSanitizer on a machine without a GPU.
// Each function checks its arguments as the library does and reads or
// writes every byte of the caller's buffers that the library's call reads or
// writes, so that a wrong size or a freed buffer on the facade's side is
// reported.  Nothing is rendered: frames hold a pattern.
#include <stdint.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "../../include/ocr.h"

struct ocr_ctx {
  int device = 0;
};
struct ocr_scene {
  std::vector<ocr_triangle> tris;
  std::vector<ocr_obj_desc> objs;
  std::vector<ocr_light> lights;
  uint32_t light_dim = 0;
  uint64_t texel_sum = 0;
};
struct ocr_frame {
  uint32_t w = 0, h = 0;
  std::vector<uint32_t> depth, tri, rgba;
  std::vector<uint16_t> normals;
  bool rendered = false;
};

namespace {
thread_local std::string g_err;
int fail(int code, const char* what) {
  g_err = what;
  return code;
}
volatile uint64_t g_sink = 0;
void touch(const void* p, size_t n) {
  const uint8_t* b = static_cast<const uint8_t*>(p);
  uint64_t s = 0;
  for (size_t i = 0; i < n; i++) s += b[i];
  g_sink = g_sink + s;
}
}  // namespace

extern "C" {

const char* ocr_last_error(void) { return g_err.c_str(); }

int ocr_create(int device, ocr_ctx** out) {
  if (!out || device < 0) return fail(OCR_ERR_INVALID, "ocr_create");
  *out = new ocr_ctx();
  (*out)->device = device;
  return OCR_OK;
}
void ocr_destroy(ocr_ctx* ctx) { delete ctx; }

int ocr_scene_create(ocr_ctx* ctx, const ocr_triangle* tris, uint32_t ntri, const ocr_obj_desc* objs, uint32_t nobj,
                     ocr_scene** out) {
  if (!ctx || !out || (ntri && !tris) || (nobj && !objs)) return fail(OCR_ERR_INVALID, "ocr_scene_create");
  ocr_scene* s = new ocr_scene();
  s->tris.assign(tris, tris + ntri);
  s->objs.assign(objs, objs + nobj);
  for (const ocr_triangle& t : s->tris)
    if (t.vertices[0].object_id >= nobj) {
      delete s;
      return fail(OCR_ERR_INVALID, "ocr_scene_create: object id out of range");
    }
  *out = s;
  return OCR_OK;
}
int ocr_scene_update_objs_async(ocr_scene* s, uint32_t first, uint32_t n, const ocr_obj_desc* objs, void*) {
  if (!s || (n && !objs) || (uint64_t)first + n > s->objs.size()) return fail(OCR_ERR_INVALID, "update");
  memcpy(s->objs.data() + first, objs, sizeof(ocr_obj_desc) * n);
  return OCR_OK;
}
int ocr_atlas_upload(ocr_scene* s, const ocr_tex_src* texs, uint32_t ntex) {
  if (!s || (ntex && !texs)) return fail(OCR_ERR_INVALID, "ocr_atlas_upload");
  for (uint32_t i = 0; i < ntex; i++) {
    if (!texs[i].rgba || texs[i].width == 0 || texs[i].height == 0 || texs[i].width > 2048 || texs[i].height > 2048)
      return fail(OCR_ERR_INVALID, "ocr_atlas_upload: bad texture");
    touch(texs[i].rgba, (size_t)texs[i].width * texs[i].height * 4);
  }
  return OCR_OK;
}
int ocr_lights_upload(ocr_scene* s, const ocr_light* lights, uint32_t n, uint32_t dim) {
  if (!s || (n && !lights) || dim < 8 || dim > 4000) return fail(OCR_ERR_INVALID, "ocr_lights_upload");
  s->lights.assign(lights, lights + n);
  s->light_dim = dim;
  return OCR_OK;
}
int ocr_shadow_pass(ocr_scene* s, void*) { return s ? OCR_OK : fail(OCR_ERR_INVALID, "ocr_shadow_pass"); }
void ocr_scene_destroy(ocr_scene* s) { delete s; }

int ocr_frame_create(ocr_ctx* ctx, uint32_t w, uint32_t h, ocr_frame** out) {
  if (!ctx || !out || w < 4 || h < 4 || (uint64_t)w * h > 16000000ull) return fail(OCR_ERR_INVALID, "frame");
  ocr_frame* f = new ocr_frame();
  f->w = w;
  f->h = h;
  const size_t n = (size_t)w * h;
  f->depth.assign(n, 0xFFFFFFFFu);
  f->tri.assign(n, OCR_NO_TRIANGLE);
  f->rgba.assign(n, 0u);
  f->normals.assign(2 * n, 0);
  *out = f;
  return OCR_OK;
}
void ocr_frame_destroy(ocr_frame* f) { delete f; }
int ocr_frame_set_raster(ocr_frame* f, const ocr_raster_config* c) {
  if (!f || !c || c->mode > OCR_RASTER_TILES) return fail(OCR_ERR_INVALID, "ocr_frame_set_raster");
  return OCR_OK;
}

int ocr_render(ocr_scene* s, const ocr_view* v, ocr_frame* f, void*) {
  if (!s || !v || !f || v->width != f->w || v->height != f->h) return fail(OCR_ERR_INVALID, "ocr_render");
  const uint32_t nt = (uint32_t)s->tris.size();
  for (size_t i = 0; i < f->depth.size(); i++) {
    f->tri[i] = nt ? (uint32_t)(i % nt) : OCR_NO_TRIANGLE;
    f->depth[i] = nt ? (uint32_t)(i * 2654435761u) >> 1 : 0xFFFFFFFFu;
    f->rgba[i] = (uint32_t)i * 40503u;
  }
  f->rendered = true;
  return OCR_OK;
}
int ocr_frame_read(ocr_frame* f, uint32_t* depth, uint32_t* tri_id, uint8_t* rgba8, uint16_t* normals, void*) {
  if (!f) return fail(OCR_ERR_INVALID, "ocr_frame_read");
  const size_t n = f->depth.size();
  if (depth) memcpy(depth, f->depth.data(), 4 * n);
  if (tri_id) memcpy(tri_id, f->tri.data(), 4 * n);
  if (rgba8) memcpy(rgba8, f->rgba.data(), 4 * n);
  if (normals) memcpy(normals, f->normals.data(), 4 * n);
  return OCR_OK;
}
int ocr_frame_blend_with_depth(ocr_frame* src, ocr_frame* dst, void*) {
  if (!src || !dst || src->w != dst->w || src->h != dst->h || !src->rendered || !dst->rendered)
    return fail(OCR_ERR_INVALID, "ocr_frame_blend_with_depth");
  for (size_t i = 0; i < dst->rgba.size(); i++)
    if (src->depth[i] <= dst->depth[i]) dst->rgba[i] = src->rgba[i];
  return OCR_OK;
}
int ocr_frame_pseudo_aa(ocr_frame* f, void*) {
  if (!f || !f->rendered) return fail(OCR_ERR_INVALID, "ocr_frame_pseudo_aa");
  return OCR_OK;
}
int ocr_frame_motion_blur(ocr_frame* f, ocr_scene* s, const ocr_view* v, const float c_pos_old[4],
                          const float c_rot_old[4], uint32_t, float, float, void*) {
  if (!f || !s || !v || !c_pos_old || !c_rot_old || !f->rendered || v->width != f->w || v->height != f->h)
    return fail(OCR_ERR_INVALID, "ocr_frame_motion_blur");
  touch(c_pos_old, 16);
  touch(c_rot_old, 16);
  return OCR_OK;
}

// The HIP event calls facade.cpp declares (hipError_t is an int enum).
int hipEventCreateWithFlags(void** event, unsigned) {
  if (!event) return 1;
  *event = new int(0);
  return 0;
}
int hipEventRecord(void* event, void*) { return event ? 0 : 1; }
int hipEventSynchronize(void* event) { return event ? 0 : 1; }
int hipEventDestroy(void* event) {
  delete static_cast<int*>(event);
  return 0;
}
int hipDeviceSynchronize(void) { return 0; }

}  // extern "C"
