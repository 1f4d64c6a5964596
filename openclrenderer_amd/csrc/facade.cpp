// facade.cpp — the reference's draw API (engine / object_context /
// objects_container / texture_context / light) implemented over the C ABI.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <thread>
#include <algorithm>
#include <map>
#include <set>
#include <string>

#include "../../include/ocr_engine.hpp"
#include "../../include/ocr_obj.h"

// The HIP runtime header declares a `texture` template that collides with the
// draw API's texture class, so the few runtime entry points used here are
// declared by their C ABI (hipError_t is an int enum; events and streams are
// opaque pointers).
extern "C" {
int hipEventCreateWithFlags(void** event, unsigned flags);
int hipEventRecord(void* event, void* stream);
int hipEventSynchronize(void* event);
int hipEventDestroy(void* event);
int hipDeviceSynchronize(void);
}
static const unsigned kHipEventDisableTiming = 0x2;

// ------------------------------------------------------------------ statics
cl_uint engine::width = 0, engine::height = 0, engine::depth = 0;
cl_uint engine::l_size = 1024;
cl_float4 engine::c_pos = {{0, 0, 0, 0}};
cl_float4 engine::c_rot = {{0, 0, 0, 0}};

std::vector<light*> light::lightlist;
std::vector<cl_uint> light::active;
namespace {
// The global list owns the lights add_light copies; those never removed are
// freed at exit (destroyed before the list, which is defined first).
struct LightListOwner {
  ~LightListOwner() {
    for (light* l : light::lightlist) delete l;
    light::lightlist.clear();
    light::active.clear();
  }
} g_light_list_owner;
}  // namespace
bool light::static_lights_are_dirty = true;
bool light::dirty_gpu = true;

static ocr_ctx* g_ctx = nullptr;

static void report(const char* where, int rc) {
  if (rc != OCR_OK) fprintf(stderr, "[ocr] %s failed (%d): %s\n", where, rc, ocr_last_error());
}

ocr_ctx* engine::ctx() {
  if (!g_ctx) report("ocr_create", ocr_create(0, &g_ctx));
  return g_ctx;
}

void compute_event::wait() const {
  if (ev) (void)hipEventSynchronize(ev.get());
}

static compute_event record_event(void* stream) {
  void* e = nullptr;
  compute_event out;
  if (hipEventCreateWithFlags(&e, kHipEventDisableTiming) != 0) return out;
  if (hipEventRecord(e, stream) != 0) {
    (void)hipEventDestroy(e);
    return out;
  }
  out.ev = std::shared_ptr<void>(e, [](void* p) { (void)hipEventDestroy(p); });
  return out;
}

// ---------------------------------------------------------- objects_container
void objects_container::set_pos(cl_float4 p) {
  pos = p;
  for (auto& o : objs) o.set_pos(pos);
}
void objects_container::set_rot_quat(quaternion q) {
  rot_quat = q;
  for (auto& o : objs) o.set_rot_quat(q);
}
void objects_container::set_dynamic_scale(float s) {
  dynamic_scale = s;
  for (auto& o : objs) o.set_dynamic_scale(s);
}
void objects_container::set_specular(float v) { for (auto& o : objs) o.specular = v; }
void objects_container::set_spec_mult(float v) { for (auto& o : objs) o.spec_mult = v; }
void objects_container::set_diffuse(float v) { for (auto& o : objs) o.diffuse = v; }
void objects_container::set_two_sided(bool v) { for (auto& o : objs) o.set_two_sided(v); }
void objects_container::set_outlined(bool v) { for (auto& o : objs) o.set_outlined(v); }
void objects_container::set_ss_reflective(int v) { for (auto& o : objs) o.set_ss_reflective(v != 0); }
void objects_container::set_is_static(bool v) { for (auto& o : objs) o.set_is_static(v); }
void objects_container::set_does_not_receive_dynamic_shadows(bool v) {
  for (auto& o : objs) o.set_does_not_receive_dynamic_shadows(v);
}
void objects_container::set_children_texture_id(cl_uint tid) { for (auto& o : objs) o.tid = tid; }

// ------------------------------------------------------------------- texture
void texture::set_create_colour(uint8_t r, uint8_t g, uint8_t b, uint8_t a, cl_uint w, cl_uint h) {
  width = w;
  height = h;
  rgba.resize((size_t)w * h * 4);
  for (size_t i = 0; i < (size_t)w * h; i++) {
    rgba[i * 4] = r; rgba[i * 4 + 1] = g; rgba[i * 4 + 2] = b; rgba[i * 4 + 3] = a;
  }
}

void texture::set_raw(cl_uint w, cl_uint h, const uint8_t* data) {
  width = w;
  height = h;
  rgba.assign(data, data + (size_t)w * h * 4);
}

bool texture::set_location(const std::string& path) {
  uint32_t w = 0, h = 0;
  uint8_t* px = nullptr;
  if (ocr_png_load(path.c_str(), &w, &h, &px) != OCR_OK) {
    std::fprintf(stderr, "[ocr facade] texture %s could not be decoded\n", path.c_str());
    set_create_colour(255, 0, 255, 255, 32, 32);
    return false;
  }
  set_raw(w, h, px);
  ocr_png_free(px);
  return true;
}

texture* texture_context::make_new_cached(const std::string& loc) {
  for (auto& t : all_textures)
    if (t->cacheable && t->texture_location == loc) return t.get();
  texture* t = make_new();
  t->texture_location = loc;
  return t;
}

texture* texture_context::make_new() {
  all_textures.emplace_back(new texture());
  all_textures.back()->id = gid++;
  return all_textures.back().get();
}

texture* texture_context::id_to_tex(int id) {
  for (auto& t : all_textures)
    if (t->id == id) return t.get();
  return nullptr;
}

int texture_context::get_gpu_position_id(int id) const {
  for (size_t i = 0; i < texture_id_orders.size(); i++)
    if (texture_id_orders[i] == id) return (int)i;
  return -1;
}

void texture_context::plan(object_context& ctx, std::vector<int>& order, std::vector<ocr_tex_src>& srcs) {
  // texture_context.cpp:350-420: textures used by active objects, plus
  // force_load ones, in ascending id order (a std::set in the reference)
  std::set<int> in_use;
  for (auto* c : ctx.containers) {
    if (!c->isactive) continue;
    for (auto& o : c->objs)
      if (o.isactive && id_to_tex((int)o.tid)) in_use.insert((int)o.tid);
  }
  for (auto& t : all_textures)
    if (t->force_load) in_use.insert(t->id);
  order.assign(in_use.begin(), in_use.end());
  srcs.clear();
  for (int id : order) {
    texture* t = id_to_tex(id);
    srcs.push_back(ocr_tex_src{t->width, t->height, t->rgba.data()});
  }
}

int texture_context::alloc_gpu(object_context& ctx, ocr_scene* scene) {
  std::vector<ocr_tex_src> srcs;
  plan(ctx, texture_id_orders, srcs);
  mipmap_start = (int)texture_id_orders.size();
  for (size_t i = 0; i < texture_id_orders.size(); i++) id_to_tex(texture_id_orders[i])->gpu_id = (int)i;
  if (!scene) return OCR_OK;
  int rc = ocr_atlas_upload(scene, srcs.data(), (uint32_t)srcs.size());
  report("ocr_atlas_upload", rc);
  return rc;
}

// --------------------------------------------------------------------- light
light* light::add_light(const light* l) {
  light* n = new light(*l);
  lightlist.push_back(n);
  active.push_back(1);
  dirty_gpu = true;
  if (n->shadow && n->is_static) static_lights_are_dirty = true;
  return n;
}

void light::remove_light(light* l) {
  for (size_t i = 0; i < lightlist.size(); i++)
    if (lightlist[i] == l) {
      delete l;
      lightlist.erase(lightlist.begin() + (long)i);
      active.erase(active.begin() + (long)i);
      dirty_gpu = true;
      return;
    }
}

void light::set_active(bool a) {
  for (size_t i = 0; i < lightlist.size(); i++)
    if (lightlist[i] == this) active[i] = a ? 1 : 0;
  dirty_gpu = true;
}

int light::get_num_shadowcasting_lights() {
  int n = 0;
  for (size_t i = 0; i < lightlist.size(); i++) n += active[i] && lightlist[i]->shadow == 1;
  return n;
}

int light::get_num_static_shadowcasters() {
  int n = 0;
  for (size_t i = 0; i < lightlist.size(); i++) n += active[i] && lightlist[i]->shadow && lightlist[i]->is_static;
  return n;
}

// light::build (light.cpp:145-198): the active lights packed in list order.
light_gpu light::build(light_gpu* old_dat) {
  static cl_uint generation = 0;
  light_gpu dat;
  for (size_t i = 0; i < lightlist.size(); i++) {
    if (!active[i]) continue;
    const light& l = *lightlist[i];
    ocr_light g;
    memset(&g, 0, sizeof(g));
    for (int k = 0; k < 4; k++) { g.pos[k] = l.pos.s[k]; g.col[k] = l.col.s[k]; }
    g.shadow = l.shadow;
    g.brightness = l.brightness;
    g.radius = l.radius;
    g.diffuse = l.diffuse;
    g.godray_intensity = l.godray_intensity;
    g.is_static = l.is_static;
    dat.lights.push_back(g);
  }
  dat.generation = ++generation;
  dirty_gpu = false;
  (void)old_dat;
  return dat;
}

// ---------------------------------------------------------- object_context
object_context::object_context() {}

// An async build: the scene is made on a builder thread from inputs it
// owns; flip() makes it current once `done` is set.
struct object_context::Pending {
  std::thread th;
  std::atomic<bool> done{false};
  int rc = OCR_OK;
  ocr_scene* scene = nullptr;
  std::vector<object*> objs;
  std::vector<int> order;
  cl_uint tri_num = 0, obj_num = 0;
};

object_context::~object_context() {
  if (pending) {
    if (pending->th.joinable()) pending->th.join();
    if (pending->scene) ocr_scene_destroy(pending->scene);
    pending.reset();
  }
  if (gpu_dat.frame) ocr_frame_destroy(gpu_dat.frame);
  if (gpu_dat.scene) ocr_scene_destroy(gpu_dat.scene);
}

// obj_load.cpp:181-568 on the parsed mesh: one object per usemtl part.
void obj_load(objects_container* c) {
  if (!c || !c->parent) return;
  ocr_obj_mesh* m = nullptr;
  if (ocr_obj_load(c->file.c_str(), c->requested_scale, &m) != OCR_OK) {
    std::fprintf(stderr, "[ocr facade] %s could not be loaded\n", c->file.c_str());
    return;
  }
  std::vector<triangle> tris(ocr_obj_ntri(m));
  std::vector<ocr_obj_part> parts(ocr_obj_nparts(m));
  ocr_obj_get(m, tris.data(), parts.data());
  ocr_obj_free(m);
  texture_context& tc = c->parent->tex_ctx;
  for (const ocr_obj_part& p : parts) {
    texture* tex;
    if (p.texture_kind == OCR_OBJ_TEX_FILE) {
      tex = c->textures_are_unique ? tc.make_new() : tc.make_new_cached(p.texture_cache);
      if (tex->rgba.empty()) tex->set_location(p.texture_path);
    } else if (p.texture_kind == OCR_OBJ_TEX_COLOUR) {
      tex = c->textures_are_unique ? tc.make_new() : tc.make_new_cached(p.texture_cache);
      if (tex->rgba.empty()) tex->set_create_colour(p.colour[0], p.colour[1], p.colour[2], 255, 32, 32);
    } else {
      tex = tc.make_new();
      tex->set_create_colour(255, 0, 255, 255, 32, 32);
    }
    object o;
    o.tri_list.assign(tris.begin() + p.first_tri, tris.begin() + p.first_tri + p.ntri);
    o.object_name = p.group_name;
    o.tid = (cl_uint)tex->id;
    if (p.bump_path[0]) {
      texture* bump = tc.make_new_cached(p.bump_path);
      bump->type = 1;
      if (bump->rgba.empty()) bump->set_location(p.bump_path);
      o.has_bump = 1;
    }
    if (!c->normal_map.empty()) {
      texture* nm = tc.make_new_cached(c->normal_map);
      if (nm->rgba.empty()) nm->set_location(c->normal_map);
      o.rid = (cl_uint)nm->id;
    }
    o.pos = c->pos;
    o.rot_quat = c->rot_quat;
    o.isloaded = 1;
    c->objs.push_back(o);
  }
  c->current_scale = c->requested_scale;
  c->requested_scale = 1.f;
}

objects_container* object_context::make_new() {
  owned.emplace_back(new objects_container());
  objects_container* c = owned.back().get();
  c->parent = this;
  c->id = (cl_uint)(owned.size() - 1);
  containers.push_back(c);
  return c;
}

void object_context::destroy(objects_container* obj) {
  containers.erase(std::remove(containers.begin(), containers.end(), obj), containers.end());
  for (auto it = owned.begin(); it != owned.end(); ++it)
    if (it->get() == obj) {
      owned.erase(it);
      break;
    }
}

// object_context::load_active (object_context.cpp:127): run each active,
// unloaded container's load function.
void object_context::load_active() {
  for (auto* c : containers) {
    if (!c->isactive || c->isloaded) continue;
    c->call_load_func(c);
    c->isloaded = true;
    for (auto& o : c->objs) {
      o.isloaded = 1;
      o.set_pos(c->pos);
      o.set_rot_quat(c->rot_quat);
      if (c->dynamic_scale != 1.f) o.set_dynamic_scale(c->dynamic_scale);
    }
  }
}

// generate_gpu_object_descriptor (object_context.cpp:228-339)
void object_context::write_descriptors(std::vector<ocr_obj_desc>& out) {
  write_descriptors(out, built_objects, tex_ctx.texture_id_orders);
}

void object_context::write_descriptors(std::vector<ocr_obj_desc>& out, const std::vector<object*>& objs,
                                       const std::vector<int>& tex_order) {
  out.clear();
  for (object* o : objs) {
    ocr_obj_desc d;
    memset(&d, 0, sizeof(d));
    for (int k = 0; k < 3; k++) d.world_pos[k] = o->pos.s[k];
    d.world_rot_quat[0] = o->rot_quat.x; d.world_rot_quat[1] = o->rot_quat.y;
    d.world_rot_quat[2] = o->rot_quat.z; d.world_rot_quat[3] = o->rot_quat.w;
    memcpy(d.old_world_pos_1, d.world_pos, 16);
    memcpy(d.old_world_pos_2, d.world_pos, 16);
    memcpy(d.old_world_rot_quat_1, d.world_rot_quat, 16);
    memcpy(d.old_world_rot_quat_2, d.world_rot_quat, 16);
    d.scale = o->dynamic_scale;
    int tid = -1;
    for (size_t i = 0; i < tex_order.size(); i++)
      if (tex_order[i] == (int)o->tid) tid = (int)i;
    d.tid = tid < 0 ? 0u : (uint32_t)tid;
    d.rid = o->rid;
    d.ssid = o->ssid;
    d.has_bump = o->has_bump;
    d.specular = o->specular;
    d.spec_mult = o->spec_mult;
    d.diffuse = o->diffuse;
    d.feature_flag = o->feature_flag;
    out.push_back(d);
  }
}

void object_context::build(bool force, bool async, compute_event* async_render_event) {
  (void)force; (void)async_render_event;
  // object_context.cpp:656-664: a build in flight completes and flips first
  if (pending) {
    pending->th.join();
    flip();
  }
  std::vector<object*> objs;
  std::vector<triangle> tris;
  cl_uint n = 0;
  // alloc_gpu (object_context.cpp:398-453): active containers' objects in
  // order; every triangle's vertices[0].object_id is its object's index
  for (auto* c : containers) {
    if (!c->isactive) continue;
    for (auto& o : c->objs) {
      o.object_g_id = (cl_int)n;
      objs.push_back(&o);
      for (const triangle& t : o.tri_list) {
        tris.push_back(t);
        tris.back().vertices[0].object_id = n;
      }
      n++;
    }
  }
  std::vector<int> order;
  std::vector<ocr_tex_src> srcs;
  tex_ctx.plan(*this, order, srcs);
  std::vector<ocr_obj_desc> descs;
  write_descriptors(descs, objs, order);
  auto p = std::make_shared<Pending>();
  p->objs = std::move(objs);
  p->order = std::move(order);
  p->tri_num = (cl_uint)tris.size();
  p->obj_num = (cl_uint)descs.size();
  auto make = [p](std::vector<triangle> t, std::vector<ocr_obj_desc> d, std::vector<std::vector<uint8_t>> px,
                  std::vector<ocr_tex_src> s) {
    for (size_t i = 0; i < s.size() && !px.empty(); i++) s[i].rgba = px[i].data();
    p->rc = ocr_scene_create(engine::ctx(), t.data(), (uint32_t)t.size(), d.data(), (uint32_t)d.size(), &p->scene);
    report("ocr_scene_create", p->rc);
    if (!p->rc) {
      p->rc = ocr_atlas_upload(p->scene, s.data(), (uint32_t)s.size());
      report("ocr_atlas_upload", p->rc);
    }
    p->done = true;
  };
  pending = p;
  if (!async) {
    make(std::move(tris), std::move(descs), {}, srcs);
    flip();
    return;
  }
  // the builder owns copies of the texture pixels
  std::vector<std::vector<uint8_t>> px;
  for (const ocr_tex_src& t : srcs) px.emplace_back(t.rgba, t.rgba + (size_t)t.width * t.height * 4);
  p->th = std::thread(make, std::move(tris), std::move(descs), std::move(px), srcs);
}

void object_context::flip() {
  if (!pending || !pending->done) return;
  if (pending->th.joinable()) pending->th.join();
  std::shared_ptr<Pending> p = std::move(pending);
  pending.reset();
  if (p->rc) {
    if (p->scene) ocr_scene_destroy(p->scene);
    return;
  }
  if (gpu_dat.scene) ocr_scene_destroy(gpu_dat.scene);
  gpu_dat.scene = p->scene;
  gpu_dat.tri_num = p->tri_num;
  gpu_dat.obj_num = p->obj_num;
  built_objects = std::move(p->objs);
  tex_ctx.texture_id_orders = std::move(p->order);
  tex_ctx.mipmap_start = (int)tex_ctx.texture_id_orders.size();
  for (size_t i = 0; i < tex_ctx.texture_id_orders.size(); i++)
    if (texture* t = tex_ctx.id_to_tex(tex_ctx.texture_id_orders[i])) t->gpu_id = (int)i;
}

void object_context::build_tick(bool async, compute_event* async_render_event) {
  if (!request_dirty) return;
  build(false, async, async_render_event);
  request_dirty = false;
}

void object_context::flush_locations(bool force, compute_event* render_event) {
  (void)force; (void)render_event;
  if (!gpu_dat.scene || built_objects.empty()) return;
  std::vector<ocr_obj_desc> descs;
  write_descriptors(descs);
  // enqueued like the reference's descriptor writes: the next draw follows it
  report("ocr_scene_update_objs_async",
         ocr_scene_update_objs_async(gpu_dat.scene, 0, (uint32_t)descs.size(), descs.data(), nullptr));
}

bool object_context_data::ensure_screen_buffers(int w, int h) {
  if (frame && s_w == w && s_h == h) return true;
  if (frame) ocr_frame_destroy(frame);
  frame = nullptr;
  int rc = ocr_frame_create(engine::ctx(), (uint32_t)w, (uint32_t)h, &frame);
  report("ocr_frame_create", rc);
  if (rc) return false;
  s_w = w;
  s_h = h;
  return true;
}

// -------------------------------------------------------------------- engine
// calculate_fov_constant_from_hfov (engine.cpp:119-133), then the value goes
// through std::to_string(fov) + "f" into the kernel build (engine.cpp:474-477).
float engine::get_fov_const() const {
  float fov_radians = (horizontal_fov_degrees / 360.f) * 2 * M_PI;
  float triangle_angle = fov_radians / 2;
  float fov_constant = ((float)width / 2) / tanf(triangle_angle);
  return strtof(std::to_string(fov_constant).c_str(), nullptr);
}

void engine::load(cl_uint w, cl_uint h, cl_uint d, const std::string& name, const std::string& loc,
                  bool only_3d, bool fullscreen) {
  (void)name; (void)loc; (void)only_3d; (void)fullscreen;
  width = w;
  height = h;
  depth = d;
  l_size = 1024;
  (void)ctx();
}

void engine::sync_lights(object_context_data& dat) {
  if (!dat.scene || !light_data) return;
  if (uploaded_light_generation == light_data->generation && lights_scene == dat.scene) return;
  int rc = ocr_lights_upload(dat.scene, light_data->lights.data(), (uint32_t)light_data->lights.size(), l_size);
  report("ocr_lights_upload", rc);
  last_status = rc;
  uploaded_light_generation = light_data->generation;
  lights_scene = dat.scene;
}

compute_event engine::generate_realtime_shadowing(object_context_data& dat) {
  sync_lights(dat);
  if (!dat.scene) return compute_event();
  last_status = ocr_shadow_pass(dat.scene, nullptr);
  report("ocr_shadow_pass", last_status);
  return record_event(nullptr);
}

compute_event engine::blend_with_depth(object_context_data& src, object_context_data& dst) {
  if (!src.frame || !dst.frame) return compute_event();
  last_status = ocr_frame_blend_with_depth(src.frame, dst.frame, nullptr);
  report("ocr_frame_blend_with_depth", last_status);
  return record_event(nullptr);
}

compute_event engine::do_pseudo_aa() {
  if (!obj_data || !obj_data->frame) return compute_event();
  last_status = ocr_frame_pseudo_aa(obj_data->frame, nullptr);
  report("ocr_frame_pseudo_aa", last_status);
  return record_event(nullptr);
}

void object_context_data::swap_buffers() {
  c_pos_old = engine::c_pos;
  c_rot_old = engine::c_rot;
}

compute_event engine::do_motion_blur(object_context_data& dat, cl_float strength, cl_float camera_contribution) {
  if (!dat.scene || !dat.frame) return compute_event();
  last_status = ocr_frame_motion_blur(dat.frame, dat.scene, &last_view, dat.c_pos_old.s, dat.c_rot_old.s,
                                      dat.frame_id, strength, camera_contribution, nullptr);
  report("ocr_frame_motion_blur", last_status);
  return record_event(nullptr);
}

compute_event engine::draw_tiled_deferred(object_context_data& dat) {
  if (!dat.scene || !dat.ensure_screen_buffers((int)width, (int)height)) return compute_event();
  int twl = 3;
  while (twl < 6 && (1 << twl) < tile_size) twl++;
  ocr_raster_config rc;
  memset(&rc, 0, sizeof(rc));
  rc.mode = OCR_RASTER_TILES;
  rc.tile_w_log2 = (uint32_t)twl;
  rc.tile_h_log2 = (uint32_t)std::min(twl, 11 - twl);
  last_status = ocr_frame_set_raster(dat.frame, &rc);
  report("ocr_frame_set_raster", last_status);
  if (last_status) return compute_event();
  const compute_event ev = draw(dat);
  // draw_bulk_objs_n keeps the sample-space raster passes
  rc.mode = OCR_RASTER_SAMPLES;
  (void)ocr_frame_set_raster(dat.frame, &rc);
  return ev;
}

compute_event engine::draw_bulk_objs_n(object_context_data& dat) {
  if (!dat.scene || !dat.ensure_screen_buffers((int)width, (int)height)) return compute_event();
  return draw(dat);
}

compute_event engine::draw(object_context_data& dat) {
  sync_lights(dat);
  if (!light_data) {
    last_status = ocr_lights_upload(dat.scene, nullptr, 0, l_size);
    report("ocr_lights_upload", last_status);
  }
  ocr_view v;
  memset(&v, 0, sizeof(v));
  v.width = width;
  v.height = height;
  v.fov_const = get_fov_const();
  v.use_linear_rendering = dat.use_linear_rendering;
  for (int k = 0; k < 4; k++) {
    v.c_pos[k] = c_pos.s[k];
    v.c_rot[k] = c_rot.s[k];
    v.clear_col[k] = dat.clear_col.s[k];
  }
  last_status = ocr_render(dat.scene, &v, dat.frame, nullptr);
  report("ocr_render", last_status);
  dat.frame_id++;
  last_view = v;
  return record_event(nullptr);
}

void engine::blit_to_screen(object_context_data& dat) {
  if (!dat.frame) return;
  size_t npix = (size_t)dat.s_w * dat.s_h;
  dat.depth.resize(npix);
  dat.tri_id.resize(npix);
  dat.rgba8.resize(npix * 4);
  int rc = ocr_frame_read(dat.frame, dat.depth.data(), dat.tri_id.data(), dat.rgba8.data(), nullptr, nullptr);
  if (rc == OCR_ERR_OVERFLOW) {
    // the work lists grew; render the same view again
    rc = ocr_render(dat.scene, &last_view, dat.frame, nullptr);
    if (rc == OCR_OK) rc = ocr_frame_read(dat.frame, dat.depth.data(), dat.tri_id.data(), dat.rgba8.data(), nullptr, nullptr);
  }
  report("ocr_frame_read", rc);
  last_status = rc;
}

void engine::render_block() { (void)hipDeviceSynchronize(); }
