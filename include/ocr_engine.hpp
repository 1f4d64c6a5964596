// ocr_engine.hpp — C++ draw API over the C ABI of ocr.h.
//
// Keeps the reference's engine / object_context / objects_container / object /
// texture_context / texture / light names, signatures and call order
// (engine.hpp:215-271, object_context.hpp:187-273, objects_container.hpp:63-150,
// texture_context.hpp:39-72, light.hpp:25-69; call sites main.cpp:68-293), so
// an application's draw loop ports by swapping headers:
//
//   object_context context;
//   objects_container* c = context.make_new();
//   c->set_load_func(...); c->set_active(true);
//   engine window; window.load(w, h, 1000, "name", "cl2.cl", true);
//   context.load_active(); context.build(true);
//   window.set_object_data(*context.fetch());
//   light l; ...; light::add_light(&l); auto ld = light::build(); window.set_light_data(ld);
//   loop: window.generate_realtime_shadowing(*context.fetch());
//         window.draw_bulk_objs_n(*context.fetch()); window.blit_to_screen(*context.fetch());
//         window.render_block(); window.flip();
//
// What does not carry over: OpenGL presentation (the frame stays offscreen and
// blit_to_screen reads it back, engine::get_screen_rgba8), SFML windows and
// input, runtime OpenCL program options (the kernels are compiled in; the
// -D constants of main.cpp:79-85 are fixed).  set_file loads OBJ/MTL files
// with PNG textures through ocr_obj_load (include/ocr_obj.h) unless
// objects_container::set_load_func names another loader.
#ifndef OCR_ENGINE_HPP_INCLUDED
#define OCR_ENGINE_HPP_INCLUDED

#include <stdint.h>

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "ocr.h"

#ifndef __OPENCL_CL_H
typedef uint32_t cl_uint;
typedef int32_t cl_int;
typedef float cl_float;
typedef union {
  float s[4];
  struct { float x, y, z, w; };
} cl_float4;
#endif

// Stand-in for vec.hpp's quaternion as the draw API uses it (x, y, z, w).
struct quaternion {
  float x = 0.f, y = 0.f, z = 0.f, w = 1.f;
  quaternion() = default;
  quaternion(float x_, float y_, float z_, float w_) : x(x_), y(y_), z(z_), w(w_) {}
};

// triangle / vertex: byte-compatible with the reference's (ocr.h).
typedef ocr_vertex vertex;
typedef ocr_triangle triangle;

struct object_context;
struct objects_container;

// A device-side event: replaces compute::event (engine.hpp:241).
struct compute_event {
  std::shared_ptr<void> ev;  // hipEvent_t, destroyed with the last copy
  void wait() const;
};

struct object {
  std::string object_name;
  cl_float4 pos{{0, 0, 0, 0}};
  quaternion rot_quat;
  float dynamic_scale = 1.f;
  bool isactive = true;
  int isloaded = 0;
  std::vector<triangle> tri_list;
  cl_uint tid = 0;           // texture id (texture::id) on load
  cl_uint rid = 0xFFFFFFFFu;
  cl_uint ssid = 0xFFFFFFFFu;
  cl_uint has_bump = 0;
  float specular = 0.9f;     // object.cpp:72-74
  float spec_mult = 1.f;
  float diffuse = 1.f;
  cl_int feature_flag = 0;
  cl_int object_g_id = -1;   // descriptor index, assigned by object_context::build

  void set_pos(cl_float4 p) { pos = p; }
  void set_rot_quat(quaternion q) { rot_quat = q; }
  void set_dynamic_scale(float s) { dynamic_scale = s; }
  void set_two_sided(bool v) { set_flag(OCR_FEATURE_TWO_SIDED, v); }
  void set_outlined(bool v) { set_flag(OCR_FEATURE_OUTLINE, v); }
  void set_ss_reflective(bool v) { set_flag(OCR_FEATURE_SS_REFLECTIVE, v); }
  void set_is_static(bool v) { set_flag(OCR_FEATURE_IS_STATIC, v); }
  void set_does_not_receive_dynamic_shadows(bool v) { set_flag(OCR_FEATURE_DOES_NOT_RECEIVE_DYNAMIC_SHADOWS, v); }
  void set_flag(cl_int flag, bool v) { feature_flag = v ? (feature_flag | flag) : (feature_flag & ~flag); }
};

struct objects_container;
// obj_load (obj_load.cpp:181-568): the default load function, reading
// `file` (Wavefront OBJ + MTL, include/ocr_obj.h) into one object per usemtl
// group with its texture (a PNG map_Kd, a Kd colour or magenta).
void obj_load(objects_container* c);

struct objects_container {
  object_context* parent = nullptr;
  cl_uint id = 0;
  std::string file;
  float requested_scale = 1.f;       // applied to positions at load
  float current_scale = 1.f;
  bool textures_are_unique = false;
  std::string normal_map;
  bool isactive = false;
  bool isloaded = false;
  cl_float4 pos{{0, 0, 0, 0}};
  quaternion rot_quat;
  float dynamic_scale = 1.f;
  std::vector<object> objs;
  std::function<void(objects_container*)> fp = obj_load;   // objects_container.cpp:22

  void set_file(const std::string& f) { file = f; }
  void request_scale(float f) { requested_scale = f; }
  void set_unique_textures(bool u) { textures_are_unique = u; }
  void set_normal(const std::string& n) { normal_map = n; }
  void set_load_func(std::function<void(objects_container*)> f) { fp = std::move(f); }
  void call_load_func(objects_container* c) { if (fp) fp(c); }
  void set_active(bool param) { isactive = param; }
  void set_pos(cl_float4 p);
  void set_rot_quat(quaternion q);
  void set_dynamic_scale(float s);
  void set_specular(float v);
  void set_spec_mult(float v);
  void set_diffuse(float v);
  void set_two_sided(bool v);
  void set_outlined(bool v);
  void set_ss_reflective(int v);
  void set_is_static(bool v);
  void set_does_not_receive_dynamic_shadows(bool v);
  void set_children_texture_id(cl_uint tid);
  void unload_tris() { objs.clear(); isloaded = false; }
};

struct texture {
  int id = -1;
  std::string texture_location;    // cache key of make_new_cached
  bool cacheable = true;
  int type = 0;                    // 1 = bump map
  cl_uint width = 0, height = 0;
  std::vector<uint8_t> rgba;       // row 0 first, tightly packed
  bool force_load = false;
  int gpu_id = -1;
  // texture::set_create_colour (texture.hpp): a solid-colour texture
  void set_create_colour(uint8_t r, uint8_t g, uint8_t b, uint8_t a, cl_uint w, cl_uint h);
  void set_raw(cl_uint w, cl_uint h, const uint8_t* data);
  // decodes a PNG file (ocr_png_load); magenta 32x32 if it cannot
  bool set_location(const std::string& path);
};

struct texture_context {
  std::vector<std::unique_ptr<texture>> all_textures;
  std::vector<int> texture_id_orders;
  int mipmap_start = 0;
  int gid = 0;
  texture* make_new();
  // texture_context.cpp:27-45: the texture already made for `loc`, if any
  texture* make_new_cached(const std::string& loc);
  texture* id_to_tex(int id);
  // position of a texture id in the last upload order (get_gpu_position_id)
  int get_gpu_position_id(int id) const;
  // the upload order and sources alloc_gpu would use, without uploading
  void plan(object_context& ctx, std::vector<int>& order, std::vector<ocr_tex_src>& srcs);
  // uploads every texture used by an active object (or force_load) into the
  // scene's atlas, ordered by id as the reference's std::set is
  int alloc_gpu(object_context& ctx, ocr_scene* scene);
};

struct object_context_data {
  ocr_scene* scene = nullptr;
  cl_uint tri_num = 0;
  cl_uint obj_num = 0;
  cl_uint frame_id = 0;
  cl_uint use_linear_rendering = 1;
  cl_float4 clear_col{{0, 0, 0, 0}};
  // camera of the last swap_buffers (object_context.hpp:89-90), the
  // previous frame's camera for do_motion_blur
  cl_float4 c_pos_old{{0, 0, 0, 0}};
  cl_float4 c_rot_old{{0, 0, 0, 0}};
  ocr_frame* frame = nullptr;        // offscreen framebuffer (ensure_screen_buffers)
  int s_w = 0, s_h = 0;
  bool ensure_screen_buffers(int w, int h);
  // object_context.cpp:17-25: records engine::c_pos / c_rot as the old
  // camera (the depth double buffer is the frame's own)
  void swap_buffers();
  // host read-back of the last blit (blit_to_screen)
  std::vector<uint32_t> depth, tri_id;
  std::vector<uint8_t> rgba8;
};

struct light_gpu {
  std::vector<ocr_light> lights;     // packed active lights (light.cpp:145-198)
  cl_uint generation = 0;
};

struct light {
  cl_float4 pos{{0, 0, 0, 0}};
  cl_float4 col{{1, 1, 1, 0}};
  cl_uint shadow = 0;
  cl_float brightness = 1.f;        // light.cpp:88-93
  cl_float radius = 3.4028235e38f / 100000.0f;
  cl_float diffuse = 1.f;
  cl_float godray_intensity = 0.f;
  cl_int is_static = 0;

  void set_pos(cl_float4 p) { pos = p; dirty_gpu = true; }
  void set_col(cl_float4 c) { col = c; dirty_gpu = true; }
  void set_shadow_casting(cl_uint s) { shadow = s; dirty_gpu = true; }
  void set_brightness(cl_float b) { brightness = b; dirty_gpu = true; }
  void set_radius(cl_float r) { radius = r; dirty_gpu = true; }
  void set_diffuse(cl_float d) { diffuse = d; dirty_gpu = true; }
  void set_godray_intensity(cl_float g) { godray_intensity = g; dirty_gpu = true; }
  void set_is_static(bool s) { is_static = s; static_lights_are_dirty = true; dirty_gpu = true; }
  void set_active(bool a);

  static std::vector<light*> lightlist;
  static std::vector<cl_uint> active;
  static bool static_lights_are_dirty;
  static bool dirty_gpu;
  static light* add_light(const light* l);
  static void remove_light(light* l);
  static light_gpu build(light_gpu* old_dat = nullptr);
  static int get_num_shadowcasting_lights();
  static int get_num_static_shadowcasters();
};

struct object_context {
  texture_context tex_ctx;
  std::vector<objects_container*> containers;
  object_context();
  ~object_context();
  objects_container* make_new();
  void destroy(objects_container* obj);
  void load_active();
  // object_context::build (object_context.cpp:646): concatenates every active
  // container's objects, stamps object ids, writes descriptors, uploads
  // geometry and atlas.  Synchronous by default: the new scene is current on
  // return.  With async the upload runs on a builder thread while the
  // current scene keeps rendering, and flip() makes it current once it has
  // finished (the reference's second queue and ready_to_flip callback,
  // object_context.cpp:596-612, :486-594); a further build first waits for
  // the one in flight.  Textures must not change while a build is in flight.
  void build(bool force = false, bool async = false, compute_event* async_render_event = nullptr);
  void build_request() { request_dirty = true; }
  void build_tick(bool async = false, compute_event* async_render_event = nullptr);
  bool build_pending() const { return pending != nullptr; }
  object_context_data* fetch() { return &gpu_dat; }
  void set_clear_colour(const cl_float4& col) { gpu_dat.clear_col = col; }
  // pushes pos / rotation / scale / flags of every object (object::g_flush)
  void flush_locations(bool force = false, compute_event* render_event = nullptr);
  void flip();   // object_context.cpp:839-842: applies a finished async build
  void set_use_linear_rendering(cl_uint v) { gpu_dat.use_linear_rendering = v; }
  object_context_data gpu_dat;

 private:
  std::vector<std::unique_ptr<objects_container>> owned;
  std::vector<object*> built_objects;
  bool request_dirty = false;
  struct Pending;
  std::shared_ptr<Pending> pending;
  void write_descriptors(std::vector<ocr_obj_desc>& out);
  void write_descriptors(std::vector<ocr_obj_desc>& out, const std::vector<object*>& objs,
                         const std::vector<int>& tex_order);
};

struct engine {
  static cl_uint width, height, depth;
  static cl_uint l_size;             // light cube-map size (engine.cpp:471)
  static cl_float4 c_pos, c_rot;
  float horizontal_fov_degrees = 120;
  object_context_data* obj_data = nullptr;
  light_gpu* light_data = nullptr;

  // engine::load (engine.cpp:290): creates the device context (device 0)
  void load(cl_uint width, cl_uint height, cl_uint depth, const std::string& name,
            const std::string& loc, bool only_3d, bool fullscreen = false);
  void set_fov(float horizontal_degrees) { horizontal_fov_degrees = horizontal_degrees; }
  void set_camera_pos(cl_float4 p) { c_pos = p; }
  void set_camera_rot(cl_float4 r) { c_rot = r; }
  void set_light_data(light_gpu& l) { light_data = &l; }
  void set_object_data(object_context_data& d) { obj_data = &d; }
  void append_opencl_extra_command_line(const std::string&) {}
  compute_event generate_realtime_shadowing(object_context_data& dat);
  compute_event draw_bulk_objs_n(object_context_data& dat);
  // engine::draw_tiled_deferred (engine.cpp:2447-2535): the same draw with
  // the depth and id passes resolved per screen tile of tile_size x
  // min(tile_size, 2048 / tile_size) pixels (ocr_frame_set_raster,
  // OCR_RASTER_TILES; tile_size a power of two, 16..64).  The reference's
  // experimental tiled draw stops at a depth buffer of its own coverage
  // rule; this one draws the whole frame, bit for bit as draw_bulk_objs_n.
  compute_event draw_tiled_deferred(object_context_data& dat);
  cl_int tile_size = 32;             // engine.hpp:139 (the reference sets 64, engine.cpp:684)
 private:
  compute_event draw(object_context_data& dat);   // the render both draws share
 public:
  // engine.cpp:2550-2570: src's screen over dst's where src is at least as
  // near (both contexts drawn at this engine's size)
  compute_event blend_with_depth(object_context_data& src, object_context_data& dst);
  // engine.cpp:1513-1516: edge smoothing of the last draw's screen
  compute_event do_pseudo_aa();
  // engine.cpp:1518-1536: motion blur of dat's last draw against its old
  // camera and each object's pose of the last pass (frame parity dat.frame_id)
  compute_event do_motion_blur(object_context_data& dat, cl_float strength, cl_float camera_contribution);
  void blit_to_screen(object_context_data& dat);
  void render_block();
  void flip() { frame_count++; }
  void clear_depth_buffer(object_context_data&) {}
  float get_fov_const() const;
  const std::vector<uint8_t>& get_screen_rgba8() const { return obj_data->rgba8; }
  int last_status = OCR_OK;          // status of the last C-ABI call
  uint64_t frame_count = 0;

  static ocr_ctx* ctx();

 private:
  ocr_view last_view{};
  cl_uint uploaded_light_generation = 0xFFFFFFFFu;
  ocr_scene* lights_scene = nullptr;
  void sync_lights(object_context_data& dat);
};

#endif  // OCR_ENGINE_HPP_INCLUDED
