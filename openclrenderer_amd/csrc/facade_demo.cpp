// facade_demo.cpp — a main.cpp-style program (main.cpp:53-342) driving the
// draw API of ocr_engine.hpp: containers with load functions, textures,
// lights, shadow pass, draw, blit, read-back.
//
// usage: ocr_facade_demo <scene.bin> <out.bin> [frames] [async|motion|tiled]
//        ocr_facade_demo <scene.bin> <out.bin> <frames> obj <model.obj> <scale>
//   obj: one more container loads <model.obj> (OBJ + MTL + PNG textures)
//   through its default load function, objects_container::set_file and
//   request_scale (main.cpp:68-87); the scene file's containers come first.
//   scene.bin: "OCRSCN01", u32 ncontainers, per container: u32 nobj, per
//   object: u32 ntri, u32 tex_index, obj_desc (144 B), ntri x triangle (144 B);
//   u32 ntex, per texture: u32 w, u32 h, w*h*4 bytes; u32 nlights, n x 64 B;
//   view (64 B), f32 hfov.
//   out.bin: the view actually rendered (64 B) then depth, tri id (u32 x W*H)
//   and RGBA8 (W*H*4).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/ocr_engine.hpp"

struct Obj {
  ocr_obj_desc desc;
  uint32_t tex;
  std::vector<triangle> tris;
};

static bool rd(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s scene.bin out.bin [frames] [async|motion|tiled|obj model.obj scale]\n", argv[0]);
    return 2;
  }
  int frames = argc > 3 ? atoi(argv[3]) : 2;
  // async: build the first half of the containers, render, then add the
  // rest with an async build while rendering, flipping when it is ready
  const bool async_build = argc > 4 && !strcmp(argv[4], "async");
  // tiled: every frame drawn by engine::draw_tiled_deferred
  const bool tiled = argc > 4 && !strcmp(argv[4], "tiled");
  // motion: after the frames, swap_buffers (main.cpp:283), move the camera
  // by (40, -25, 30), draw once more and apply do_motion_blur (main.cpp:270)
  const bool motion = argc > 4 && !strcmp(argv[4], "motion");
  const bool obj_mode = argc > 6 && !strcmp(argv[4], "obj");
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 3;
  char magic[8];
  uint32_t ncont = 0;
  if (!rd(f, magic, 8) || memcmp(magic, "OCRSCN01", 8) || !rd(f, &ncont, 4)) return 4;
  std::vector<std::vector<Obj>> conts(ncont);
  for (auto& c : conts) {
    uint32_t nobj = 0;
    if (!rd(f, &nobj, 4)) return 4;
    c.resize(nobj);
    for (auto& o : c) {
      uint32_t ntri = 0;
      if (!rd(f, &ntri, 4) || !rd(f, &o.tex, 4) || !rd(f, &o.desc, sizeof(o.desc))) return 4;
      o.tris.resize(ntri);
      if (ntri && !rd(f, o.tris.data(), sizeof(triangle) * ntri)) return 4;
    }
  }
  uint32_t ntex = 0;
  if (!rd(f, &ntex, 4)) return 4;
  object_context context;
  std::vector<texture*> texs;
  for (uint32_t t = 0; t < ntex; t++) {
    uint32_t wh[2];
    if (!rd(f, wh, 8)) return 4;
    std::vector<uint8_t> px((size_t)wh[0] * wh[1] * 4);
    if (!rd(f, px.data(), px.size())) return 4;
    texture* tx = context.tex_ctx.make_new();
    tx->set_raw(wh[0], wh[1], px.data());
    texs.push_back(tx);
  }
  uint32_t nl = 0;
  if (!rd(f, &nl, 4)) return 4;
  std::vector<ocr_light> ls(nl);
  if (nl && !rd(f, ls.data(), sizeof(ocr_light) * nl)) return 4;
  ocr_view view;
  float hfov = 60.f;
  if (!rd(f, &view, sizeof(view)) || !rd(f, &hfov, 4)) return 4;
  fclose(f);

  // main.cpp:68-87: containers with load functions, then the engine
  for (auto& c : conts) {
    objects_container* oc = context.make_new();
    oc->set_load_func([&c, &texs](objects_container* self) {
      for (auto& o : c) {
        object ob;
        ob.tri_list = o.tris;
        ob.tid = (cl_uint)texs[o.tex]->id;
        self->objs.push_back(ob);
      }
    });
    oc->set_active(!async_build || oc->id < conts.size() / 2);
  }
  if (obj_mode) {
    objects_container* oc = context.make_new();
    oc->set_file(argv[5]);
    oc->request_scale((float)atof(argv[6]));
    oc->set_active(true);
  }
  engine window;
  window.set_fov(hfov);
  window.load(view.width, view.height, 1000, "demo", "cl2.cl", true);
  window.set_camera_pos(cl_float4{{view.c_pos[0], view.c_pos[1], view.c_pos[2], 0}});
  window.set_camera_rot(cl_float4{{view.c_rot[0], view.c_rot[1], view.c_rot[2], 0}});
  // per-object placement and material after load (objects_container setters)
  auto place = [&]() {
    context.load_active();
    size_t ci = 0;
    for (auto* oc : context.containers) {
      if (ci >= conts.size()) break;   // the model's container keeps its loaded defaults
      for (size_t k = 0; k < oc->objs.size(); k++) {
        const ocr_obj_desc& d = conts[ci][k].desc;
        object& ob = oc->objs[k];
        ob.set_pos(cl_float4{{d.world_pos[0], d.world_pos[1], d.world_pos[2], 0}});
        ob.set_rot_quat(quaternion(d.world_rot_quat[0], d.world_rot_quat[1], d.world_rot_quat[2], d.world_rot_quat[3]));
        ob.set_dynamic_scale(d.scale);
        ob.specular = d.specular;
        ob.spec_mult = d.spec_mult;
        ob.diffuse = d.diffuse;
        ob.feature_flag = d.feature_flag;
      }
      ci++;
    }
  };
  place();
  context.set_use_linear_rendering(view.use_linear_rendering);
  context.set_clear_colour(cl_float4{{view.clear_col[0], view.clear_col[1], view.clear_col[2], view.clear_col[3]}});
  context.build(true);
  window.set_object_data(*context.fetch());
  // main.cpp:141-171: lights through the global list
  for (auto& g : ls) {
    light l;
    l.set_pos(cl_float4{{g.pos[0], g.pos[1], g.pos[2], g.pos[3]}});
    l.set_col(cl_float4{{g.col[0], g.col[1], g.col[2], g.col[3]}});
    l.set_shadow_casting(g.shadow);
    l.set_brightness(g.brightness);
    l.set_radius(g.radius);
    l.set_diffuse(g.diffuse);
    l.set_is_static(g.is_static != 0);
    light::add_light(&l);
  }
  light_gpu light_data = light::build();
  window.set_light_data(light_data);
  // main.cpp:243-342: the frame loop
  auto frame = [&]() {
    window.generate_realtime_shadowing(*context.fetch());
    compute_event ev = tiled ? window.draw_tiled_deferred(*context.fetch()) : window.draw_bulk_objs_n(*context.fetch());
    window.blit_to_screen(*context.fetch());
    window.render_block();
    window.flip();
    context.flip();
    ev.wait();
  };
  for (int i = 0; i < frames; i++) frame();
  if (async_build) {
    for (auto* oc : context.containers) oc->set_active(true);
    place();
    context.build(false, true);
    int during = 0;
    while (context.build_pending() && during < 100000) {
      frame();                     // the previous scene, until the flip
      during++;
    }
    if (context.build_pending()) {
      fprintf(stderr, "async build did not finish\n");
      return 7;
    }
    printf("async build: %d frames rendered while it ran\n", during);
    for (int i = 0; i < frames; i++) frame();
  }
  if (motion) {
    context.fetch()->swap_buffers();
    cl_float4 p = engine::c_pos;
    p.s[0] += 40.f;
    p.s[1] -= 25.f;
    p.s[2] += 30.f;
    window.set_camera_pos(p);
    window.generate_realtime_shadowing(*context.fetch());
    window.draw_bulk_objs_n(*context.fetch());
    compute_event ev = window.do_motion_blur(*context.fetch(), 1.f, 1.f);
    window.blit_to_screen(*context.fetch());
    window.render_block();
    window.flip();
    ev.wait();
    printf("motion blur: frame_id %u\n", context.fetch()->frame_id);
  }
  if (window.last_status != OCR_OK) {
    fprintf(stderr, "render failed: %s\n", ocr_last_error());
    return 5;
  }
  object_context_data* dat = context.fetch();
  FILE* o = fopen(argv[2], "wb");
  if (!o) return 6;
  ocr_view used = view;
  used.fov_const = window.get_fov_const();
  for (int k = 0; k < 4; k++) used.c_pos[k] = engine::c_pos.s[k];
  fwrite(&used, sizeof(used), 1, o);
  fwrite(dat->depth.data(), 4, dat->depth.size(), o);
  fwrite(dat->tri_id.data(), 4, dat->tri_id.size(), o);
  fwrite(dat->rgba8.data(), 1, dat->rgba8.size(), o);
  fclose(o);
  printf("facade demo: %u x %u, %u triangles, %u objects, %d frames\n", view.width, view.height,
         dat->tri_num, dat->obj_num, frames);
  return 0;
}
