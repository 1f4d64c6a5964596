// Drives server.hip's request layer (built with the host-only stand-ins of
// server_stub.hip) from three client threads: two submit 400 requests each
// (five frame sizes, retrying while the 4 MiB budget refuses them) and one
// polls, checks every completion's host buffer against the pattern its view
// implies, and releases it; the main thread reads the server's counters
// meanwhile.  Exit status 0 when every request completed correctly.
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

#include "../../openclrenderer_amd/csrc/ocr_internal.h"

int main() {
  static ocr_ctx ctx;
  ctx.device = 0;
  std::vector<ocr_scene*> scenes;
  static uint8_t fake_page;
  for (int i = 0; i < 3; i++) {
    ocr_scene* s = new ocr_scene();
    s->ctx = &ctx;
    s->ntri = 1000u << i;
    s->atlas.pages = &fake_page;
    s->lights = reinterpret_cast<ocr_light*>(&fake_page);
    scenes.push_back(s);
  }
  ocr_server_config cfg{};
  cfg.max_batch_views = 6;
  cfg.max_inflight = 3;
  cfg.max_batch_tris = 100000;
  cfg.hbm_fraction = 0.5f;
  cfg.read_mask = OCR_SERVE_READ_RGBA | OCR_SERVE_READ_DEPTH | OCR_SERVE_READ_IDKEY;
  ocr_server* srv = nullptr;
  if (ocr_server_create(&ctx, &cfg, &srv) != OCR_OK) return 2;
  uint32_t slot[3];
  for (int i = 0; i < 3; i++)
    if (ocr_server_add_scene(srv, scenes[i], &slot[i]) != OCR_OK) return 3;
  static const uint32_t sizes[5][2] = {{64, 48}, {96, 64}, {128, 96}, {200, 150}, {256, 192}};
  const int per_client = 400;
  std::atomic<int> submitted{0}, refused{0}, bad{0}, done{0};
  auto client = [&](int c) {
    for (int i = 0; i < per_client; i++) {
      ocr_view v{};
      const int k = (i * 7 + c * 3) % 5;
      v.width = sizes[k][0];
      v.height = sizes[k][1];
      v.fov_const = 500.f;
      v.c_pos[0] = (float)(c * per_client + i);   // the request's tag
      uint64_t id = 0;
      for (;;) {
        const int rc = ocr_server_submit(srv, slot[(i + c) % 3], &v, &id);
        if (rc == OCR_OK) break;
        if (rc != OCR_ERR_BUSY) {
          bad++;
          return;
        }
        refused++;
        std::this_thread::sleep_for(std::chrono::microseconds(200));
      }
      submitted++;
    }
  };
  auto poller = [&]() {
    ocr_completion cs[16];
    const int total = 2 * per_client;
    auto t_end = std::chrono::steady_clock::now() + std::chrono::seconds(120);
    while (done.load() < total && std::chrono::steady_clock::now() < t_end) {
      uint32_t n = 0;
      if (ocr_server_poll(srv, cs, 16, &n, 5) != OCR_OK) {
        bad++;
        return;
      }
      for (uint32_t i = 0; i < n; i++) {
        const ocr_completion& c = cs[i];
        const size_t npix = (size_t)c.width * c.height;
        const uint32_t* rgba = static_cast<const uint32_t*>(c.host);
        const uint32_t* depth = rgba + npix;
        const unsigned long long* key = reinterpret_cast<const unsigned long long*>(depth + npix);
        const uint32_t tag0 = (uint32_t)(key[0] >> 32);
        bool ok = c.status == OCR_OK;
        for (size_t k = 0; ok && k < npix; k++)
          ok = depth[k] == tag0 + (uint32_t)k && rgba[k] == (tag0 ^ (uint32_t)k) &&
               key[k] == (((unsigned long long)tag0 << 32) | k);
        if (!ok) bad++;
        if (ocr_server_release(srv, c.request_id) != OCR_OK) bad++;
        done++;
      }
    }
  };
  std::thread t1(client, 0), t2(client, 1), tp(poller);
  ocr_server_info info{};
  for (int i = 0; i < 50; i++) {
    ocr_server_get_info(srv, &info);
    std::this_thread::sleep_for(std::chrono::milliseconds(2));
  }
  t1.join();
  t2.join();
  tp.join();
  ocr_server_get_info(srv, &info);
  printf("submitted %d done %d refused %d bad %d batches %llu splits %llu evicted %llu frames %u\n",
         submitted.load(), done.load(), refused.load(), bad.load(), (unsigned long long)info.n_batches,
         (unsigned long long)info.n_splits, (unsigned long long)info.n_evicted, info.n_frames);
  ocr_server_destroy(srv);
  for (ocr_scene* s : scenes) delete s;
  return bad.load() == 0 && done.load() == 2 * per_client ? 0 : 1;
}
