// server.hip — the served form: render requests against resident scenes
// (their shadow maps rendered once when the scene is added), admitted
// against an HBM budget and dispatched in continuous batches.
//
// The reference renders one view per draw call from its main loop
// (main.cpp:254-342, engine.cpp:2356-2445); there is no request layer.  Here
// a dispatcher thread takes every pending request (up to the batch limits)
// as soon as one of `max_inflight` batch slots is free, renders them as one
// multi-view batch on that slot's stream (one launch per pass), copies the
// requested outputs to pinned host buffers on the same stream and records an
// event.  A completion thread waits for the slots' events in dispatch order
// and publishes completions with steady-clock timestamps.
//
// Host code only: no kernels live here.
#include <string.h>

#include <algorithm>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ocr_internal.h"

using namespace ocr;

namespace {

uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// HBM held by one request's frame: depth 4 + id key 8 + RGBA8 4 + normal 4.
uint64_t frame_bytes(uint32_t w, uint32_t h) { return (uint64_t)w * h * 20; }

struct Request {
  uint64_t id = 0;
  uint32_t scene = 0;
  Scene* sc = nullptr;             // resolved under the lock at submit time
  ocr_view view{};
  uint64_t t_submit = 0, t_start = 0, t_done = 0;
  int32_t status = OCR_OK;
  uint32_t batch_views = 0;
  uint32_t max_views = 0xFFFFFFFFu;   // lowered when a batch holding it overflowed
  uint64_t hbm = 0;
  ocr_frame* frame = nullptr;
  void* host = nullptr;            // pinned: rgba8 | depth | id key, as read_mask asks
  size_t host_bytes = 0;
};

struct Slot {
  ocr_batch* batch = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<Request*> reqs;
  bool busy = false;
};

}  // namespace

struct ocr_server {
  Ctx* ctx = nullptr;
  ocr_server_config cfg{};
  std::vector<Scene*> scenes;
  std::mutex mu;
  std::condition_variable cv_dispatch, cv_complete, cv_poll;
  std::deque<Request*> pending;
  std::deque<Request*> completed;
  std::deque<int> busy_order;                       // slots in dispatch order
  std::unordered_map<uint64_t, Request*> live;      // admitted, not yet released
  std::vector<Slot> slots;
  std::map<uint64_t, std::vector<ocr_frame*>> idle_frames;   // key w << 32 | h
  std::vector<ocr_frame*> all_frames;
  std::map<size_t, std::vector<void*>> idle_host;
  std::vector<void*> all_host;
  uint64_t budget = 0, held = 0;
  uint64_t pooled = 0;             // HBM of every frame in all_frames
  uint64_t n_evicted = 0;
  uint64_t next_id = 1;
  uint64_t n_batches = 0, n_done = 0, n_refused = 0, n_splits = 0;
  bool stop = false;
  int error = OCR_OK;
  std::thread dispatcher, completer;
};

namespace {

// Servers still alive at process exit are stopped (threads joined) before the
// HIP runtime's own teardown, which runs after this handler.
std::mutex g_registry_mu;
std::vector<ocr_server*> g_live_servers;
bool g_exit_hook = false;

void stop_threads(ocr_server* s);

void stop_live_servers_at_exit() {
  std::vector<ocr_server*> live;
  {
    std::lock_guard<std::mutex> g(g_registry_mu);
    live.swap(g_live_servers);
  }
  for (ocr_server* s : live) stop_threads(s);
}

void register_server(ocr_server* s) {
  std::lock_guard<std::mutex> g(g_registry_mu);
  if (!g_exit_hook) {
    std::atexit(stop_live_servers_at_exit);
    g_exit_hook = true;
  }
  g_live_servers.push_back(s);
}

void unregister_server(ocr_server* s) {
  std::lock_guard<std::mutex> g(g_registry_mu);
  g_live_servers.erase(std::remove(g_live_servers.begin(), g_live_servers.end(), s), g_live_servers.end());
}

size_t host_bytes_for(const ocr_server* s, uint32_t w, uint32_t h) {
  const size_t npix = (size_t)w * h;
  size_t b = 0;
  if (s->cfg.read_mask & OCR_SERVE_READ_RGBA) b += 4 * npix;
  if (s->cfg.read_mask & OCR_SERVE_READ_DEPTH) b += 4 * npix;
  if (s->cfg.read_mask & OCR_SERVE_READ_IDKEY) b += 8 * npix;
  return b;
}

// Frames come from a pool per size.  The pool's HBM stays within the budget:
// a frame in use belongs to an admitted request, whose bytes admission
// already counted in `held` (<= budget), so freeing idle frames of other sizes
// always makes room for a new one.  Frees synchronise the device, so they
// happen only when the size mix changes under a full budget.  Called by the
// dispatcher thread without the lock held.
int take_frame(ocr_server* s, uint32_t w, uint32_t h, ocr_frame** out) {
  const uint64_t key = (uint64_t)w << 32 | h;
  const uint64_t need = frame_bytes(w, h);
  std::vector<ocr_frame*> evict;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->idle_frames.find(key);
    if (it != s->idle_frames.end() && !it->second.empty()) {
      *out = it->second.back();
      it->second.pop_back();
      return OCR_OK;
    }
    for (auto jt = s->idle_frames.begin(); jt != s->idle_frames.end() && s->pooled + need > s->budget; ++jt) {
      while (!jt->second.empty() && s->pooled + need > s->budget) {
        ocr_frame* f = jt->second.back();
        jt->second.pop_back();
        s->pooled -= frame_bytes(f->width, f->height);
        s->all_frames.erase(std::find(s->all_frames.begin(), s->all_frames.end(), f));
        evict.push_back(f);
        s->n_evicted++;
      }
    }
    s->pooled += need;   // reserved for the frame allocated below
  }
  for (ocr_frame* f : evict) ocr_frame_destroy(f);
  ocr_frame* f = new (std::nothrow) ocr_frame();
  int rc = f ? OCR_OK : fail(OCR_ERR_OOM, "server: host allocation");
  if (f) {
    f->ctx = s->ctx;
    f->width = w;
    f->height = h;
    // one output set: server frames are only rendered in batches
    rc = frame_alloc_set(f, 0);
  }
  std::lock_guard<std::mutex> g(s->mu);
  if (rc) {
    if (f) ocr_frame_destroy(f);
    s->pooled -= need;
    return rc;
  }
  f->select(0);
  s->all_frames.push_back(f);
  *out = f;
  return OCR_OK;
}

int take_host(ocr_server* s, size_t bytes, void** out) {
  *out = nullptr;
  if (!bytes) return OCR_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->idle_host.find(bytes);
    if (it != s->idle_host.end() && !it->second.empty()) {
      *out = it->second.back();
      it->second.pop_back();
      return OCR_OK;
    }
  }
  void* p = nullptr;
  hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(OCR_ERR_OOM, std::string("server pinned buffer: ") + hipGetErrorString(e));
  }
  std::lock_guard<std::mutex> g(s->mu);
  s->all_host.push_back(p);
  *out = p;
  return OCR_OK;
}

// With the lock held: the request's frame goes back to the idle pool.
void pool_frame(ocr_server* s, Request* r) {
  if (!r->frame) return;
  s->idle_frames[(uint64_t)r->view.width << 32 | r->view.height].push_back(r->frame);
  r->frame = nullptr;
}

// Renders one batch on a slot; returns the HIP/ABI status (the requests are
// completed with it either way).
int launch_batch(ocr_server* s, Slot& slot, std::vector<Request*>& reqs) {
  std::vector<ocr_scene*> sc(reqs.size());
  std::vector<ocr_view> vw(reqs.size());
  std::vector<ocr_frame*> fr(reqs.size());
  for (size_t i = 0; i < reqs.size(); i++) {
    Request* r = reqs[i];
    int rc = take_frame(s, r->view.width, r->view.height, &r->frame);
    if (rc) return rc;
    if (!r->host) {   // a request re-dispatched after a split keeps its buffer
      rc = take_host(s, host_bytes_for(s, r->view.width, r->view.height), &r->host);
      if (rc) return rc;
      r->host_bytes = host_bytes_for(s, r->view.width, r->view.height);
    }
    sc[i] = static_cast<ocr_scene*>(r->sc);
    vw[i] = r->view;
    fr[i] = r->frame;
  }
  int rc = ocr_batch_set_views(slot.batch, sc.data(), vw.data(), fr.data(), (uint32_t)reqs.size());
  if (rc) return rc;
  const uint64_t t0 = now_ns();
  for (Request* r : reqs) {
    r->t_start = t0;
    r->batch_views = (uint32_t)reqs.size();
  }
  rc = ocr_batch_render(slot.batch, slot.stream);
  if (rc) return rc;
  for (Request* r : reqs) {
    const size_t npix = (size_t)r->view.width * r->view.height;
    uint8_t* h = static_cast<uint8_t*>(r->host);
    if (s->cfg.read_mask & OCR_SERVE_READ_RGBA) {
      OCR_HIP_CHECK(hipMemcpyAsync(h, r->frame->rgba, 4 * npix, hipMemcpyDeviceToHost, slot.stream));
      h += 4 * npix;
    }
    if (s->cfg.read_mask & OCR_SERVE_READ_DEPTH) {
      OCR_HIP_CHECK(hipMemcpyAsync(h, r->frame->depth, 4 * npix, hipMemcpyDeviceToHost, slot.stream));
      h += 4 * npix;
    }
    if (s->cfg.read_mask & OCR_SERVE_READ_IDKEY)
      OCR_HIP_CHECK(hipMemcpyAsync(h, r->frame->idkey, 8 * npix, hipMemcpyDeviceToHost, slot.stream));
  }
  OCR_HIP_CHECK(hipEventRecord(slot.done, slot.stream));
  return OCR_OK;
}

void dispatcher_main(ocr_server* s) {
  (void)hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    int free_slot = -1;
    s->cv_dispatch.wait(lk, [&] {
      if (s->stop) return true;
      if (s->pending.empty()) return false;
      for (size_t i = 0; i < s->slots.size(); i++)
        if (!s->slots[i].busy) { free_slot = (int)i; return true; }
      return false;
    });
    if (s->stop) return;
    // continuous batching: everything pending, up to the batch limits (and
    // up to the cap a split left on the oldest request)
    std::vector<Request*> reqs;
    uint64_t tris = 0;
    const uint32_t limit = std::min(s->cfg.max_batch_views, s->pending.front()->max_views);
    while (!s->pending.empty() && reqs.size() < limit) {
      Request* r = s->pending.front();
      const uint64_t t = r->sc->ntri;
      if (!reqs.empty() && tris + t > s->cfg.max_batch_tris) break;
      reqs.push_back(r);
      tris += t;
      s->pending.pop_front();
    }
    Slot& slot = s->slots[free_slot];
    slot.busy = true;
    lk.unlock();
    int rc = launch_batch(s, slot, reqs);
    lk.lock();
    if (rc) {
      // the batch never ran: complete its requests with the error
      const uint64_t t = now_ns();
      for (Request* r : reqs) {
        r->status = rc;
        r->t_done = t;
        if (r->frame) pool_frame(s, r);
        r->frame = nullptr;
        s->completed.push_back(r);
      }
      slot.busy = false;
      s->cv_poll.notify_all();
      continue;
    }
    slot.reqs = std::move(reqs);
    s->busy_order.push_back(free_slot);
    s->n_batches++;
    s->cv_complete.notify_all();
  }
}

void completer_main(ocr_server* s) {
  (void)hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_complete.wait(lk, [&] { return !s->busy_order.empty() || s->stop; });
    if (s->busy_order.empty()) return;   // stop with nothing in flight
    const int si = s->busy_order.front();
    s->busy_order.pop_front();
    Slot& slot = s->slots[si];
    lk.unlock();
    const hipError_t e = hipEventSynchronize(slot.done);
    ocr_frame_stats st;
    int rc = e == hipSuccess ? ocr_batch_get_stats(slot.batch, slot.stream, &st) : OCR_ERR_HIP;
    const uint64_t t = now_ns();
    lk.lock();
    if (rc == OCR_ERR_UNSUPPORTED && slot.reqs.size() > 1) {
      // The views together needed more raster samples than one render
      // holds: put them back at the head of the queue, in order, to be
      // rendered in batches of at most half as many views.
      const uint32_t half = (uint32_t)std::max<size_t>(1, slot.reqs.size() / 2);
      for (size_t i = slot.reqs.size(); i-- > 0;) {
        Request* r = slot.reqs[i];
        pool_frame(s, r);
        r->frame = nullptr;
        r->max_views = std::min(r->max_views, half);
        s->pending.push_front(r);
      }
      s->n_splits++;
      slot.reqs.clear();
      slot.busy = false;
      s->cv_dispatch.notify_all();
      continue;
    }
    for (Request* r : slot.reqs) {
      r->t_done = t;
      r->status = rc;
      pool_frame(s, r);
      r->frame = nullptr;
      s->completed.push_back(r);
      s->n_done++;
    }
    slot.reqs.clear();
    slot.busy = false;
    s->cv_dispatch.notify_all();
    s->cv_poll.notify_all();
  }
}

}  // namespace

extern "C" {

int ocr_server_create(ocr_ctx* ctx, const ocr_server_config* cfg, ocr_server** out) {
  if (!ctx || !cfg || !out) return fail(OCR_ERR_INVALID, "ocr_server_create: null argument");
  if (cfg->max_batch_views == 0 || cfg->max_batch_views > kMaxViews || cfg->max_inflight == 0 ||
      cfg->max_inflight > 8 || cfg->max_batch_tris == 0 || !(cfg->hbm_fraction > 0.f && cfg->hbm_fraction <= 1.f) ||
      (cfg->read_mask & ~(uint32_t)(OCR_SERVE_READ_RGBA | OCR_SERVE_READ_DEPTH | OCR_SERVE_READ_IDKEY)))
    return fail(OCR_ERR_INVALID, "ocr_server_create: bad configuration");
  *out = nullptr;
  Ctx* c = static_cast<Ctx*>(ctx);
  OCR_HIP_CHECK(hipSetDevice(c->device));
  ocr_server* s = new (std::nothrow) ocr_server();
  if (!s) return fail(OCR_ERR_OOM, "ocr_server_create: host allocation");
  s->ctx = c;
  s->cfg = *cfg;
  s->slots.resize(cfg->max_inflight);
  for (Slot& sl : s->slots) {
    int rc = ocr_batch_create(ctx, cfg->max_batch_views, &sl.batch);
    if (!rc) rc = batch_reserve(sl.batch, cfg->max_batch_tris);
    hipError_t e = hipSuccess;
    if (!rc) e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    if (!rc && e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (rc || e != hipSuccess) {
      ocr_server_destroy(s);
      return rc ? rc : fail(OCR_ERR_HIP, std::string("ocr_server_create: ") + hipGetErrorString(e));
    }
  }
  size_t free_b = 0, total_b = 0;
  hipError_t e = hipMemGetInfo(&free_b, &total_b);
  if (e != hipSuccess) {
    ocr_server_destroy(s);
    return fail(OCR_ERR_HIP, std::string("ocr_server_create: ") + hipGetErrorString(e));
  }
  s->budget = (uint64_t)((double)free_b * cfg->hbm_fraction);
  s->dispatcher = std::thread(dispatcher_main, s);
  s->completer = std::thread(completer_main, s);
  register_server(s);
  *out = s;
  return OCR_OK;
}

int ocr_server_add_scene(ocr_server* s, ocr_scene* scene, uint32_t* slot) {
  if (!s || !scene || !slot) return fail(OCR_ERR_INVALID, "ocr_server_add_scene: null argument");
  if (!scene->atlas.pages) return fail(OCR_ERR_INVALID, "ocr_server_add_scene: scene has no atlas");
  OCR_HIP_CHECK(hipSetDevice(s->ctx->device));
  if (!scene->lights) {
    int rc = ocr_lights_upload(scene, nullptr, 0, scene->lbuf_dim ? scene->lbuf_dim : 1024);
    if (rc) return rc;
  }
  // resident shadow maps: rendered once, shared by every request
  int rc = ocr_shadow_pass(scene, nullptr);
  if (rc) return rc;
  OCR_HIP_CHECK(hipDeviceSynchronize());
  std::lock_guard<std::mutex> g(s->mu);
  *slot = (uint32_t)s->scenes.size();
  s->scenes.push_back(scene);
  return OCR_OK;
}

int ocr_server_submit(ocr_server* s, uint32_t scene_slot, const ocr_view* view, uint64_t* request_id) {
  if (!s || !view || !request_id) return fail(OCR_ERR_INVALID, "ocr_server_submit: null argument");
  if (view->width < 4 || view->height < 4 || (uint64_t)view->width * view->height > kMaxTargetPixels ||
      !(view->fov_const >= 0x1p-60f && view->fov_const <= 0x1p60f))
    return fail(OCR_ERR_INVALID, "ocr_server_submit: bad view");
  Request* r = new (std::nothrow) Request();
  if (!r) return fail(OCR_ERR_OOM, "ocr_server_submit: host allocation");
  r->scene = scene_slot;
  r->view = *view;
  r->hbm = frame_bytes(view->width, view->height);
  std::lock_guard<std::mutex> g(s->mu);
  if (scene_slot >= s->scenes.size()) {
    delete r;
    return fail(OCR_ERR_INVALID, "ocr_server_submit: unknown scene slot");
  }
  // the dispatcher reads the scene through this pointer without the lock
  // (ocr_server_add_scene may grow the table meanwhile)
  r->sc = s->scenes[scene_slot];
  if (s->held + r->hbm > s->budget) {
    s->n_refused++;
    delete r;
    return fail(OCR_ERR_BUSY, "ocr_server_submit: HBM budget full; retry after releasing completions");
  }
  s->held += r->hbm;
  r->id = s->next_id++;
  r->t_submit = now_ns();
  s->live[r->id] = r;
  s->pending.push_back(r);
  *request_id = r->id;
  s->cv_dispatch.notify_one();
  return OCR_OK;
}

int ocr_server_poll(ocr_server* s, ocr_completion* out, uint32_t max, uint32_t* n, int wait_ms) {
  if (!s || !out || !n) return fail(OCR_ERR_INVALID, "ocr_server_poll: null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->completed.empty() && wait_ms > 0)
    s->cv_poll.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->completed.empty(); });
  uint32_t k = 0;
  while (k < max && !s->completed.empty()) {
    Request* r = s->completed.front();
    s->completed.pop_front();
    ocr_completion& c = out[k++];
    memset(&c, 0, sizeof(c));
    c.request_id = r->id;
    c.scene_slot = r->scene;
    c.width = r->view.width;
    c.height = r->view.height;
    c.status = r->status;
    c.batch_views = r->batch_views;
    c.t_submit_ns = r->t_submit;
    c.t_start_ns = r->t_start;
    c.t_done_ns = r->t_done;
    c.host = r->host;
  }
  *n = k;
  return OCR_OK;
}

int ocr_server_release(ocr_server* s, uint64_t request_id) {
  if (!s) return fail(OCR_ERR_INVALID, "ocr_server_release: null server");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->live.find(request_id);
  if (it == s->live.end()) return fail(OCR_ERR_INVALID, "ocr_server_release: unknown request");
  Request* r = it->second;
  if (r->t_done == 0) return fail(OCR_ERR_INVALID, "ocr_server_release: request not completed");
  if (r->host) s->idle_host[r->host_bytes].push_back(r->host);
  s->held -= r->hbm;
  s->live.erase(it);
  delete r;
  return OCR_OK;
}

uint64_t ocr_server_now_ns(void) { return now_ns(); }

}  // extern "C"

namespace {

// Lets in-flight batches finish, then stops and joins both threads; requests
// that never started stay queued (destroy frees them).
void stop_threads(ocr_server* s) {
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_poll.wait_for(lk, std::chrono::seconds(60), [&] { return s->busy_order.empty(); });
    s->stop = true;
    s->cv_dispatch.notify_all();
    s->cv_complete.notify_all();
  }
  if (s->dispatcher.joinable()) s->dispatcher.join();
  if (s->completer.joinable()) s->completer.join();
}

}  // namespace

extern "C" {

int ocr_server_get_info(ocr_server* s, ocr_server_info* info) {
  if (!s || !info) return fail(OCR_ERR_INVALID, "ocr_server_get_info: null argument");
  std::lock_guard<std::mutex> g(s->mu);
  memset(info, 0, sizeof(*info));
  info->hbm_budget = s->budget;
  info->hbm_held = s->held;
  info->n_batches = s->n_batches;
  info->n_done = s->n_done;
  info->n_refused = s->n_refused;
  info->n_splits = s->n_splits;
  info->n_pending = (uint32_t)s->pending.size();
  info->n_frames = (uint32_t)s->all_frames.size();
  info->frame_bytes = s->pooled;
  info->n_evicted = s->n_evicted;
  return OCR_OK;
}

void ocr_server_destroy(ocr_server* s) {
  if (!s) return;
  unregister_server(s);
  stop_threads(s);
  (void)hipSetDevice(s->ctx->device);
  (void)hipDeviceSynchronize();
  // every admitted request stays in `live` until released
  for (auto& kv : s->live) delete kv.second;
  for (Slot& sl : s->slots) {
    if (sl.batch) ocr_batch_destroy(sl.batch);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
  }
  for (ocr_frame* f : s->all_frames) ocr_frame_destroy(f);
  for (void* p : s->all_host) (void)hipHostFree(p);
  delete s;
}

}  // extern "C"
