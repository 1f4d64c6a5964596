// Host-only stand-ins for the HIP runtime calls and the library functions
// that server.hip's request layer uses (batches, frames, shadow pass), so
// that its threads, pools and admission run under ThreadSanitizer on a
// machine without a GPU.  A "batch render" writes, into each of its frames'
// host-memory buffers, a pattern derived from the view (the driver checks
// every completion against it) after a short random pause; a batch of more
// than kSplitViews views whose total pixels pass kSplitPixels reports
// OCR_ERR_UNSUPPORTED from its statistics, as a render that ran out of
// raster samples does, so that the server's split path runs.
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../../openclrenderer_amd/csrc/ocr_internal.h"

using namespace ocr;

constexpr uint32_t kSplitViews = 2;
constexpr uint64_t kSplitPixels = 60000;

struct ocr_batch {
  std::vector<ocr_frame*> frames;
  std::vector<ocr_view> views;
  uint64_t pixels = 0;
};

namespace {
thread_local std::mt19937 g_rng(std::hash<std::thread::id>()(std::this_thread::get_id()));
void pause_a_little() {
  std::this_thread::sleep_for(std::chrono::microseconds(g_rng() % 300));
}
}  // namespace

namespace ocr {
thread_local std::string g_stub_err;
int fail(int code, const std::string& msg) {
  g_stub_err = msg;
  return code;
}
int batch_reserve(ocr_batch* b, size_t) { return b ? OCR_OK : OCR_ERR_INVALID; }
int frame_alloc_set(Frame* f, int set) {
  const size_t n = (size_t)f->width * f->height;
  Frame::OutBuf& o = f->out[set];
  o.depth = static_cast<uint32_t*>(calloc(n, 4));
  o.idkey = static_cast<unsigned long long*>(calloc(n, 8));
  o.rgba = static_cast<uint32_t*>(calloc(n, 4));
  o.normals = static_cast<uint32_t*>(calloc(n, 4));
  return o.depth && o.idkey && o.rgba && o.normals ? OCR_OK : OCR_ERR_OOM;
}
}  // namespace ocr

extern "C" {

void ocr_frame_destroy(ocr_frame* f) {
  if (!f) return;
  for (auto& o : f->out) {
    free(o.depth);
    free(o.idkey);
    free(o.rgba);
    free(o.normals);
  }
  delete f;
}

int ocr_batch_create(ocr_ctx*, uint32_t, ocr_batch** out) {
  *out = new ocr_batch();
  return OCR_OK;
}
void ocr_batch_destroy(ocr_batch* b) { delete b; }
int ocr_batch_set_views(ocr_batch* b, ocr_scene* const* scenes, const ocr_view* views, ocr_frame* const* frames,
                        uint32_t n) {
  if (!b || !scenes || !views || !frames || !n) return OCR_ERR_INVALID;
  b->frames.assign(frames, frames + n);
  b->views.assign(views, views + n);
  b->pixels = 0;
  for (uint32_t i = 0; i < n; i++) {
    if (!scenes[i] || frames[i]->width != views[i].width || frames[i]->height != views[i].height)
      return OCR_ERR_INVALID;
    b->pixels += (uint64_t)views[i].width * views[i].height;
  }
  return OCR_OK;
}
int ocr_batch_render(ocr_batch* b, void*) {
  pause_a_little();
  for (size_t i = 0; i < b->frames.size(); i++) {
    ocr_frame* f = b->frames[i];
    const ocr_view& v = b->views[i];
    const size_t n = (size_t)v.width * v.height;
    const uint32_t tag = (uint32_t)(v.c_pos[0] * 16.f);
    for (size_t k = 0; k < n; k++) {
      f->depth[k] = tag + (uint32_t)k;
      f->rgba[k] = tag ^ (uint32_t)k;
      f->idkey[k] = ((unsigned long long)tag << 32) | k;
    }
  }
  return OCR_OK;
}
int ocr_batch_get_stats(ocr_batch* b, void*, ocr_frame_stats* st) {
  memset(st, 0, sizeof(*st));
  return b->views.size() > kSplitViews && b->pixels > kSplitPixels ? OCR_ERR_UNSUPPORTED : OCR_OK;
}
int ocr_lights_upload(ocr_scene*, const ocr_light*, uint32_t, uint32_t) { return OCR_OK; }
int ocr_shadow_pass(ocr_scene*, void*) { return OCR_OK; }

hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "stub"; }
hipError_t hipHostMalloc(void** p, size_t n, unsigned int) {
  *p = malloc(n);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* p) {
  free(p);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind, hipStream_t) {
  memcpy(dst, src, n);
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = reinterpret_cast<hipStream_t>(new int(0));
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete reinterpret_cast<int*>(s);
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) {
  *e = reinterpret_cast<hipEvent_t>(new int(0));
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete reinterpret_cast<int*>(e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) {
  pause_a_little();
  return hipSuccess;
}
hipError_t hipMemGetInfo(size_t* free_b, size_t* total_b) {
  *free_b = (size_t)8 << 20;     // with hbm_fraction 0.5: a 4 MiB budget
  *total_b = (size_t)288 << 30;
  return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) { return hipSuccess; }

}  // extern "C"
