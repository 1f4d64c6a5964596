// batch.hip — many views of several resident scenes in one render: one
// launch per pass (clear, prearrange, depth, id, shade) covers every view,
// with per-view work ranges found from device view tables, shared work lists,
// hipGraph capture of the whole multi-view frame, and an autotuner for the
// raster grid size and the shading tile shape.
//
// The reference renders one view per draw call with screen size, FOV and
// light-map size baked into the program (ocl.h:227-240, engine.cpp:1794-2161);
// here they are per-view runtime values so views of different sizes share a
// launch.
#pragma clang fp contract(off)

#include <string.h>

#include <algorithm>
#include <chrono>
#include <new>
#include <vector>

#include "ocr_internal.h"

using namespace ocr;

struct ocr_batch {
  Ctx* ctx = nullptr;
  uint32_t max_views = 0, nviews = 0;
  std::vector<Scene*> scenes;        // distinct scenes, table order
  std::vector<DevView> host_views;
  DevView* d_views = nullptr;
  DevScene* d_scenes = nullptr;
  // the generation of each scene that d_scenes was built from (a scene's
  // generation moves when its geometry version, atlas or lights are
  // replaced), and the pinned staging of its refresh (reused once the last
  // refresh copy ended)
  std::vector<uint64_t> bound_gen;
  DevScene* ds_host = nullptr;
  hipEvent_t ds_done = nullptr;
  bool ds_pending = false;
  uint32_t tri_slots = 0, pix_slots = 0, tiles = 0;
  uint64_t sum_tris = 0, sum_pix = 0;
  CutTri* cuts = nullptr;
  uint32_t* counters = nullptr;
  uint32_t cap_cut = 0;
  // tiled raster path (tune.raster_tile != 0): bins over every view's tiles
  TileBins tb{};
  uint32_t box_cap = 0, bin_tiles = 0, want_bin = 0;
  // covered-sample entries of the depth pass (read by the id pass); fixed
  // at 256 MiB for a batch reserved for serving
  CoverEntries ent{};
  bool ent_fixed = false;
  ocr_batch_tuning tune{8, 4, 1, 0};
  // tune.groups > 1: sub-batches over pixel-balanced groups of the views,
  // each rendered on a stream of its own between a fork and a join event
  std::vector<ocr_batch*> subs;
  std::vector<hipStream_t> sub_streams;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr;
  // the views as last set, to regroup them when the tuning changes
  std::vector<ocr_scene*> set_scenes;
  std::vector<ocr_view> set_views;
  std::vector<ocr_frame*> set_frames;
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t graph_unshadowed = 0;     // the shading variant the graph was captured with
  bool profiling = false;
  std::vector<hipEvent_t> ev;        // kRing x 2 (begin, end of a batch)
  int n_profiled = 0;
};

static constexpr int kBatchRing = 1024;

static void drop_graph(ocr_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

static PassArgs batch_pass(const ocr_batch* b) {
  PassArgs a;
  memset(&a, 0, sizeof(a));
  a.views = b->d_views;
  a.scenes = b->d_scenes;
  a.nviews = b->nviews;
  a.tri_slots = b->tri_slots;
  a.pix_slots = b->pix_slots;
  a.tiles = b->tiles;
  a.tile_w_log2 = b->tune.shade_tile_w_log2;
  a.unshadowed = 1u;
  for (const Scene* s : b->scenes)
    if (s->n_shadow) a.unshadowed = 0u;
  a.lists = make_lists(b->cuts, b->counters, b->cap_cut);
  a.lists.ent = b->ent;
  if (b->tune.raster_tile) a.lists.tb = b->tb;
  return a;
}

static uint32_t raster_twl(const ocr_batch* b) { return b->tune.raster_tile & 0xFFu; }
static uint32_t raster_thl(const ocr_batch* b) { return b->tune.raster_tile >> 8; }

// Recompute the per-view ranges (they depend on the tile shape) and upload
// the view table.
static int upload_views(ocr_batch* b) {
  uint32_t tri = 0, pix = 0, tiles = 0, bins = 0;
  const bool tiled = b->tune.raster_tile != 0;
  for (auto& v : b->host_views) {
    const Scene* s = b->scenes[v.scene];
    v.tri_begin = tri;
    v.pix_begin = pix;
    v.tile_begin = tiles;
    v.bin_begin = bins;
    v.bin_tiles_x = tiled ? bin_tiles_x(v.v.width, raster_twl(b)) : 0u;
    tri += (s->ntri + 255) / 256 * 256;
    pix += (v.v.width * v.v.height + 255) / 256 * 256;
    tiles += view_tiles(v.v.width, v.v.height, b->tune.shade_tile_w_log2);
    if (tiled) bins += bin_tiles(v.v.width, v.v.height, raster_twl(b), raster_thl(b));
  }
  b->tri_slots = tri;
  b->pix_slots = pix;
  b->tiles = tiles;
  b->bin_tiles = bins;
  OCR_HIP_CHECK(hipMemcpy(b->d_views, b->host_views.data(), sizeof(DevView) * b->nviews, hipMemcpyHostToDevice));
  return OCR_OK;
}

// At most 2 cut pieces per triangle of every view, so the list never
// overflows.
static int ensure_bins_of(ocr_batch* b) {
  if (!b->tune.raster_tile) return OCR_OK;
  const uint64_t dflt = std::max<uint64_t>(2ull * b->cap_cut, 4096ull);
  const uint32_t cap = (uint32_t)std::min<uint64_t>(b->want_bin ? (uint64_t)b->want_bin : dflt, 0xFFFFFFF0ull);
  const TileBins before = b->tb;
  const int rc = ensure_bins(b->tb, b->box_cap, b->cap_cut, b->bin_tiles, raster_twl(b), raster_thl(b), cap,
                             b->want_bin != 0);
  // a captured graph holds the bins' addresses and sizes
  if (memcmp(&before, &b->tb, sizeof(TileBins)) != 0) drop_graph(b);
  return rc;
}

// Covered-sample entries for 10 raster samples per pixel of the views (80
// bytes per pixel); a batch reserved for serving (batch_reserve) keeps a
// fixed 256 MiB and sweeps the chunks beyond.  They only grow, so batches
// of changing views keep theirs.
constexpr uint32_t kBatchEntMaxChunks = (uint32_t)((256ull << 20) / (kEntCap * sizeof(uint2)));

static int ensure_entries(ocr_batch* b) {
  const uint64_t need = (b->sum_pix * 10 + kChunk - 1) / kChunk;
  const uint32_t want = b->ent_fixed ? kBatchEntMaxChunks
                                     : (uint32_t)std::min<uint64_t>(std::max<uint64_t>(need, 1), 0x7FFFFFFFull / kChunk);
  if (b->ent.ent && b->ent.chunks >= want) return OCR_OK;
  OCR_HIP_CHECK(hipDeviceSynchronize());
  drop_graph(b);
  if (b->ent.ent) (void)hipFree(b->ent.ent);
  if (b->ent.cnt) (void)hipFree(b->ent.cnt);
  b->ent = CoverEntries{};
  hipError_t e = hipMalloc((void**)&b->ent.ent, sizeof(uint2) * (size_t)kEntCap * want);
  if (e == hipSuccess) e = hipMalloc((void**)&b->ent.cnt, sizeof(uint32_t) * (size_t)want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (b->ent.ent) (void)hipFree(b->ent.ent);
    b->ent = CoverEntries{};
    return fail(OCR_ERR_OOM, std::string("batch covered-sample entries: ") + hipGetErrorString(e));
  }
  b->ent.chunks = want;
  return OCR_OK;
}

static int ensure_lists(ocr_batch* b) {
  uint64_t want_cut = std::min<uint64_t>(b->sum_tris * 2 + 64, 0xFFFFFFF0ull);
  if (b->cuts && b->cap_cut >= want_cut) {
    const int rc = ensure_entries(b);
    return rc ? rc : ensure_bins_of(b);
  }
  OCR_HIP_CHECK(hipDeviceSynchronize());
  drop_graph(b);
  if (b->cuts) (void)hipFree(b->cuts);
  b->cuts = nullptr;
  b->cap_cut = 0;
  hipError_t e = hipMalloc((void**)&b->cuts, sizeof(CutTri) * want_cut * 2);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(OCR_ERR_OOM, std::string("batch cut list: ") + hipGetErrorString(e));
  }
  b->cap_cut = (uint32_t)want_cut;
  const int rc = ensure_entries(b);
  return rc ? rc : ensure_bins_of(b);
}

// Sizes the lists ahead of use (a server's batch slots): the cut list for
// sum_tris triangles and the covered-sample entries at their largest, so
// batches of any views later allocate nothing.
int ocr::batch_reserve(ocr_batch* b, uint64_t sum_tris) {
  OCR_HIP_CHECK(hipSetDevice(b->ctx->device));
  const uint64_t keep = b->sum_tris;
  b->sum_tris = sum_tris;
  b->ent_fixed = true;
  const int rc = ensure_lists(b);
  b->sum_tris = keep;
  return rc;
}

// A batch reads its scenes' geometry and shadow maps: it follows their last
// stream-ordered descriptor update and shadow pass, wherever enqueued.
static hipError_t wait_scenes(ocr_batch* b, hipStream_t st) {
  for (Scene* s : b->scenes) {
    hipError_t e = s->wait_geometry(st);
    if (e == hipSuccess) e = s->wait_shadow_maps(st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Holds the version locks of the batch's scenes (address order, so two
// batches sharing scenes cannot deadlock).
static void lock_scenes(const std::vector<Scene*>& scenes, std::vector<std::unique_lock<std::mutex>>& locks) {
  std::vector<Scene*> order = scenes;
  std::sort(order.begin(), order.end());
  order.erase(std::unique(order.begin(), order.end()), order.end());
  for (Scene* s : order) locks.emplace_back(s->ver_mu);
}

// With the scenes locked: every scene still has what the kernels read
// through the scene table (an atlas or light upload that failed after the
// views were set leaves a scene without one of them).
static int check_scenes(const ocr_batch* b) {
  for (const Scene* s : b->scenes) {
    if (!s->atlas.pages) return fail(OCR_ERR_INVALID, "batch: a scene lost its atlas (failed ocr_atlas_upload)");
    if (s->nlights && !s->lights) return fail(OCR_ERR_INVALID, "batch: a scene's lights are missing");
    if (s->n_shadow && !s->shadow_maps) return fail(OCR_ERR_INVALID, "batch: a scene's shadow maps are missing");
    if (s->nobj && s->max_tid >= std::max<uint32_t>(s->atlas.mip_start, 1))
      return fail(OCR_ERR_INVALID, "batch: an object's texture id is outside its scene's atlas");
  }
  return OCR_OK;
}

// With the scenes locked: points d_scenes at the scenes' current buffers (a
// copy on `st` when an update, atlas or light upload replaced one since the
// last launch) and makes `st` follow the scenes' updates and shadow passes.
static hipError_t bind_scenes(ocr_batch* b, hipStream_t st) {
  bool changed = false;
  for (size_t i = 0; i < b->scenes.size(); i++) changed |= b->bound_gen[i] != b->scenes[i]->gen;
  if (changed) {
    if (b->ds_pending) {
      hipError_t e = hipEventSynchronize(b->ds_done);
      if (e != hipSuccess) return e;
    }
    for (size_t i = 0; i < b->scenes.size(); i++) {
      b->ds_host[i] = dev_scene(b->scenes[i]);
      b->bound_gen[i] = b->scenes[i]->gen;
    }
    hipError_t e = hipMemcpyAsync(b->d_scenes, b->ds_host, sizeof(DevScene) * b->scenes.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(b->ds_done, st);
    if (e != hipSuccess) return e;
    b->ds_pending = true;
  }
  return wait_scenes(b, st);
}

// With the scenes locked, after the launch: the scenes' current geometry
// versions and shadow maps are read by work on `st`.
static hipError_t note_scene_reads(ocr_batch* b, hipStream_t st) {
  for (Scene* s : b->scenes) {
    hipError_t e = s->note_read(st, true);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

static hipError_t enqueue_batch(ocr_batch* b, hipStream_t st) {
  const PassArgs pa = batch_pass(b);
  hipError_t e;
  if (b->tune.raster_tile) {
    // counters and tile counts, prearrange (+ boxes and counts), bins, then
    // depth and ids per tile for every view (the tiles write every pixel)
    e = launch_bin_reset(pa.lists.tb, pa.lists.counters, st);
    if (e == hipSuccess) e = launch_prearrange(pa, st);
    if (e == hipSuccess) e = launch_bins(pa, st);
    if (e == hipSuccess) e = launch_tile_raster(pa, st);
  } else {
    e = launch_clear(pa, st);
    if (e == hipSuccess) e = launch_prearrange(pa, st);
    if (e == hipSuccess) e = launch_depth(pa, (int)b->tune.raster_blocks_per_cu, st);
    if (e == hipSuccess) e = launch_id(pa, (int)b->tune.raster_blocks_per_cu, st);
  }
  if (e == hipSuccess) e = launch_shade(pa, b->ctx->face_sc, st);
  return e;
}

// Groups in use: the tuning's, but never more than there are views.
static uint32_t group_count(const ocr_batch* b) {
  return std::min<uint32_t>(std::max<uint32_t>(b->tune.groups, 1), std::max<uint32_t>(b->nviews, 1));
}

// Splits the views set last into tune.groups contiguous, pixel-balanced
// groups and sets each on a sub-batch (created on first use) with the
// batch's raster and shading tuning.
static int regroup(ocr_batch* b) {
  const uint32_t g = group_count(b);
  if (g <= 1 || b->set_views.empty()) return OCR_OK;
  if (b->cuts) {
    // the groups hold the work lists now; the batch's own are sized for
    // every view and would double the memory
    OCR_HIP_CHECK(hipDeviceSynchronize());
    drop_graph(b);
    OCR_HIP_CHECK(hipFree(b->cuts));
    b->cuts = nullptr;
    b->cap_cut = 0;
    if (b->ent.ent) OCR_HIP_CHECK(hipFree(b->ent.ent));
    if (b->ent.cnt) OCR_HIP_CHECK(hipFree(b->ent.cnt));
    b->ent = CoverEntries{};
  }
  while (b->subs.size() < g) {
    // a group's sub-batch, stream and join event exist together or not at all
    hipStream_t q = nullptr;
    hipEvent_t e = nullptr;
    hipError_t he = hipStreamCreateWithFlags(&q, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    ocr_batch* sb = nullptr;
    int rc = he == hipSuccess ? ocr_batch_create(static_cast<ocr_ctx*>(b->ctx), b->max_views, &sb)
                              : fail(OCR_ERR_HIP, std::string("batch groups: ") + hipGetErrorString(he));
    if (rc) {
      if (e) (void)hipEventDestroy(e);
      if (q) (void)hipStreamDestroy(q);
      return rc;
    }
    b->subs.push_back(sb);
    b->sub_streams.push_back(q);
    b->joins.push_back(e);
  }
  if (!b->fork) OCR_HIP_CHECK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  const uint32_t n = (uint32_t)b->set_views.size();
  uint64_t total = 0;
  for (const ocr_view& v : b->set_views) total += (uint64_t)v.width * v.height;
  ocr_batch_tuning st = b->tune;
  st.groups = 1;
  uint32_t begin = 0;
  uint64_t acc = 0;
  for (uint32_t k = 0; k < g; k++) {
    // group k ends where the running pixel count reaches (k+1)/g of the
    // total, keeping at least one view for each later group
    uint32_t end = begin + 1;
    acc += (uint64_t)b->set_views[begin].width * b->set_views[begin].height;
    while (end < n - (g - 1 - k) && acc * g < total * (k + 1)) {
      acc += (uint64_t)b->set_views[end].width * b->set_views[end].height;
      end++;
    }
    if (k == g - 1) end = n;
    ocr_batch* sb = b->subs[k];
    int rc = ocr_batch_set_tuning(sb, &st);
    if (!rc) rc = ocr_batch_set_views(sb, &b->set_scenes[begin], &b->set_views[begin], &b->set_frames[begin], end - begin);
    if (rc) return rc;
    begin = end;
  }
  return OCR_OK;
}

// Renders the groups: each sub-batch on its own stream after `st`'s work so
// far, `st` after all of them.
static int render_groups(ocr_batch* b, hipStream_t st, bool graph) {
  OCR_HIP_CHECK(hipEventRecord(b->fork, st));
  for (uint32_t k = 0; k < group_count(b); k++) {
    OCR_HIP_CHECK(hipStreamWaitEvent(b->sub_streams[k], b->fork, 0));
    const int rc = graph ? ocr_batch_launch_graph(b->subs[k], b->sub_streams[k])
                         : ocr_batch_render(b->subs[k], b->sub_streams[k]);
    if (rc) return rc;
    OCR_HIP_CHECK(hipEventRecord(b->joins[k], b->sub_streams[k]));
    OCR_HIP_CHECK(hipStreamWaitEvent(st, b->joins[k], 0));
  }
  return OCR_OK;
}

extern "C" {

int ocr_batch_create(ocr_ctx* ctx, uint32_t max_views, ocr_batch** out) {
  if (!ctx || !out || max_views == 0 || max_views > kMaxViews)
    return fail(OCR_ERR_INVALID, "ocr_batch_create: bad argument");
  *out = nullptr;
  Ctx* c = static_cast<Ctx*>(ctx);
  OCR_HIP_CHECK(hipSetDevice(c->device));
  ocr_batch* b = new (std::nothrow) ocr_batch();
  if (!b) return fail(OCR_ERR_OOM, "ocr_batch_create: host allocation");
  b->ctx = c;
  b->max_views = max_views;
  hipError_t e = hipMalloc((void**)&b->d_views, sizeof(DevView) * max_views);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_scenes, sizeof(DevScene) * max_views);
  if (e == hipSuccess) e = hipMalloc((void**)&b->counters, 4 * kCounterWords);
  if (e == hipSuccess) e = hipMemset(b->counters, 0, 4 * kCounterWords);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->ds_host, sizeof(DevScene) * max_views);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ds_done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    ocr_batch_destroy(b);
    return fail(OCR_ERR_HIP, std::string("ocr_batch_create: ") + hipGetErrorString(e));
  }
  *out = b;
  return OCR_OK;
}

void ocr_batch_destroy(ocr_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipDeviceSynchronize();
  for (ocr_batch* sb : b->subs) ocr_batch_destroy(sb);
  for (hipStream_t q : b->sub_streams) (void)hipStreamDestroy(q);
  for (hipEvent_t e : b->joins) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  drop_graph(b);
  if (b->d_views) (void)hipFree(b->d_views);
  if (b->d_scenes) (void)hipFree(b->d_scenes);
  if (b->cuts) (void)hipFree(b->cuts);
  if (b->ent.ent) (void)hipFree(b->ent.ent);
  if (b->ent.cnt) (void)hipFree(b->ent.cnt);
  free_bins(b->tb, b->box_cap);
  if (b->counters) (void)hipFree(b->counters);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  if (b->ds_host) (void)hipHostFree(b->ds_host);
  if (b->ds_done) (void)hipEventDestroy(b->ds_done);
  for (hipEvent_t e : b->ev)
    if (e) (void)hipEventDestroy(e);
  delete b;
}

int ocr_batch_set_views(ocr_batch* b, ocr_scene* const* scenes, const ocr_view* views, ocr_frame* const* frames,
                        uint32_t n) {
  if (!b || !scenes || !views || !frames || n == 0 || n > b->max_views)
    return fail(OCR_ERR_INVALID, "ocr_batch_set_views: bad argument");
  OCR_HIP_CHECK(hipSetDevice(b->ctx->device));
  std::vector<Scene*> uniq;
  std::vector<DevView> dv(n);
  uint64_t sum_tris = 0, sum_pix = 0;
  for (uint32_t i = 0; i < n; i++) {
    Scene* s = static_cast<Scene*>(scenes[i]);
    Frame* f = static_cast<Frame*>(frames[i]);
    if (!s || !f) return fail(OCR_ERR_INVALID, "ocr_batch_set_views: null scene or frame");
    if (views[i].width != f->width || views[i].height != f->height)
      return fail(OCR_ERR_INVALID, "ocr_batch_set_views: view size differs from its frame");
    if (!(views[i].fov_const >= 0x1p-60f && views[i].fov_const <= 0x1p60f))
      return fail(OCR_ERR_INVALID, "ocr_batch_set_views: fov_const must be in [2^-60, 2^60]");
    if (!s->atlas.pages || (s->nlights && !s->lights))
      return fail(OCR_ERR_INVALID, "ocr_batch_set_views: scene needs its atlas and lights uploaded");
    if (s->nobj && s->max_tid >= std::max<uint32_t>(s->atlas.mip_start, 1))
      return fail(OCR_ERR_INVALID, "ocr_batch_set_views: an object's texture id is outside the atlas");
    for (uint32_t j = 0; j < i; j++)
      if (frames[j] == frames[i]) return fail(OCR_ERR_INVALID, "ocr_batch_set_views: a frame appears twice");
    if (f->use == Frame::kRender)
      return fail(OCR_ERR_INVALID, "ocr_batch_set_views: a frame was rendered by ocr_render (frames serve one or the other)");
    uint32_t si = 0;
    while (si < uniq.size() && uniq[si] != s) si++;
    if (si == uniq.size()) uniq.push_back(s);
    dv[i] = dev_view(&views[i], si, f);
    sum_tris += s->ntri;
    sum_pix += (uint64_t)views[i].width * views[i].height;
  }
  for (size_t i = 0; i < uniq.size(); i++) {
    if (!uniq[i]->lights) {
      int rc = ocr_lights_upload(static_cast<ocr_scene*>(uniq[i]), nullptr, 0, uniq[i]->lbuf_dim);
      if (rc) return rc;
    }
  }
  std::vector<DevScene> ds(uniq.size());
  std::vector<uint64_t> gens(uniq.size());
  {
    std::vector<std::unique_lock<std::mutex>> locks;
    lock_scenes(uniq, locks);
    for (size_t i = 0; i < uniq.size(); i++) {
      ds[i] = dev_scene(uniq[i]);
      gens[i] = uniq[i]->gen;
    }
  }
  for (uint32_t i = 0; i < n; i++) static_cast<Frame*>(frames[i])->use = Frame::kBatch;
  drop_graph(b);
  b->scenes = uniq;
  b->bound_gen = gens;
  if (b->ds_pending) {
    OCR_HIP_CHECK(hipEventSynchronize(b->ds_done));   // a refresh copy still reading the old table
    b->ds_pending = false;
  }
  b->host_views = dv;
  b->nviews = n;
  b->sum_tris = sum_tris;
  b->sum_pix = sum_pix;
  OCR_HIP_CHECK(hipMemcpy(b->d_scenes, ds.data(), sizeof(DevScene) * ds.size(), hipMemcpyHostToDevice));
  b->set_scenes.assign(scenes, scenes + n);
  b->set_views.assign(views, views + n);
  b->set_frames.assign(frames, frames + n);
  int rc = upload_views(b);
  if (rc) return rc;
  if (group_count(b) > 1) return regroup(b);   // the groups own the work lists
  return ensure_lists(b);
}

int ocr_batch_set_tuning(ocr_batch* b, const ocr_batch_tuning* t) {
  if (!b || !t || t->raster_blocks_per_cu == 0 || t->raster_blocks_per_cu > 32 || t->shade_tile_w_log2 > 8 ||
      t->groups > 8)
    return fail(OCR_ERR_INVALID, "ocr_batch_set_tuning: bad tuning");
  if (t->raster_tile) {
    const uint32_t a = t->raster_tile & 0xFFu, c = t->raster_tile >> 8;
    if (a < 3 || a > 6 || c < 3 || c > 6 || a + c < 8 || a + c > 11)
      return fail(OCR_ERR_INVALID, "ocr_batch_set_tuning: raster tiles are 2^3..2^6 on a side and 256..2048 pixels");
  }
  b->tune = *t;
  b->tune.groups = std::max<uint32_t>(t->groups, 1);
  drop_graph(b);
  if (!b->nviews) return OCR_OK;
  int rc = upload_views(b);
  if (!rc && group_count(b) > 1) rc = regroup(b);
  if (!rc && group_count(b) == 1) rc = ensure_lists(b);
  return rc;
}

int ocr_batch_get_tuning(ocr_batch* b, ocr_batch_tuning* t) {
  if (!b || !t) return fail(OCR_ERR_INVALID, "ocr_batch_get_tuning: bad argument");
  *t = b->tune;
  return OCR_OK;
}

int ocr_batch_set_profiling(ocr_batch* b, int enable) {
  if (!b) return fail(OCR_ERR_INVALID, "ocr_batch_set_profiling: null batch");
  b->profiling = enable != 0;
  b->n_profiled = 0;
  if (b->profiling && b->ev.empty()) {
    OCR_HIP_CHECK(hipSetDevice(b->ctx->device));
    b->ev.assign(2 * kBatchRing, nullptr);
    for (auto& e : b->ev) OCR_HIP_CHECK(hipEventCreate(&e));
  }
  return OCR_OK;
}

int ocr_batch_render(ocr_batch* b, void* stream) {
  if (!b || b->nviews == 0) return fail(OCR_ERR_INVALID, "ocr_batch_render: no views");
  OCR_HIP_CHECK(hipSetDevice(b->ctx->device));
  hipStream_t st = (hipStream_t)stream;
  if (group_count(b) > 1) {
    const bool prof = b->profiling && b->n_profiled < kBatchRing;
    if (prof) OCR_HIP_CHECK(hipEventRecord(b->ev[2 * b->n_profiled], st));
    const int rg = render_groups(b, st, false);
    if (rg) return rg;
    if (prof) OCR_HIP_CHECK(hipEventRecord(b->ev[2 * b->n_profiled++ + 1], st));
    return OCR_OK;
  }
  int rc = ensure_lists(b);
  if (rc) return rc;
  std::vector<std::unique_lock<std::mutex>> locks;
  lock_scenes(b->scenes, locks);
  if ((rc = check_scenes(b))) return rc;
  OCR_HIP_CHECK(bind_scenes(b, st));
  const bool prof = b->profiling && b->n_profiled < kBatchRing;
  if (prof) OCR_HIP_CHECK(hipEventRecord(b->ev[2 * b->n_profiled], st));
  OCR_HIP_CHECK(enqueue_batch(b, st));
  OCR_HIP_CHECK(note_scene_reads(b, st));
  if (prof) OCR_HIP_CHECK(hipEventRecord(b->ev[2 * b->n_profiled++ + 1], st));
  return OCR_OK;
}

int ocr_batch_launch_graph(ocr_batch* b, void* stream) {
  if (!b || b->nviews == 0) return fail(OCR_ERR_INVALID, "ocr_batch_launch_graph: no views");
  OCR_HIP_CHECK(hipSetDevice(b->ctx->device));
  if (group_count(b) > 1) {
    hipStream_t gst = (hipStream_t)stream;
    const bool prof = b->profiling && b->n_profiled < kBatchRing;
    if (prof) OCR_HIP_CHECK(hipEventRecord(b->ev[2 * b->n_profiled], gst));
    const int rg = render_groups(b, gst, true);
    if (rg) return rg;
    if (prof) OCR_HIP_CHECK(hipEventRecord(b->ev[2 * b->n_profiled++ + 1], gst));
    return OCR_OK;
  }
  int rc = ensure_lists(b);
  if (rc) return rc;
  // the graph holds one shading kernel: capture again when a light upload
  // changed whether any scene casts shadows
  if (b->exec && batch_pass(b).unshadowed != b->graph_unshadowed) drop_graph(b);
  if (!b->exec) {
    // Capture on the batch's own stream (the legacy stream cannot be
    // captured); the passes use fill kernels, never memset nodes.
    OCR_HIP_CHECK(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue_batch(b, b->capture_stream);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->capture_stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(OCR_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    b->graph = g;
    b->graph_unshadowed = batch_pass(b).unshadowed;
    OCR_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  hipStream_t st = (hipStream_t)stream;
  std::vector<std::unique_lock<std::mutex>> locks;
  lock_scenes(b->scenes, locks);
  if ((rc = check_scenes(b))) return rc;
  OCR_HIP_CHECK(bind_scenes(b, st));   // the graph reads d_scenes when it runs
  const bool prof = b->profiling && b->n_profiled < kBatchRing;
  if (prof) OCR_HIP_CHECK(hipEventRecord(b->ev[2 * b->n_profiled], st));
  OCR_HIP_CHECK(hipGraphLaunch(b->exec, st));
  OCR_HIP_CHECK(note_scene_reads(b, st));
  if (prof) OCR_HIP_CHECK(hipEventRecord(b->ev[2 * b->n_profiled++ + 1], st));
  return OCR_OK;
}

int ocr_batch_get_stats(ocr_batch* b, void* stream, ocr_frame_stats* out) {
  if (!b || !out) return fail(OCR_ERR_INVALID, "ocr_batch_get_stats: bad argument");
  OCR_HIP_CHECK(hipSetDevice(b->ctx->device));
  OCR_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  uint32_t c[kCounterWords];
  memset(c, 0, sizeof(c));
  if (group_count(b) > 1) {
    // the groups' counters, summed (their streams joined `stream`)
    for (uint32_t k = 0; k < group_count(b); k++) {
      uint32_t g[kCounterWords];
      OCR_HIP_CHECK(hipMemcpy(g, b->subs[k]->counters, sizeof(g), hipMemcpyDeviceToHost));
      c[kPacked + 1] += g[kPacked + 1];
      c[kFragCount] += g[kFragCount];
      c[kOverflow] |= g[kOverflow];
      // a group whose bins overflowed grows them for its next render
      if (g[kOverflow] & kOverflowBins)
        b->subs[k]->want_bin = (uint32_t)std::min<uint64_t>((uint64_t)g[kBinTotal] + g[kBinTotal] / 4 + 64, 0xFFFFFFF0ull);
    }
  } else {
    OCR_HIP_CHECK(hipMemcpy(c, b->counters, sizeof(c), hipMemcpyDeviceToHost));
    if (c[kOverflow] & kOverflowBins)
      b->want_bin = (uint32_t)std::min<uint64_t>((uint64_t)c[kBinTotal] + c[kBinTotal] / 4 + 64, 0xFFFFFFF0ull);
  }
  memset(out, 0, sizeof(*out));
  out->n_cut = c[kPacked + 1];
  out->n_frag = c[kFragCount];
  out->overflow = c[kOverflow] ? 1u : 0u;
  out->n_profiled = (uint32_t)b->n_profiled;
  if (b->n_profiled) {
    double sum = 0;
    for (int r = 0; r < b->n_profiled; r++) {
      float ms = 0.f;
      OCR_HIP_CHECK(hipEventElapsedTime(&ms, b->ev[2 * r], b->ev[2 * r + 1]));
      sum += ms;
    }
    out->ms_total = (float)(sum / b->n_profiled);
  }
  if (c[kOverflow] & kOverflowSamples)
    return fail(OCR_ERR_UNSUPPORTED, "ocr_batch_get_stats: the batch needs more than 2^31 - 1 raster samples");
  if (c[kOverflow]) return fail(OCR_ERR_OVERFLOW, "ocr_batch_get_stats: a work list overflowed (it grows for the next render)");
  return OCR_OK;
}

// Times candidate tunings on the current views (graph launches, after one
// warm-up each) and keeps the fastest: raster blocks per CU x shading tile
// width x view groups with the sample-space raster passes, then, with the
// best of those, the tiled raster path at each raster tile shape.
int ocr_batch_autotune(ocr_batch* b, void* stream, uint32_t reps, ocr_batch_tuning* best_out) {
  if (!b || b->nviews == 0) return fail(OCR_ERR_INVALID, "ocr_batch_autotune: no views");
  OCR_HIP_CHECK(hipSetDevice(b->ctx->device));
  if (reps == 0) reps = 5;
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0, e1;
  OCR_HIP_CHECK(hipEventCreate(&e0));
  OCR_HIP_CHECK(hipEventCreate(&e1));
  const bool prof = b->profiling;
  b->profiling = false;
  // one candidate: set it, one warm-up launch, then reps timed launches
  auto time_one = [&](const ocr_batch_tuning& t, float& ms) -> int {
    int status = ocr_batch_set_tuning(b, &t);
    if (!status) status = ocr_batch_launch_graph(b, stream);
    if (status) return status;
    hipError_t e = hipSuccess;
    if (group_count(b) > 1) {
      e = hipEventRecord(e0, st);
      for (uint32_t r = 0; r < reps && e == hipSuccess && status == OCR_OK; r++)
        status = ocr_batch_launch_graph(b, stream);
      if (status) return status;
    } else {
      // the timed replays read the scenes like any launch
      std::vector<std::unique_lock<std::mutex>> locks;
      lock_scenes(b->scenes, locks);
      if ((status = check_scenes(b))) return status;
      e = bind_scenes(b, st);
      if (e == hipSuccess) e = hipEventRecord(e0, st);
      for (uint32_t r = 0; r < reps && e == hipSuccess; r++) e = hipGraphLaunch(b->exec, st);
      if (e == hipSuccess) e = note_scene_reads(b, st);
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) return fail(OCR_ERR_HIP, std::string("ocr_batch_autotune: ") + hipGetErrorString(e));
    // a candidate whose lists overflowed is not a result (they grew)
    ocr_frame_stats fs;
    status = ocr_batch_get_stats(b, stream, &fs);
    if (status == OCR_ERR_OVERFLOW) {
      ms = 1e30f;
      status = OCR_OK;
    }
    return status;
  };
  const uint32_t blocks[] = {2, 4, 8, 16};
  const uint32_t tiles[] = {3, 4, 5, 6};     // shading tile widths 8, 16, 32, 64
  const uint32_t groups[] = {1, 2, 4};
  // raster tiles (w_log2 | h_log2 << 8): 16x16, 32x16, 16x32, 32x32, 64x32
  const uint32_t raster_tiles[] = {4u | 4u << 8, 5u | 4u << 8, 4u | 5u << 8, 5u | 5u << 8, 6u | 5u << 8};
  ocr_batch_tuning best = b->tune;
  float best_ms = 1e30f;
  int status = OCR_OK;
  for (uint32_t gr : groups) {
    if (gr > b->nviews || status) break;
    for (uint32_t bl : blocks)
      for (uint32_t tl : tiles) {
        if (status) break;
        const ocr_batch_tuning t{bl, tl, gr, 0};
        float ms = 0.f;
        status = time_one(t, ms);
        if (!status && ms < best_ms) {
          best_ms = ms;
          best = t;
        }
      }
  }
  const ocr_batch_tuning base = best;
  for (uint32_t rt : raster_tiles) {
    if (status) break;
    ocr_batch_tuning t = base;
    t.raster_tile = rt;
    float ms = 0.f;
    status = time_one(t, ms);
    if (!status && ms < best_ms) {
      best_ms = ms;
      best = t;
    }
  }
  b->profiling = prof;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (status) return status;
  if (best_out) *best_out = best;
  return ocr_batch_set_tuning(b, &best);
}

}  // extern "C"
