// ocr_internal.h — device records, handles and launch helpers shared by the
// runtime and the kernels.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ocr.h"
#include "ocr_math.h"

namespace ocr {

// Triangle positions + owning object: the 40 bytes prearrange and the raster
// passes need (vertex_pos + vertices[0].object_id, cl2.cl:4300-4304).
struct alignas(8) TriPos {
  float p[9];
  uint32_t oid;
};
static_assert(sizeof(TriPos) == 40, "TriPos layout");

// Shading attributes of a triangle (normals, vt, vertex colours), read only
// for visible triangles by the shading pass (cl2.cl:5872-5938).
struct TriAttr {
  float n[9];
  float vt[6];
  uint32_t col[3];
};
static_assert(sizeof(TriAttr) == 72, "TriAttr layout");

// The obj_g_descriptor fields the path reads (52 bytes), and what kernel3
// derives from them identically for every pixel of the object, formed once
// per descriptor upload with the same operations (pack_obj):
//   qn  = q * (1 / sqrt(q.q))                 rot_quat's normalised quaternion
//   qb  = n(c / (c.c)), c = conj(q)           back_rot_quat's, normalised by rot_quat
//   rough2 = r * r, r = clamp(1 - specular, 0.001, 10), and RN(1 / rough2)
// (cl2.cl:350-370, :6232-6240).  96 bytes.
struct alignas(16) ObjDesc {
  float wp[3];
  float scale;
  float q[4];
  uint32_t tid;
  float specular;
  float spec_mult;
  float diffuse;
  int32_t flags;
  float rough2;
  float inv_rough2;
  uint32_t _pad;
  float qn[4];
  float qb[4];
};
static_assert(sizeof(ObjDesc) == 96, "ObjDesc layout");

// A clipped, culled, projected triangle: screen x, y and camera z of its
// three vertices, its 1/det constant, its sort key (2*tri + sub for a camera
// view, cube face for a shadow map), object id, view, and for camera views
// the screen-space plane of 1/z (get_vtdiff's interpolate_get_const,
// cl2.cl:5691-5760), which shading needs once per visible triangle.  Records
// are in sample order: prearrange allocates records and sample ranges
// together, once per block.
struct alignas(16) CutTri {
  float x[3];
  float y[3];
  float z[3];
  float rconst;
  uint32_t key;
  uint32_t oid;
  uint32_t view;
  float da, db, dc;   // 1/z plane (0 for shadow faces)
};
static_assert(sizeof(CutTri) == 64, "CutTri layout");

// The raster setup of a cut triangle, computed once by prearrange and read
// by every raster pass: bounding-box origin and width, 1/width, the 1/z plane
// (cl2.cl:5044) and point_in_tri's loop-invariant terms with the edge sign
// folded in (negation is exact, so a*sign + (b*sign)*x + (c*sign)*y has the
// bits of (a + b*x + c*y) * sign).
struct alignas(16) RasterTri {
  float min_x;
  uint32_t key;     // | kKeyBoxForm when the samples are the box's pixels
  float min_y;
  float w;          // max_x - min_x
  float iw;         // 1 / w
  float A, B, C;    // 1/z plane
  float ek[7];      // s0, s1, s2, t0, t1, t2 (signed), 2.0001 * |area|
  uint32_t sample_begin;
};
static_assert(sizeof(RasterTri) == 64, "RasterTri layout");
// Set in RasterTri::key (keys are below 2^31) when the record's samples are
// its bounding box's pixels in row-major order, w * rows of them, instead of
// the walk's: for triangles whose walk visits exactly that box
// (walk_is_rect).  The depth and id passes keep a minimum and a maximum per
// pixel, so only the set of pixels a triangle visits matters, not the order
// or the repeats (the walk visits every fragment's first index twice).
constexpr uint32_t kKeyBoxForm = 1u << 31;

// Everything a kernel needs to know about one view.
struct ViewParams {
  uint32_t width, height;
  float fov;
  uint32_t linear;
  float cpos[3];
  float ifov;               // RN(1 / fov); views keep fov in [2^-60, 2^60]
  SinCos cam;
  float clear[4];
};

// One scene as the kernels see it.
struct DevScene {
  const TriPos* tri_pos;
  const TriPos* world;      // world-space vertices (object transform applied)
  const TriAttr* tri_attr;
  const ObjDesc* objs;
  const uint8_t* pages;
  const uint32_t* nums;
  const uint32_t* sizes;
  const ocr_light* lights;
  const float* light_lin;   // per light: gamma_fwd of its colour (x, y, z, 0)
  const uint32_t* shadow_maps;
  const uint32_t* static_maps;
  uint32_t ntri, nobj, mip_start, nlights, L, _pad;
};

// One view of a render.  Per-view work ranges start on 256-slot boundaries
// so every 256-thread block of the clear, prearrange and shade launches
// belongs to exactly one view.
struct DevView {
  ViewParams v;
  uint32_t scene;          // index into the scene table
  uint32_t tri_begin;      // first triangle slot of this view (multiple of 256)
  uint32_t pix_begin;      // first pixel slot (multiple of 256)
  uint32_t tile_begin;     // first 16x16 shading tile
  uint32_t bin_begin;      // first raster tile (tiled raster path)
  uint32_t bin_tiles_x;    // raster tiles per row
  uint32_t* depth;
  unsigned long long* idkey;
  uint32_t* rgba;
  uint32_t* normals;
};

// Screen-tile bins of the tiled raster path (tiles.hip): every cut record
// is listed in each tile of tw x th pixels its bounding box touches.
// Prearrange writes each record's box and counts the tiles' entries, a scan
// turns the counts into offsets, a scatter fills the bins, and one
// workgroup per tile resolves its depth and ids in LDS.
struct TileBins {
  uint32_t* cnt;      // per tile: entries (zeroed before prearrange)
  uint32_t* off;      // per tile: first entry; off[ntiles] = total
  uint32_t* cur;      // per tile: scatter cursor
  uint32_t* sums;     // per scan chunk: its tiles' entries
  uint32_t* bin;      // cut-record indices, grouped by tile
  uint4* box;         // per cut record: see TileBox
  uint32_t ntiles, cap_bin;
  uint32_t tw_log2, th_log2;
};
// A cut record's tiles: x = tx0 | tx1 << 16, y = ty0 | ty1 << 16 (inclusive
// tile columns / rows of its view), z = the view's first tile, w = max_y |
// tiles_x << 16 | walk_is_rect << 31.
constexpr uint32_t kBoxRect = 1u << 31;

// Covered-sample entries of a render: the depth pass writes, per chunk of
// kChunk raster samples, each covered sample that can still fall in the id
// pass's window ({x | y << 16, record}), up to kEntCap, and the chunk's
// count; the id pass reads the entries, forms each one's depth again from
// its record's plane, instead of sweeping the chunk's samples again (chunks
// with more, or past `chunks`, are swept).
// 2048-sample chunks: half the wavefronts of 1024-sample ones, which leaves
// room for the other renders' shading beside the raster passes (1080p/1M
// frame 0.183 -> 0.178 ms, though the depth pass alone takes 84 us, not 75;
// 512 and 4096 measured slower).
constexpr uint32_t kChunk = 64 * 32;   // raster samples per wavefront work item
constexpr uint32_t kEntCap = kChunk / 2;
struct CoverEntries {
  uint2* ent;         // chunks x kEntCap
  uint32_t* cnt;      // per chunk
  uint32_t chunks;
};

// The cut list and counters shared by the views of one render.  One
// allocation of 2 * cap_cut records holds the CutTri records followed by
// their RasterTri setups.  tb.bin is null unless the render is tiled.
struct Lists {
  CutTri* cuts;
  RasterTri* rast;
  uint32_t* counters;
  uint32_t cap_cut;
  uint32_t count_frags;   // keep the reference's fragment count (statistics)
  TileBins tb;
  CoverEntries ent;
};
inline Lists make_lists(CutTri* cuts, uint32_t* counters, uint32_t cap_cut, bool count_frags = true) {
  return Lists{cuts, reinterpret_cast<RasterTri*>(cuts + cap_cut), counters, cap_cut, count_frags ? 1u : 0u};
}
constexpr uint32_t kMaxViews = 4096;

// Everything the camera passes need.  A single view travels inline in the
// kernel arguments (views == nullptr); a batch reads device tables.
struct PassArgs {
  DevView view0;
  DevScene scene0;
  const DevView* views;
  const DevScene* scenes;
  uint32_t nviews;
  uint32_t tri_slots, pix_slots, tiles;
  uint32_t tile_w_log2;    // shading tile = (1 << tile_w_log2) x (256 >> tile_w_log2)
  uint32_t unshadowed;     // no scene of the pass has a shadow-casting light
  Lists lists;
};

// Shading tiles of one view for a tile width of 1 << log2 (256 pixels each).
__host__ __device__ inline uint32_t view_tiles(uint32_t w, uint32_t h, uint32_t log2w) {
  const uint32_t tw = 1u << log2w, th = 256u >> log2w;
  return ((w + tw - 1) / tw) * ((h + th - 1) / th);
}

// Device counters of one render: words 0-1 hold a u64 whose high half is the
// cut-triangle count and low half the raster sample count.  Every block of
// prearrange adds to it; atomics on one cache line serialise (about 12 ns
// each), so the fragment statistic lives on a line of its own.
enum Counter : int {
  kPacked = 0,
  kOverflow = 2,
  kValidEnd = 16,          // shadow passes: u64, end of the list's valid prefix, byte 64
  kFragCount = 32,         // reference fragments (statistics), byte 128
  kBinTotal = 48,          // tiled path: bin entries the render needed, byte 192
  kCounterWords = 64,
};
// Shadow-pass lists are split into kShadowParts partitions (one per lane of
// the raster pass's wavefronts), each with its own record range, sample
// space, packed counter and valid-prefix word on a 128-byte line of its own
// after the render counters: prearrange block b allocates in partition
// b mod kShadowParts.  (One counter for the whole pass serialised the
// prearrange's wavefront allocations on one address.)
constexpr uint32_t kShadowParts = 64;
constexpr int kPartLineWords = 32;
constexpr int kShadowCounterWords = kCounterWords + (int)kShadowParts * kPartLineWords;
// partition p: packed counter (u64) at word kPartWord(p), valid prefix at +2
__host__ __device__ constexpr int kPartWord(uint32_t p) { return kCounterWords + (int)p * kPartLineWords; }
// Bits of the overflow word: the cut list is full / the sample space is full.
constexpr uint32_t kOverflowCuts = 1u, kOverflowSamples = 2u, kOverflowBins = 4u;
// Raster samples per render, and per cut triangle.  Frames and shadow faces
// hold at most 16,000,000 pixels, so a triangle has at most
// area * (1 + 1/op) + op + 1 < 2^24 samples: sample indices, bounding-box
// products and row starts are exact in binary32.
constexpr uint32_t kMaxSamples = 0x7FFFFFFFu;
constexpr uint32_t kMaxTriSamples = (1u << 24) - 1;
constexpr uint64_t kMaxTargetPixels = 16000000ull;

// Shadow-pass lists pack each partition's counter as (records << 40 |
// samples): a wavefront allocates at most 2^25 samples at once (a lane's two
// pieces hold fewer than 2^25) and only if the counter it loads first is
// below the limits, and at most 8192 wavefronts are resident, so the sample
// part stays below 2^30 + 2^13 * 2^25 < 2^40 and the record part below
// kShadowMaxRecords + 2^13 * 128 < 2^24.
constexpr int kShadowSampleBits = 40;
constexpr unsigned long long kShadowSampleMask = (1ull << kShadowSampleBits) - 1;
constexpr uint32_t kShadowWaveSamples = 1u << 25;
constexpr uint32_t kShadowMaxSamples = 1u << 30;
constexpr uint32_t kShadowMaxRecords = 14u << 20;

struct Ctx {
  int device = 0;
  hipDeviceProp_t prop{};
  SinCos face_sc[6];
  // device address of the no-allocation render's scratch on this device
  // (hipGetSymbolAddress at ocr_create)
  unsigned char* noalloc_scratch = nullptr;
};

struct Atlas {
  uint8_t* pages = nullptr;   // npages x 2048 x 2048 x RGBA8
  uint32_t* nums = nullptr;
  uint32_t* sizes = nullptr;
  uint32_t npages = 0, nnums = 0, nsizes = 0, mip_start = 0;
};

struct Scene {
  Ctx* ctx = nullptr;
  uint32_t ntri = 0, nobj = 0;
  uint32_t max_tid = 0;              // largest descriptor texture id (validated at render)
  TriPos* tri_pos = nullptr;
  TriPos* world = nullptr;           // world-space copy, refreshed when objects change
  TriAttr* tri_attr = nullptr;
  ObjDesc* objs = nullptr;
  // motion-blur poses, 16 floats per object: old_world_pos_1, old_world_pos_2,
  // old_world_rot_quat_1, old_world_rot_quat_2 (obj_g_descriptor.hpp:15-18);
  // set at creation, then written only by the motion-blur pass
  float* motion = nullptr;
  Atlas atlas;
  ocr_light* lights = nullptr;
  float* light_lin = nullptr;        // 4 floats per light
  uint32_t nlights = 0, n_shadow = 0, n_static = 0, lbuf_dim = 1024;
  std::vector<ocr_light> host_lights;
  uint32_t* shadow_maps = nullptr;   // n_shadow x 6 x dim^2
  uint32_t* static_maps = nullptr;   // n_static x 6 x dim^2
  bool static_dirty = false;
  // work list of the shadow passes (raster setups only)
  RasterTri* sh_rast = nullptr;
  uint32_t* sh_counters = nullptr;
  uint32_t sh_cap_cut = 0;
  // recorded on the stream of the last shadow pass after its last launch;
  // renders wait for it (they run on streams of their frame)
  hipEvent_t written = nullptr;
  std::atomic<bool> written_valid{false};
  // recorded after the last stream-ordered descriptor update and its
  // world-space refresh; renders and shadow passes wait for it before
  // reading the geometry
  hipEvent_t geom_written = nullptr;
  std::atomic<bool> geom_valid{false};
  // pinned staging of stream-ordered updates: a slot is reused once the copy
  // that read it has completed
  static constexpr int kUpdRing = 4;
  ObjDesc* upd_host[kUpdRing] = {};
  hipEvent_t upd_done[kUpdRing] = {};
  bool upd_used[kUpdRing] = {};
  int upd_next = 0;
  // Geometry versions.  objs / world above alias objs_v[ver] / world_v[ver].
  // A stream-ordered update builds the other version on upd_stream, after
  // the passes that read that version (its readers) have finished, and then
  // makes it current, so it does not wait for the passes still reading the
  // current one.  Every pass that reads objs or world enqueues its kernels
  // and records its reader event with ver_mu held (a server's dispatcher
  // thread renders while the application updates).
  struct Reader {
    hipStream_t stream;
    hipEvent_t ev;   // recorded on `stream` after the stream's last pass that read the version
  };
  ObjDesc* objs_v[2] = {};
  TriPos* world_v[2] = {};
  int ver = 0;
  // bumped whenever a buffer the kernels reach through DevScene is replaced
  // (atlas or light upload, geometry version), so that scene tables kept in
  // device memory (batches) are rebuilt before their next launch
  uint64_t gen = 0;
  std::vector<Reader> readers[2];
  std::vector<hipEvent_t> ev_pool;
  hipStream_t upd_stream = nullptr;
  std::mutex ver_mu;
  // Dynamic shadow-map versions, the same way: shadow_maps aliases
  // maps_v[mver]; a shadow pass renders the other version on sh_stream once
  // the passes that read it (map_readers) have finished, so it overlaps the
  // shading of the render before it.  Static maps have one version (they
  // are rendered only after a light upload).
  uint32_t* maps_v[2] = {};
  int mver = 0;
  std::vector<Reader> map_readers[2];
  hipStream_t sh_stream = nullptr;
  hipError_t note_into(std::vector<Reader>& list, hipStream_t st) {
    for (Reader& r : list)
      if (r.stream == st) return hipEventRecord(r.ev, st);
    hipEvent_t e = nullptr;
    if (!ev_pool.empty()) {
      e = ev_pool.back();
      ev_pool.pop_back();
    } else {
      hipError_t c = hipEventCreateWithFlags(&e, hipEventDisableTiming);
      if (c != hipSuccess) return c;
    }
    list.push_back(Reader{st, e});
    return hipEventRecord(e, st);
  }
  // Records that work enqueued on `st` so far reads the current geometry
  // version (and, for passes that shade, the current shadow maps).
  hipError_t note_read(hipStream_t st, bool maps = false) {
    hipError_t e = note_into(readers[ver], st);
    if (e == hipSuccess && maps) e = note_into(map_readers[mver], st);
    return e;
  }
  // Makes `st` wait for every pass recorded in `list`, then forgets them.
  hipError_t wait_readers(std::vector<Reader>& list, hipStream_t st) {
    for (const Reader& r : list) {
      hipError_t e = hipStreamWaitEvent(st, r.ev, 0);
      if (e != hipSuccess) return e;
      ev_pool.push_back(r.ev);
    }
    list.clear();
    return hipSuccess;
  }
  // what the render and batch paths wait on before they read the scene
  hipError_t wait_geometry(hipStream_t st) const {
    return geom_valid ? hipStreamWaitEvent(st, geom_written, 0) : hipSuccess;
  }
  hipError_t wait_shadow_maps(hipStream_t st) const {
    return written_valid ? hipStreamWaitEvent(st, written, 0) : hipSuccess;
  }
};

struct Frame {
  Ctx* ctx = nullptr;
  uint32_t width = 0, height = 0;
  // Output and cut-list sets: render N uses set N mod kSets on that set's
  // stream, so it can overlap the renders before it.  The pointers below
  // alias the set of the last render.
  struct OutBuf {
    uint32_t* depth = nullptr;
    unsigned long long* idkey = nullptr;
    uint32_t* rgba = nullptr;
    uint32_t* normals = nullptr;
    CutTri* cuts = nullptr;
    uint32_t* counters = nullptr;
    uint32_t cap_cut = 0;
    // tiled raster path (tiles.hip): bins of this set's cut list
    TileBins tb{};
    uint32_t box_cap = 0;   // records the box array holds
    // covered-sample entries of the sample-space depth pass
    CoverEntries ent{};
  };
  // three renders in flight: 1080p/1M frame 0.186 -> 0.183 ms (four: 0.197)
  static constexpr int kSets = 3;
  OutBuf out[kSets];
  int cur = 0;
  uint32_t* depth = nullptr;
  unsigned long long* idkey = nullptr;
  uint32_t* rgba = nullptr;
  uint32_t* normals = nullptr;
  hipStream_t ss[kSets] = {};   // the streams renders of each set run on
  hipEvent_t entry[kSets] = {}; // caller's stream when a render of the set was issued
  hipEvent_t done[kSets] = {};  // a render of the set finished
  bool entry_valid[kSets] = {};
  uint32_t want_cut = 0;             // capacity set by ocr_frame_reserve (0 = automatic)
  // raster path of ocr_render (ocr_frame_set_raster): the sample space or
  // screen tiles of 2^tw_log2 x 2^th_log2 pixels, with bins of want_bin
  // entries (0 = automatic; raised after an overflow)
  uint32_t raster_mode = OCR_RASTER_SAMPLES, tw_log2 = 4, th_log2 = 4, want_bin = 0;
  uint32_t* post_in = nullptr;       // colour snapshot of the screen-space post passes
  int profiling = 0;   // 0 off, 1 every pass, 2 the shading pass only
  static constexpr int kEvents = 7;       // pass boundaries, and the shading pass's start
  static constexpr int kRing = 1024;      // renders averaged per profiling window
  std::vector<hipEvent_t> ev;             // kRing x kEvents, created on first use
  int n_profiled = 0;
  bool rendered = false;
  // A frame is written either by ocr_render or by batches, never both: a
  // batch writes the output set that was current when its views were set,
  // with the batch's own cut list and counters, while ocr_render alternates
  // sets and decodes through the frame's lists.  The first of the two to
  // take the frame fixes its use; the other refuses it.
  enum Use : int { kUnused = 0, kRender = 1, kBatch = 2 };
  int use = kUnused;
  void select(int b) {
    cur = b;
    depth = out[b].depth;
    idkey = out[b].idkey;
    rgba = out[b].rgba;
    normals = out[b].normals;
  }
};

// Allocates output set b of a frame (idempotent); counters zeroed.
int frame_alloc_set(Frame* f, int b);
// Tile bins for a cut list of cap_cut records over `ntiles` tiles of
// 2^twl x 2^thl pixels with room for cap_bin entries: (re)allocates what is
// missing or too small (synchronising the device first), zeroes the counts.
int ensure_bins(TileBins& tb, uint32_t& box_cap, uint32_t cap_cut, uint32_t ntiles, uint32_t twl, uint32_t thl,
                uint32_t cap_bin, bool exact);
void free_bins(TileBins& tb, uint32_t& box_cap);
hipError_t launch_bin_reset(const TileBins& tb, uint32_t* counters, hipStream_t s);
uint32_t bin_scan_chunks(uint32_t ntiles);

// ---- error plumbing -------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
#define OCR_HIP_CHECK(expr)                                                     \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess)                                                       \
      return ::ocr::fail(OCR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---- kernel launchers (defined next to their kernels) --------------------
hipError_t launch_clear(const PassArgs& a, hipStream_t s);
hipError_t launch_prearrange(const PassArgs& a, hipStream_t s);
hipError_t launch_depth(const PassArgs& a, int blocks_per_cu, hipStream_t s);
hipError_t launch_id(const PassArgs& a, int blocks_per_cu, hipStream_t s);
// Tiled raster path: bin scan and scatter, then depth and ids per tile.
hipError_t launch_bins(const PassArgs& a, hipStream_t s);
hipError_t launch_tile_raster(const PassArgs& a, hipStream_t s);
// Raster tiles of a view for tiles of 2^twl x 2^thl pixels.
__host__ __device__ inline uint32_t bin_tiles_x(uint32_t w, uint32_t twl) { return (w + (1u << twl) - 1) >> twl; }
__host__ __device__ inline uint32_t bin_tiles(uint32_t w, uint32_t h, uint32_t twl, uint32_t thl) {
  return bin_tiles_x(w, twl) * ((h + (1u << thl) - 1) >> thl);
}
// sin/cos of a rotation's three angles: binary64 evaluation rounded once
void sincos_of(const float ang[3], SinCos& sc);
hipError_t launch_shade(const PassArgs& a, const SinCos face[6], hipStream_t s);
hipError_t launch_shade_count(const PassArgs& a, const SinCos face[6], uint32_t* tex_bits,
                              uint32_t* shadow_bits, uint32_t* static_bits, hipStream_t s);
// helpers shared by the plain and batched renders (runtime.hip)
DevScene dev_scene(const Scene* sc);
// Sizes a batch's cut list for `sum_tris` scene triangles ahead of use.
int batch_reserve(ocr_batch* b, uint64_t sum_tris);
DevView dev_view(const ocr_view* v, uint32_t scene, Frame* f);
int device_cu_count();
// The cube maps of up to 16 lights in one prearrange and one raster launch:
// light i at lpos[i] writes the 6 faces of map slots[i] of `maps`.
hipError_t launch_shadow_group(Scene* sc, const float (*lpos)[3], const uint32_t* slots, uint32_t n,
                               int only_static, uint32_t* maps, hipStream_t s);
hipError_t launch_atlas_upload(const Atlas& a, uint32_t tid, const uint8_t* src, uint32_t w,
                               uint32_t h, hipStream_t s);
hipError_t launch_mip(const Atlas& a, uint32_t src_id, uint32_t dst_id, uint32_t gw, uint32_t gh,
                      hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s);
hipError_t launch_zero_counters(uint32_t* counters, hipStream_t s);
hipError_t launch_light_prep(const ocr_light* lights, uint32_t n, float* lin, hipStream_t s);
hipError_t launch_world(Scene* sc, hipStream_t s);

constexpr int kWave = 64;

}  // namespace ocr

// The C ABI's opaque handles are the internal records.
struct ocr_ctx : public ocr::Ctx {};
struct ocr_scene : public ocr::Scene {};
struct ocr_frame : public ocr::Frame {};
