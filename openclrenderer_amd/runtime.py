"""Python host binding of the C ABI (include/ocr.h) over ctypes.

The library is the in-tree ``openclrenderer_amd/libocr.so`` built by
``openclrenderer_amd.build``.  There is no fallback: if the library is missing
or a HIP call fails, these wrappers raise.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libocr.so")

# Every entry point include/ocr.h declares.
EXPORTS = [
    "ocr_create", "ocr_destroy", "ocr_last_error", "ocr_query_free_hbm",
    "ocr_scene_create", "ocr_scene_update_objs", "ocr_scene_update_objs_async", "ocr_atlas_upload", "ocr_lights_upload",
    "ocr_shadow_pass", "ocr_scene_destroy", "ocr_scene_atlas_info", "ocr_scene_read_atlas",
    "ocr_scene_read_shadow_maps", "ocr_scene_read_motion", "ocr_frame_create", "ocr_frame_destroy",
    "ocr_frame_set_profiling", "ocr_render", "ocr_render_noalloc", "ocr_frame_read", "ocr_frame_device_buffers",
    "ocr_frame_blend_with_depth", "ocr_frame_pseudo_aa", "ocr_frame_motion_blur",
    "ocr_frame_get_stats", "ocr_frame_count_texels", "ocr_frame_reserve", "ocr_frame_set_raster",
    "ocr_frame_get_raster",
    "ocr_batch_create", "ocr_batch_destroy", "ocr_batch_set_views", "ocr_batch_set_tuning",
    "ocr_batch_get_tuning", "ocr_batch_set_profiling", "ocr_batch_render", "ocr_batch_launch_graph",
    "ocr_batch_get_stats", "ocr_batch_autotune",
    "ocr_server_create", "ocr_server_add_scene", "ocr_server_submit", "ocr_server_poll",
    "ocr_server_release", "ocr_server_get_info", "ocr_server_now_ns", "ocr_server_destroy",
]


class OcrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ocr status {code}: {msg}")
        self.code = code


class TexSrc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.c_void_p)]


class BatchTuning(C.Structure):
    _fields_ = [("raster_blocks_per_cu", C.c_uint32), ("shade_tile_w_log2", C.c_uint32),
                ("groups", C.c_uint32), ("raster_tile", C.c_uint32)]


class RasterConfig(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("tile_w_log2", C.c_uint32), ("tile_h_log2", C.c_uint32),
                ("bin_capacity", C.c_uint32)]


RASTER_SAMPLES, RASTER_TILES = 0, 1


class FrameStats(C.Structure):
    _fields_ = [("n_cut", C.c_uint32), ("n_frag", C.c_uint32), ("overflow", C.c_uint32),
                ("n_profiled", C.c_uint32), ("ms_clear", C.c_float), ("ms_prearrange", C.c_float),
                ("ms_depth", C.c_float), ("ms_id", C.c_float), ("ms_shade", C.c_float),
                ("ms_total", C.c_float)]


class ServerConfig(C.Structure):
    _fields_ = [("max_batch_views", C.c_uint32), ("max_inflight", C.c_uint32),
                ("max_batch_tris", C.c_uint64), ("hbm_fraction", C.c_float), ("read_mask", C.c_uint32)]


class Completion(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("t_submit_ns", C.c_uint64), ("t_start_ns", C.c_uint64),
                ("t_done_ns", C.c_uint64), ("host", C.c_void_p), ("status", C.c_int32),
                ("scene_slot", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("batch_views", C.c_uint32), ("_pad", C.c_uint32)]


class ServerInfo(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("hbm_held", C.c_uint64), ("n_batches", C.c_uint64),
                ("n_done", C.c_uint64), ("n_refused", C.c_uint64), ("n_splits", C.c_uint64),
                ("n_pending", C.c_uint32), ("n_frames", C.c_uint32), ("frame_bytes", C.c_uint64),
                ("n_evicted", C.c_uint64)]


READ_RGBA, READ_DEPTH, READ_IDKEY = 1, 2, 4

_lib = None


def load_library(path=LIB_PATH):
    """Load libocr.so once; raise if it is absent (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library missing: {path}; run `python -m openclrenderer_amd.build`")
    L = C.CDLL(path)
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    pp = C.POINTER(C.c_void_p)
    L.ocr_last_error.restype = C.c_char_p
    L.ocr_create.argtypes = [i32, pp]
    L.ocr_destroy.argtypes = [vp]
    L.ocr_destroy.restype = None
    L.ocr_query_free_hbm.argtypes = [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    L.ocr_scene_create.argtypes = [vp, vp, u32, vp, u32, pp]
    L.ocr_scene_update_objs.argtypes = [vp, u32, u32, vp]
    L.ocr_scene_update_objs_async.argtypes = [vp, u32, u32, vp, vp]
    L.ocr_atlas_upload.argtypes = [vp, vp, u32]
    L.ocr_lights_upload.argtypes = [vp, vp, u32, u32]
    L.ocr_shadow_pass.argtypes = [vp, vp]
    L.ocr_scene_destroy.argtypes = [vp]
    L.ocr_scene_destroy.restype = None
    L.ocr_scene_atlas_info.argtypes = [vp] + [C.POINTER(C.c_uint32)] * 4
    L.ocr_scene_read_atlas.argtypes = [vp, vp, vp, vp]
    L.ocr_scene_read_shadow_maps.argtypes = [vp, vp, vp]
    L.ocr_scene_read_motion.argtypes = [vp, u32, u32, vp]
    L.ocr_frame_create.argtypes = [vp, u32, u32, pp]
    L.ocr_frame_destroy.argtypes = [vp]
    L.ocr_frame_destroy.restype = None
    L.ocr_frame_set_profiling.argtypes = [vp, i32]
    L.ocr_frame_reserve.argtypes = [vp, u32]
    L.ocr_frame_set_raster.argtypes = [vp, C.POINTER(RasterConfig)]
    L.ocr_frame_get_raster.argtypes = [vp, C.POINTER(RasterConfig)]
    L.ocr_render.argtypes = [vp, vp, vp, vp]
    L.ocr_render_noalloc.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.ocr_frame_read.argtypes = [vp, vp, vp, vp, vp, vp]
    L.ocr_frame_device_buffers.argtypes = [vp, pp, pp, pp, pp]
    L.ocr_frame_blend_with_depth.argtypes = [vp, vp, vp]
    L.ocr_frame_pseudo_aa.argtypes = [vp, vp]
    L.ocr_frame_motion_blur.argtypes = [vp, vp, vp, vp, vp, u32, C.c_float, C.c_float, vp]
    L.ocr_frame_get_stats.argtypes = [vp, vp, C.POINTER(FrameStats)]
    L.ocr_frame_count_texels.argtypes = [vp, vp, vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), vp]
    L.ocr_batch_create.argtypes = [vp, u32, pp]
    L.ocr_batch_destroy.argtypes = [vp]
    L.ocr_batch_destroy.restype = None
    L.ocr_batch_set_views.argtypes = [vp, vp, vp, vp, u32]
    L.ocr_batch_set_tuning.argtypes = [vp, C.POINTER(BatchTuning)]
    L.ocr_batch_get_tuning.argtypes = [vp, C.POINTER(BatchTuning)]
    L.ocr_batch_set_profiling.argtypes = [vp, i32]
    L.ocr_batch_render.argtypes = [vp, vp]
    L.ocr_batch_launch_graph.argtypes = [vp, vp]
    L.ocr_batch_get_stats.argtypes = [vp, vp, C.POINTER(FrameStats)]
    L.ocr_batch_autotune.argtypes = [vp, vp, u32, C.POINTER(BatchTuning)]
    L.ocr_server_create.argtypes = [vp, C.POINTER(ServerConfig), pp]
    L.ocr_server_add_scene.argtypes = [vp, vp, C.POINTER(C.c_uint32)]
    L.ocr_server_submit.argtypes = [vp, u32, vp, C.POINTER(C.c_uint64)]
    L.ocr_server_poll.argtypes = [vp, C.POINTER(Completion), u32, C.POINTER(C.c_uint32), i32]
    L.ocr_server_release.argtypes = [vp, C.c_uint64]
    L.ocr_server_get_info.argtypes = [vp, C.POINTER(ServerInfo)]
    L.ocr_server_now_ns.argtypes = []
    L.ocr_server_now_ns.restype = C.c_uint64
    L.ocr_server_destroy.argtypes = [vp]
    L.ocr_server_destroy.restype = None
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = load_library().ocr_last_error().decode(errors="replace")
        raise OcrError(rc, msg)
    return rc


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Context:
    def __init__(self, device=0):
        self.lib = load_library()
        h = C.c_void_p()
        check(self.lib.ocr_create(device, C.byref(h)))
        self.h = h

    def free_hbm(self):
        f, t = C.c_size_t(), C.c_size_t()
        check(self.lib.ocr_query_free_hbm(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def close(self):
        if self.h:
            self.lib.ocr_destroy(self.h)
            self.h = None


class GpuScene:
    """A device-resident scene: geometry, descriptors, atlas, lights."""

    def __init__(self, ctx, scene):
        self.ctx, self.lib = ctx, ctx.lib
        self.scene = scene
        tris = np.ascontiguousarray(scene.tris)
        objs = np.ascontiguousarray(scene.objs)
        h = C.c_void_p()
        check(self.lib.ocr_scene_create(ctx.h, _p(tris), tris.shape[0], _p(objs), objs.shape[0],
                                        C.byref(h)))
        self.h = h
        self.nobj = objs.shape[0]
        self.upload_textures(scene.textures)
        self.upload_lights(scene.lights, scene.light_buffer_dim)

    def upload_textures(self, textures):
        n = len(textures)
        srcs = (TexSrc * max(n, 1))()
        keep = []
        for i, t in enumerate(textures):
            rgba = np.ascontiguousarray(t.rgba, dtype=np.uint8)
            keep.append(rgba)
            srcs[i].width, srcs[i].height, srcs[i].rgba = t.width, t.height, rgba.ctypes.data
        check(self.lib.ocr_atlas_upload(self.h, C.cast(srcs, C.c_void_p), n))

    def upload_lights(self, lights, dim):
        lights = np.ascontiguousarray(lights, dtype=abi.LIGHT)
        check(self.lib.ocr_lights_upload(self.h, _p(lights), lights.shape[0], dim))
        self.light_buffer_dim = dim
        self.n_shadow = int(np.sum(lights["shadow"] == 1))
        self.n_static = int(np.sum((lights["shadow"] != 0) & (lights["is_static"] != 0)))

    def update_objs(self, first, objs):
        objs = np.ascontiguousarray(objs, dtype=abi.OBJ_DESC)
        check(self.lib.ocr_scene_update_objs(self.h, first, objs.shape[0], _p(objs)))

    def update_objs_async(self, first, objs, stream=None):
        """Descriptor update ordered on `stream`; later renders see it."""
        objs = np.ascontiguousarray(objs, dtype=abi.OBJ_DESC)
        check(self.lib.ocr_scene_update_objs_async(self.h, first, objs.shape[0], _p(objs), stream))

    def shadow_pass(self, stream=None):
        check(self.lib.ocr_shadow_pass(self.h, stream))

    def read_atlas(self):
        npages, nnums, nsizes, ms = (C.c_uint32() for _ in range(4))
        check(self.lib.ocr_scene_atlas_info(self.h, C.byref(npages), C.byref(nnums),
                                            C.byref(nsizes), C.byref(ms)))
        pages = np.zeros((npages.value, 2048, 2048, 4), np.uint8)
        nums = np.zeros(nnums.value, np.uint32)
        sizes = np.zeros(nsizes.value, np.uint32)
        check(self.lib.ocr_scene_read_atlas(self.h, _p(pages), _p(nums), _p(sizes)))
        return pages, nums, sizes, ms.value

    def read_shadow_maps(self):
        L = self.light_buffer_dim
        dyn = np.zeros((max(self.n_shadow, 1), 6, L, L), np.uint32)
        stat = np.zeros((max(self.n_static, 1), 6, L, L), np.uint32)
        check(self.lib.ocr_scene_read_shadow_maps(self.h, _p(dyn), _p(stat)))
        return dyn, stat

    def read_motion(self):
        """(nobj, 4, 4) float32: old_world_pos_1, old_world_pos_2,
        old_world_rot_quat_1, old_world_rot_quat_2 of every object."""
        out = np.zeros((self.nobj, 4, 4), np.float32)
        check(self.lib.ocr_scene_read_motion(self.h, 0, self.nobj, _p(out)))
        return out

    def close(self):
        if self.h:
            self.lib.ocr_scene_destroy(self.h)
            self.h = None


class Frame:
    def __init__(self, ctx, width, height, profiling=False):
        self.ctx, self.lib = ctx, ctx.lib
        self.width, self.height = width, height
        h = C.c_void_p()
        check(self.lib.ocr_frame_create(ctx.h, width, height, C.byref(h)))
        self.h = h
        if profiling:
            check(self.lib.ocr_frame_set_profiling(h, 1))

    def render(self, gscene, view, stream=None):
        v = np.ascontiguousarray(view, dtype=abi.VIEW)
        check(self.lib.ocr_render(gscene.h, _p(v), self.h, stream))

    def read(self, stream=None, depth=True, tri_id=True, rgba=True, normals=True):
        H, W = self.height, self.width
        d = np.zeros((H, W), np.uint32) if depth else None
        t = np.zeros((H, W), np.uint32) if tri_id else None
        c = np.zeros((H, W, 4), np.uint8) if rgba else None
        n = np.zeros((H, W, 2), np.uint16) if normals else None
        check(self.lib.ocr_frame_read(self.h, _p(d), _p(t), _p(c), _p(n), stream))
        return {"depth": d, "tri_id": t, "rgba": c, "normals": n}

    def render_read(self, gscene, view, stream=None, **kw):
        """Render and read back, re-rendering once if a work list overflowed
        (the first read grows the lists)."""
        for attempt in range(3):
            self.render(gscene, view, stream)
            try:
                return self.read(stream, **kw)
            except OcrError as e:
                if e.code != abi.OCR_ERR_OVERFLOW or attempt == 2:
                    raise
        raise AssertionError("unreachable")

    def stats(self, stream=None):
        s = FrameStats()
        check(self.lib.ocr_frame_get_stats(self.h, stream, C.byref(s)))
        return {k: getattr(s, k) for k, _ in FrameStats._fields_}

    def reserve(self, max_cut_triangles):
        """Fixes the cut-list capacity; a render that needs more reports
        OCR_ERR_OVERFLOW and grows it."""
        check(self.lib.ocr_frame_reserve(self.h, max_cut_triangles))

    def set_profiling(self, on=True, shade_only=False):
        """Per-pass hipEvent timings of every render (shade_only: of the
        shading pass alone, two events per render)."""
        check(self.lib.ocr_frame_set_profiling(self.h, (2 if shade_only else 1) if on else 0))

    def set_raster(self, mode, tile_w_log2=4, tile_h_log2=4, bin_capacity=0):
        """RASTER_SAMPLES (the sample-space depth and id passes) or
        RASTER_TILES (tiles of 2^tile_w_log2 x 2^tile_h_log2 pixels resolved
        in LDS); both give the same buffers."""
        c = RasterConfig(mode, tile_w_log2, tile_h_log2, bin_capacity)
        check(self.lib.ocr_frame_set_raster(self.h, C.byref(c)))

    def raster(self):
        c = RasterConfig()
        check(self.lib.ocr_frame_get_raster(self.h, C.byref(c)))
        return c.mode, c.tile_w_log2, c.tile_h_log2, c.bin_capacity

    def count_texels(self, gscene, view, stream=None):
        v = np.ascontiguousarray(view, dtype=abi.VIEW)
        a, s = C.c_uint64(), C.c_uint64()
        check(self.lib.ocr_frame_count_texels(gscene.h, _p(v), self.h, C.byref(a), C.byref(s), stream))
        return a.value, s.value

    def blend_from(self, src, stream=None):
        """engine::blend_with_depth(src, self): src over this frame where src
        is at least as near."""
        check(self.lib.ocr_frame_blend_with_depth(src.h, self.h, stream))

    def pseudo_aa(self, stream=None):
        """engine::do_pseudo_aa on this frame's last render."""
        check(self.lib.ocr_frame_pseudo_aa(self.h, stream))

    def motion_blur(self, gscene, view, c_pos_old, c_rot_old, frame_id, strength=1.0,
                    camera_contribution=1.0, stream=None):
        """engine::do_motion_blur on this frame's last render of gscene."""
        v = np.ascontiguousarray(view, dtype=abi.VIEW)
        po = np.ascontiguousarray(np.resize(np.asarray(c_pos_old, np.float32), 4))
        ro = np.ascontiguousarray(np.resize(np.asarray(c_rot_old, np.float32), 4))
        check(self.lib.ocr_frame_motion_blur(self.h, gscene.h, _p(v), _p(po), _p(ro), frame_id,
                                             strength, camera_contribution, stream))

    def device_buffers(self):
        ptrs = [C.c_void_p() for _ in range(4)]
        check(self.lib.ocr_frame_device_buffers(self.h, *[C.byref(p) for p in ptrs]))
        return [p.value for p in ptrs]

    def close(self):
        if self.h:
            self.lib.ocr_frame_destroy(self.h)
            self.h = None


class Batch:
    """Many views of several resident scenes rendered together (one launch
    per pass for the whole batch, optionally replayed from a hipGraph)."""

    def __init__(self, ctx, max_views):
        self.ctx, self.lib = ctx, ctx.lib
        h = C.c_void_p()
        check(self.lib.ocr_batch_create(ctx.h, max_views, C.byref(h)))
        self.h = h
        self.frames = []

    def set_views(self, gscenes, views, frames):
        n = len(views)
        sc = (C.c_void_p * n)(*[g.h.value for g in gscenes])
        fr = (C.c_void_p * n)(*[f.h.value for f in frames])
        vv = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=abi.VIEW).reshape(1) for v in views]))
        self._keep = (sc, fr, vv)
        self.frames = frames
        check(self.lib.ocr_batch_set_views(self.h, C.cast(sc, C.c_void_p), _p(vv), C.cast(fr, C.c_void_p), n))

    def set_tuning(self, raster_blocks_per_cu, shade_tile_w_log2, groups=1, raster_tile=0):
        """raster_tile: 0 for the sample-space raster passes, else
        tile_w_log2 | tile_h_log2 << 8 for the tiled raster path."""
        t = BatchTuning(raster_blocks_per_cu, shade_tile_w_log2, groups, raster_tile)
        check(self.lib.ocr_batch_set_tuning(self.h, C.byref(t)))

    def tuning(self):
        """(raster blocks per CU, shading tile width log2, view groups, raster tile)."""
        t = BatchTuning()
        check(self.lib.ocr_batch_get_tuning(self.h, C.byref(t)))
        return t.raster_blocks_per_cu, t.shade_tile_w_log2, max(t.groups, 1), t.raster_tile

    def set_profiling(self, on=True):
        check(self.lib.ocr_batch_set_profiling(self.h, 1 if on else 0))

    def render(self, stream=None):
        check(self.lib.ocr_batch_render(self.h, stream))

    def launch_graph(self, stream=None):
        check(self.lib.ocr_batch_launch_graph(self.h, stream))

    def stats(self, stream=None):
        s = FrameStats()
        check(self.lib.ocr_batch_get_stats(self.h, stream, C.byref(s)))
        return {k: getattr(s, k) for k, _ in FrameStats._fields_}

    def autotune(self, reps=5, stream=None):
        t = BatchTuning()
        check(self.lib.ocr_batch_autotune(self.h, stream, reps, C.byref(t)))
        return t.raster_blocks_per_cu, t.shade_tile_w_log2, max(t.groups, 1), t.raster_tile

    def render_checked(self, stream=None, graph=False):
        """Render, and once more if the shared lists overflowed (they grow)."""
        for attempt in range(3):
            (self.launch_graph if graph else self.render)(stream)
            try:
                return self.stats(stream)
            except OcrError as e:
                if e.code != abi.OCR_ERR_OVERFLOW or attempt == 2:
                    raise
        raise AssertionError("unreachable")

    def close(self):
        if self.h:
            self.lib.ocr_batch_destroy(self.h)
            self.h = None


def id_keys_to_tri(keys, depth):
    """The triangle-id image from raw id keys and depth, as ocr_frame_read
    converts it (background and empty keys give OCR_NO_TRIANGLE)."""
    keys = np.asarray(keys, dtype=np.uint64)
    tri = (((keys >> np.uint64(32)) - np.uint64(1)) >> np.uint64(1)).astype(np.uint32)
    empty = (keys == 0) | (np.asarray(depth) == 0xFFFFFFFF)
    tri[empty] = 0xFFFFFFFF
    return tri


class Server:
    """The served form: resident scenes, admission against an HBM budget,
    continuous batching over `max_inflight` in-flight batch slots."""

    def __init__(self, ctx, max_batch_views=32, max_inflight=2, max_batch_tris=16_000_000,
                 hbm_fraction=0.5, read_mask=READ_RGBA):
        self.ctx, self.lib = ctx, ctx.lib
        self.read_mask = read_mask
        cfg = ServerConfig(max_batch_views, max_inflight, max_batch_tris, hbm_fraction, read_mask)
        h = C.c_void_p()
        check(self.lib.ocr_server_create(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h
        self.scenes = []

    def add_scene(self, gscene):
        slot = C.c_uint32()
        check(self.lib.ocr_server_add_scene(self.h, gscene.h, C.byref(slot)))
        self.scenes.append(gscene)
        return slot.value

    def submit(self, slot, view):
        """Returns the request id, or None when admission refuses it (busy)."""
        v = np.ascontiguousarray(np.asarray(view, dtype=abi.VIEW).reshape(1))
        rid = C.c_uint64()
        rc = self.lib.ocr_server_submit(self.h, slot, _p(v), C.byref(rid))
        if rc == abi.OCR_ERR_BUSY:
            return None
        check(rc)
        return rid.value

    def poll(self, max_n=256, wait_ms=0):
        buf = (Completion * max_n)()
        n = C.c_uint32()
        check(self.lib.ocr_server_poll(self.h, buf, max_n, C.byref(n), wait_ms))
        return [buf[i] for i in range(n.value)]

    def outputs(self, c):
        """Copies of a completion's host outputs, keyed rgba / depth / id_key."""
        npix = c.width * c.height
        out, off = {}, 0
        base = c.host
        for bit, name, nbytes, dt in ((READ_RGBA, "rgba", 4, np.uint8), (READ_DEPTH, "depth", 4, np.uint32),
                                      (READ_IDKEY, "id_key", 8, np.uint64)):
            if self.read_mask & bit:
                raw = (C.c_uint8 * (nbytes * npix)).from_address(base + off)
                a = np.frombuffer(raw, dtype=dt).copy()
                out[name] = a.reshape(c.height, c.width, 4) if name == "rgba" else a.reshape(c.height, c.width)
                off += nbytes * npix
        return out

    def release(self, request_id):
        check(self.lib.ocr_server_release(self.h, request_id))

    def info(self):
        i = ServerInfo()
        check(self.lib.ocr_server_get_info(self.h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in ServerInfo._fields_}

    def now_ns(self):
        return int(self.lib.ocr_server_now_ns())

    def close(self):
        if self.h:
            self.lib.ocr_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
