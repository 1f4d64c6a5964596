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
    "ocr_scene_create", "ocr_scene_update_objs", "ocr_atlas_upload", "ocr_lights_upload",
    "ocr_shadow_pass", "ocr_scene_destroy", "ocr_scene_atlas_info", "ocr_scene_read_atlas",
    "ocr_scene_read_shadow_maps", "ocr_frame_create", "ocr_frame_destroy",
    "ocr_frame_set_profiling", "ocr_render", "ocr_frame_read", "ocr_frame_device_buffers",
    "ocr_frame_get_stats", "ocr_frame_count_texels", "ocr_frame_reserve",
    "ocr_batch_create", "ocr_batch_destroy", "ocr_batch_set_views", "ocr_batch_set_tuning",
    "ocr_batch_get_tuning", "ocr_batch_set_profiling", "ocr_batch_render", "ocr_batch_launch_graph",
    "ocr_batch_get_stats", "ocr_batch_autotune",
]


class OcrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ocr status {code}: {msg}")
        self.code = code


class TexSrc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.c_void_p)]


class BatchTuning(C.Structure):
    _fields_ = [("raster_blocks_per_cu", C.c_uint32), ("shade_tile_w_log2", C.c_uint32),
                ("_pad", C.c_uint32 * 2)]


class FrameStats(C.Structure):
    _fields_ = [("n_cut", C.c_uint32), ("n_frag", C.c_uint32), ("overflow", C.c_uint32),
                ("n_profiled", C.c_uint32), ("ms_clear", C.c_float), ("ms_prearrange", C.c_float),
                ("ms_depth", C.c_float), ("ms_id", C.c_float), ("ms_shade", C.c_float),
                ("ms_total", C.c_float)]


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
    L.ocr_atlas_upload.argtypes = [vp, vp, u32]
    L.ocr_lights_upload.argtypes = [vp, vp, u32, u32]
    L.ocr_shadow_pass.argtypes = [vp, vp]
    L.ocr_scene_destroy.argtypes = [vp]
    L.ocr_scene_destroy.restype = None
    L.ocr_scene_atlas_info.argtypes = [vp] + [C.POINTER(C.c_uint32)] * 4
    L.ocr_scene_read_atlas.argtypes = [vp, vp, vp, vp]
    L.ocr_scene_read_shadow_maps.argtypes = [vp, vp, vp]
    L.ocr_frame_create.argtypes = [vp, u32, u32, pp]
    L.ocr_frame_destroy.argtypes = [vp]
    L.ocr_frame_destroy.restype = None
    L.ocr_frame_set_profiling.argtypes = [vp, i32]
    L.ocr_frame_reserve.argtypes = [vp, u32]
    L.ocr_render.argtypes = [vp, vp, vp, vp]
    L.ocr_frame_read.argtypes = [vp, vp, vp, vp, vp, vp]
    L.ocr_frame_device_buffers.argtypes = [vp, pp, pp, pp, pp]
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

    def set_profiling(self, on=True):
        check(self.lib.ocr_frame_set_profiling(self.h, 1 if on else 0))

    def count_texels(self, gscene, view, stream=None):
        v = np.ascontiguousarray(view, dtype=abi.VIEW)
        a, s = C.c_uint64(), C.c_uint64()
        check(self.lib.ocr_frame_count_texels(gscene.h, _p(v), self.h, C.byref(a), C.byref(s), stream))
        return a.value, s.value

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

    def set_tuning(self, raster_blocks_per_cu, shade_tile_w_log2):
        t = BatchTuning(raster_blocks_per_cu, shade_tile_w_log2)
        check(self.lib.ocr_batch_set_tuning(self.h, C.byref(t)))

    def tuning(self):
        t = BatchTuning()
        check(self.lib.ocr_batch_get_tuning(self.h, C.byref(t)))
        return t.raster_blocks_per_cu, t.shade_tile_w_log2

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
        return t.raster_blocks_per_cu, t.shade_tile_w_log2

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
