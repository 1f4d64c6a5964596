"""ctypes binding of the oracle (oracle/_build/libocr_oracle.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the result the HIP path is compared with.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libocr_oracle.so")


class OracleScene(C.Structure):
    _fields_ = [
        ("tris", C.c_void_p), ("ntri", C.c_uint32),
        ("objs", C.c_void_p), ("nobj", C.c_uint32),
        ("lights", C.c_void_p), ("nlights", C.c_uint32),
        ("atlas", C.c_void_p), ("npages", C.c_uint32),
        ("nums", C.c_void_p), ("nnums", C.c_uint32),
        ("sizes", C.c_void_p), ("nsizes", C.c_uint32),
        ("mip_start", C.c_uint32),
        ("shadow_maps", C.c_void_p),
        ("static_maps", C.c_void_p),
        ("lbuf_dim", C.c_uint32),
    ]


class OracleOut(C.Structure):
    _fields_ = [
        ("depth", C.c_void_p), ("tri_id", C.c_void_p), ("rgba8", C.c_void_p),
        ("normals", C.c_void_p),
        ("n_cut", C.c_uint32), ("n_frag", C.c_uint32), ("n_samples", C.c_uint64),
        ("n_visible_tris", C.c_uint32), ("n_visible_objs", C.c_uint32),
        ("ms_prearrange", C.c_double), ("ms_depth", C.c_double), ("ms_id", C.c_double),
        ("ms_shade", C.c_double),
    ]


class SinCos(C.Structure):
    _fields_ = [("s", C.c_float * 3), ("c", C.c_float * 3)]


class TexSrc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.c_void_p)]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def camera_sincos(c_rot):
    """sin/cos of the camera Euler angles, binary64 rounded once to binary32."""
    sc = SinCos()
    for i in range(3):
        a = float(np.float32(c_rot[i]))
        sc.s[i] = float(np.float32(math.sin(a)))
        sc.c[i] = float(np.float32(math.cos(a)))
    return sc


class Oracle:
    def __init__(self, path=None):
        # OCR_ORACLE_LIB: another build of the same source (the sanitizer
        # build of oracle/Makefile's `asan` target)
        path = path or os.environ.get("OCR_ORACLE_LIB") or LIB_PATH
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run oracle/Makefile or build())")
        self.lib = L = C.CDLL(path)
        L.oracle_render.argtypes = [C.POINTER(OracleScene), C.c_void_p, C.POINTER(SinCos), C.c_int,
                                    C.POINTER(OracleOut)]
        L.oracle_shadow_maps.argtypes = [C.POINTER(OracleScene), C.c_int, C.c_void_p, C.c_void_p]
        L.oracle_shadow_maps_ex.argtypes = [C.POINTER(OracleScene), C.c_int, C.c_void_p, C.c_void_p,
                                            C.POINTER(C.c_uint64)]
        L.oracle_atlas_plan.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32),
                                        C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)]
        L.oracle_atlas_fill.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                        C.c_void_p]
        L.oracle_face_sincos.argtypes = [C.c_void_p]
        L.oracle_cut_list.argtypes = [C.POINTER(OracleScene), C.c_void_p, C.POINTER(SinCos), C.c_void_p,
                                      C.c_void_p, C.c_int]
        L.oracle_cut_list.restype = C.c_int
        f, u, i = C.c_float, C.c_uint32, C.c_int
        fp = C.POINTER(C.c_float)
        L.oracle_rot.argtypes = [fp, fp, C.POINTER(SinCos), fp]
        L.oracle_back_rot.argtypes = [fp, C.POINTER(SinCos), fp]
        L.oracle_rot_quat.argtypes = [fp, fp, fp]
        L.oracle_back_rot_quat.argtypes = [fp, fp, fp]
        L.oracle_clip.argtypes = [fp, fp]
        L.oracle_clip.restype = i
        L.oracle_calc_min_max.argtypes = [fp, f, f, fp]
        L.oracle_point_in_tri.argtypes = [f, f, fp, fp]
        L.oracle_point_in_tri.restype = i
        L.oracle_walk.argtypes = [fp, f, f, u, i, C.c_void_p, C.c_void_p, C.c_void_p, i]
        L.oracle_walk.restype = i
        for name in ("oracle_wang_hash", "oracle_xorshift"):
            getattr(L, name).argtypes = [u]
            getattr(L, name).restype = u
        for name in ("oracle_rational_acos", "oracle_log2_approx", "oracle_exp", "oracle_tan",
                     "oracle_powr5", "oracle_pow4"):
            getattr(L, name).argtypes = [f]
            getattr(L, name).restype = f
        L.oracle_gamma_transform.argtypes = [fp, fp]
        L.oracle_gamma_inverse.argtypes = [fp, fp]
        L.oracle_ret_cubeface.argtypes = [fp, fp]
        L.oracle_ret_cubeface.restype = i
        L.oracle_get_barycentric.argtypes = [fp, fp, fp, fp, fp, fp, fp]
        L.oracle_encode_normal.argtypes = [fp]
        L.oracle_encode_normal.restype = u
        L.oracle_texture_mod.argtypes = [fp, fp]
        L.oracle_blend_with_depth.argtypes = [u, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_pseudo_aa.argtypes = [u, u, C.c_void_p, C.c_void_p, C.c_void_p]
        vp = C.c_void_p
        L.oracle_motion_blur.argtypes = [vp, vp, fp, vp, vp, vp, vp, u, vp, u, u, C.c_float, C.c_float, vp]
        L.oracle_sample_linear.argtypes = [vp, u, u, C.c_float, C.c_float, fp]
        L.oracle_ssao.argtypes = [i, i, vp, u, u, f]
        L.oracle_ssao.restype = f
        L.oracle_shadow_term.argtypes = [fp, fp, fp, vp, u, u]
        L.oracle_shadow_term.restype = f
        L.oracle_texture_filter.argtypes = [f, f, f, f, i, u, vp, vp, vp, fp]
        L.oracle_shadow_map_one.argtypes = [C.POINTER(OracleScene), u, i, vp, vp]
        L.oracle_shadow_map_one.restype = i

    # --------------------------------------------------------------- helpers
    @staticmethod
    def _f(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        return a, a.ctypes.data_as(C.POINTER(C.c_float))

    def call_vec(self, name, *arrays, out_len=3):
        keep, ptrs = [], []
        for a in arrays:
            arr, p = self._f(a)
            keep.append(arr)
            ptrs.append(p)
        out, po = self._f(np.zeros(out_len, dtype=np.float32))
        getattr(self.lib, name)(*ptrs, po)
        return out

    def blend_with_depth(self, src_depth, dst_depth, src_rgba, dst_rgba):
        """blend_screens_with_depth on host arrays; returns the new dst RGBA8
        (H x W x 4 uint8), the inputs are left unchanged."""
        sd = np.ascontiguousarray(src_depth, dtype=np.uint32)
        dd = np.ascontiguousarray(dst_depth, dtype=np.uint32)
        sr = np.ascontiguousarray(src_rgba, dtype=np.uint8)
        out = np.array(dst_rgba, dtype=np.uint8, copy=True, order="C")
        assert sd.size == dd.size == sr.size // 4 == out.size // 4
        self.lib.oracle_blend_with_depth(sd.size, sd.ctypes.data, dd.ctypes.data, sr.ctypes.data, out.ctypes.data)
        return out

    def pseudo_aa(self, depth, normals, rgba):
        """do_pseudo_aa on a frame's host buffers (depth H x W u32, normals
        H x W x 2 u16, rgba H x W x 4 u8); returns the new RGBA8."""
        H, W = depth.shape
        d = np.ascontiguousarray(depth, dtype=np.uint32)
        n = np.ascontiguousarray(normals, dtype=np.uint16)
        out = np.array(rgba, dtype=np.uint8, copy=True, order="C")
        assert n.size == 2 * W * H and out.size == 4 * W * H
        self.lib.oracle_pseudo_aa(W, H, d.ctypes.data, n.ctypes.data, out.ctypes.data)
        return out

    def motion_blur(self, view, frame, tris, objs, c_pos_old, c_rot_old, frame_id, strength=1.0,
                    camera_contribution=1.0):
        """do_motion_blur on a frame's host buffers (dict with depth, tri_id,
        rgba as Frame.read returns).  Returns (new RGBA8, descriptors with the
        recorded poses); the inputs are left unchanged."""
        v = np.ascontiguousarray(view)
        W, H = int(v["width"][0]), int(v["height"][0])
        d = np.ascontiguousarray(frame["depth"], dtype=np.uint32)
        t = np.ascontiguousarray(frame["tri_id"], dtype=np.uint32)
        out = np.array(frame["rgba"], dtype=np.uint8, copy=True, order="C")
        assert d.size == t.size == W * H and out.size == 4 * W * H
        tr = np.ascontiguousarray(tris)
        ob = np.array(objs, copy=True, order="C")
        cam = camera_sincos(v["c_rot"][0])
        cam_old = camera_sincos(np.asarray(c_rot_old, np.float32))
        po, ppo = self._f(np.asarray(c_pos_old, np.float32)[:3])
        self.lib.oracle_motion_blur(v.ctypes.data, C.addressof(cam), ppo, C.addressof(cam_old), d.ctypes.data,
                                    t.ctypes.data, tr.ctypes.data, tr.shape[0], ob.ctypes.data, ob.shape[0],
                                    frame_id, strength, camera_contribution, out.ctypes.data)
        return out, ob

    def sample_linear(self, rgba, u, v):
        img = np.ascontiguousarray(rgba, dtype=np.uint8)
        H, W = img.shape[:2]
        out, po = self._f(np.zeros(4, np.float32))
        self.lib.oracle_sample_linear(img.ctypes.data, W, H, u, v, po)
        return out

    def ssao(self, depth, x, y, fov):
        """generate_ssao at pixel (x, y) of an H x W u32 depth buffer."""
        d = np.ascontiguousarray(depth, dtype=np.uint32)
        H, W = d.shape
        return self.lib.oracle_ssao(int(x), int(y), d.ctypes.data, W, H, float(fov))

    def shadow_term(self, gpos, normal, lpos, maps, shnum=0):
        """kernel3's shadow term of one light (0 lit .. 1 shadowed) from cube
        maps shaped (n, 6, L, L)."""
        m = np.ascontiguousarray(maps, dtype=np.uint32)
        g, pg = self._f(gpos)
        n, pn = self._f(normal)
        lp, pl = self._f(lpos)
        return self.lib.oracle_shadow_term(pg, pn, pl, m.ctypes.data, shnum, m.shape[-1])

    def texture_filter(self, atlas, vt, dx, dy, tid=0):
        """texture_filter_diff on an atlas (pages, nums, sizes, mip_start)."""
        pages, nums, sizes, mip_start = atlas
        out, po = self._f(np.zeros(4, np.float32))
        self.lib.oracle_texture_filter(float(vt[0]), float(vt[1]), float(dx), float(dy), tid, mip_start,
                                       _ptr(nums), _ptr(sizes), _ptr(pages), po)
        return out

    def shadow_map_one(self, scene, light, only_static=False):
        """One light's (6, L, L) cube map, rendered sequentially, and the
        fragment records per face."""
        L = scene.light_buffer_dim
        m = np.zeros((6, L, L), np.uint32)
        frags = np.zeros(6, np.uint64)
        dummy_atlas = (np.zeros((1, 1, 1, 4), np.uint8), np.zeros(1, np.uint32),
                       np.zeros(1, np.uint32), 0)
        s = self._scene(scene, dummy_atlas, None)
        rc = self.lib.oracle_shadow_map_one(C.byref(s), light, 1 if only_static else 0, _ptr(m), _ptr(frags))
        if rc != 0:
            raise ValueError("oracle_shadow_map_one: no such light")
        return m, frags

    def face_sincos(self):
        arr = (SinCos * 6)()
        self.lib.oracle_face_sincos(C.cast(arr, C.c_void_p))
        return arr

    # ----------------------------------------------------------------- atlas
    def atlas(self, textures):
        n = len(textures)
        if n == 0:
            return (np.zeros((1, 2048, 2048, 4), np.uint8), np.zeros(2, np.uint32),
                    np.full(2, 16, np.uint32), 0)
        w = np.array([t.width for t in textures], dtype=np.uint32)
        h = np.array([t.height for t in textures], dtype=np.uint32)
        nums = np.zeros(n * 5, dtype=np.uint32)
        sizes = np.zeros(n * 5, dtype=np.uint32)
        npages = C.c_uint32(0)
        mip_start = C.c_uint32(0)
        rc = self.lib.oracle_atlas_plan(_ptr(w), _ptr(h), n, C.byref(npages), _ptr(nums),
                                        _ptr(sizes), C.byref(mip_start))
        if rc != 0:
            raise ValueError(f"oracle_atlas_plan failed ({rc})")
        pages = np.zeros((npages.value, 2048, 2048, 4), dtype=np.uint8)
        srcs = (TexSrc * n)()
        keep = []
        for i, t in enumerate(textures):
            rgba = np.ascontiguousarray(t.rgba, dtype=np.uint8)
            keep.append(rgba)
            srcs[i].width, srcs[i].height, srcs[i].rgba = t.width, t.height, rgba.ctypes.data
        self.lib.oracle_atlas_fill(C.cast(srcs, C.c_void_p), n, _ptr(nums), _ptr(sizes),
                                   mip_start.value, _ptr(pages))
        return pages, nums, sizes[:npages.value].copy(), mip_start.value

    # ----------------------------------------------------------------- scene
    def _scene(self, scene, atlas, maps):
        pages, nums, sizes, mip_start = atlas
        dyn, stat = maps if maps is not None else (None, None)
        s = OracleScene()
        s.tris, s.ntri = _ptr(scene.tris), scene.tris.shape[0]
        s.objs, s.nobj = _ptr(scene.objs), scene.objs.shape[0]
        s.lights, s.nlights = _ptr(scene.lights), scene.lights.shape[0]
        s.atlas, s.npages = _ptr(pages), pages.shape[0]
        s.nums, s.nnums = _ptr(nums), nums.shape[0]
        s.sizes, s.nsizes = _ptr(sizes), sizes.shape[0]
        s.mip_start = mip_start
        s.shadow_maps = _ptr(dyn)
        s.static_maps = _ptr(stat)
        s.lbuf_dim = scene.light_buffer_dim
        return s

    def cut_list(self, scene, view, atlas=None):
        if atlas is None:
            atlas = (np.zeros((1, 1, 1, 4), np.uint8), np.zeros(1, np.uint32), np.zeros(1, np.uint32), 0)
        s = self._scene(scene, atlas, None)
        v = np.ascontiguousarray(view)
        cam = camera_sincos(v["c_rot"][0])
        cap = 2 * scene.tris.shape[0] + 16
        cuts = np.zeros((cap, 3, 3), np.float32)
        keys = np.zeros(cap, np.uint32)
        n = self.lib.oracle_cut_list(C.byref(s), _ptr(v), C.byref(cam), _ptr(cuts), _ptr(keys), cap)
        return cuts[:n], keys[:n]

    def shadow_maps(self, scene, nthreads=1, with_samples=False):
        """Dynamic and static cube maps; with_samples also returns the
        covered samples (depth tests) of all the maps."""
        L = scene.light_buffer_dim
        nsh = int(np.sum(scene.lights["shadow"] == 1))
        nst = int(np.sum((scene.lights["shadow"] != 0) & (scene.lights["is_static"] != 0)))
        dyn = np.zeros((max(nsh, 1), 6, L, L), dtype=np.uint32)
        stat = np.zeros((max(nst, 1), 6, L, L), dtype=np.uint32)
        dummy_atlas = (np.zeros((1, 1, 1, 4), np.uint8), np.zeros(1, np.uint32),
                       np.zeros(1, np.uint32), 0)
        s = self._scene(scene, dummy_atlas, None)
        samples = C.c_uint64(0)
        self.lib.oracle_shadow_maps_ex(C.byref(s), nthreads, _ptr(dyn), _ptr(stat), C.byref(samples))
        return (dyn, stat, samples.value) if with_samples else (dyn, stat)

    def render(self, scene, view, atlas=None, maps=None, nthreads=1, shade=True):
        if atlas is None:
            atlas = self.atlas(scene.textures)
        if maps is None and np.any(scene.lights["shadow"] != 0):
            maps = self.shadow_maps(scene, nthreads)
        v = np.ascontiguousarray(view)
        W, H = int(v["width"][0]), int(v["height"][0])
        depth = np.zeros((H, W), np.uint32)
        tri = np.zeros((H, W), np.uint32)
        rgba = np.zeros((H, W, 4), np.uint8) if shade else None
        nrm = np.zeros((H, W, 2), np.uint16) if shade else None
        out = OracleOut()
        out.depth, out.tri_id, out.rgba8, out.normals = _ptr(depth), _ptr(tri), _ptr(rgba), _ptr(nrm)
        s = self._scene(scene, atlas, maps)
        cam = camera_sincos(v["c_rot"][0])
        rc = self.lib.oracle_render(C.byref(s), _ptr(v), C.byref(cam), nthreads, C.byref(out))
        if rc != 0:
            raise RuntimeError(f"oracle_render failed ({rc})")
        stats = {k: getattr(out, k) for k, _ in OracleOut._fields_[4:]}
        return {"depth": depth, "tri_id": tri, "rgba": rgba, "normals": nrm, "stats": stats}
