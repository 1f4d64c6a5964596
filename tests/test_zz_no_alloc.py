"""ocr_render_noalloc: the render into caller buffers with the work lists in
the library's static device scratch writes the same depth, id keys, RGBA8
and normals as ocr_render, on two calls in a row (the second with a smaller
view and scene), and allocates no device memory."""
import ctypes as C

import numpy as np
import pytest

from openclrenderer_amd import abi, runtime, scenes

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Device memory the test allocates for the caller-side outputs."""

    def __init__(self, nbytes):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def read(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes == self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, self.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _free_hbm(ctx):
    runtime.load_library()
    C.CDLL("libamdhip64.so").hipDeviceSynchronize()
    return ctx.free_hbm()[0]


def _noalloc(ctx, g, view):
    lib = runtime.load_library()
    W, H = int(view["width"][0]), int(view["height"][0])
    bufs = [DeviceBuffer(4 * W * H), DeviceBuffer(8 * W * H), DeviceBuffer(4 * W * H), DeviceBuffer(4 * W * H)]
    v = np.ascontiguousarray(view, dtype=abi.VIEW)
    before = _free_hbm(ctx)
    runtime.check(lib.ocr_render_noalloc(g.h, v.ctypes.data, *[b.p for b in bufs], None))
    after = _free_hbm(ctx)
    assert after == before, f"device memory changed by {before - after} bytes"
    depth = bufs[0].read(np.uint32, (H, W))
    key = bufs[1].read(np.uint64, (H, W))
    rgba = bufs[2].read(np.uint8, (H, W, 4))
    nrm = bufs[3].read(np.uint16, (H, W, 2))
    for b in bufs:
        b.free()
    return {"depth": depth, "tri_id": runtime.id_keys_to_tri(key, depth), "rgba": rgba, "normals": nrm}


def test_noalloc_equals_render_twice(gpu_ctx):
    sc1, v1 = scenes.config1(shadow=1)
    sc2, v2 = scenes.random_scene(77, n_tris=150, width=96, height=64)
    g1, g2 = runtime.GpuScene(gpu_ctx, sc1), runtime.GpuScene(gpu_ctx, sc2)
    g1.shadow_pass()
    for g, v in ((g1, v1), (g2, v2)):       # the second call: smaller view and scene
        got = _noalloc(gpu_ctx, g, v)
        want = runtime.Frame(gpu_ctx, int(v["width"][0]), int(v["height"][0])).render_read(g, v)
        for k in ("depth", "tri_id", "rgba", "normals"):
            assert np.array_equal(got[k], want[k]), k
        assert (got["depth"] != 0xFFFFFFFF).any()


def test_noalloc_rejects_what_does_not_fit(gpu_ctx):
    lib = runtime.load_library()
    sc, v = scenes.random_scene(78, n_tris=40)
    g = runtime.GpuScene(gpu_ctx, sc)
    big = np.zeros(262_112, dtype=abi.TRIANGLE)
    big["v"]["pos"][:, :, 2] = 100.0
    gb = runtime.GpuScene(gpu_ctx, scenes.Scene(big, sc.objs, sc.lights, sc.textures, 64))
    W, H = int(v["width"][0]), int(v["height"][0])
    bufs = [DeviceBuffer(4 * W * H), DeviceBuffer(8 * W * H), DeviceBuffer(4 * W * H), DeviceBuffer(4 * W * H)]
    vv = np.ascontiguousarray(v, dtype=abi.VIEW)
    assert lib.ocr_render_noalloc(gb.h, vv.ctypes.data, *[b.p for b in bufs], None) == abi.OCR_ERR_UNSUPPORTED
    assert lib.ocr_render_noalloc(g.h, vv.ctypes.data, None, *[b.p for b in bufs[1:]], None) == abi.OCR_ERR_INVALID
    assert lib.ocr_render_noalloc(g.h, vv.ctypes.data, *[b.p for b in bufs], None) == 0
    for b in bufs:
        b.free()
