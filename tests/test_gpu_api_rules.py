"""Error paths and ownership rules of the C ABI around batches and serving:
a light upload that fails leaves a scene that batches still render safely,
a frame serves either ocr_render or batches, and a server's frame pool stays
within its HBM budget when request sizes alternate."""
import time

import numpy as np
import pytest

from openclrenderer_amd import abi, runtime, scenes

pytestmark = pytest.mark.gpu


def _batch_of_two(gpu_ctx, seed):
    sa, va = scenes.random_scene(seed, n_tris=120)
    sb, vb = scenes.random_scene(seed + 1, n_tris=90)
    ga, gb = runtime.GpuScene(gpu_ctx, sa), runtime.GpuScene(gpu_ctx, sb)
    fa = runtime.Frame(gpu_ctx, int(va["width"][0]), int(va["height"][0]))
    fb = runtime.Frame(gpu_ctx, int(vb["width"][0]), int(vb["height"][0]))
    b = runtime.Batch(gpu_ctx, 4)
    b.set_views([ga, gb], [va, vb], [fa, fb])
    return (sa, va, ga, fa), (sb, vb, gb, fb), b


def test_rejected_light_upload_leaves_scene_unchanged(gpu_ctx):
    (sa, va, ga, fa), _, b = _batch_of_two(gpu_ctx, 410)
    b.render_checked(graph=True)
    before = fa.read()
    bad = sa.lights.copy()
    bad["shadow"][0] = 2
    with pytest.raises(runtime.OcrError) as ei:
        ga.upload_lights(bad, 256)
    assert ei.value.code == abi.OCR_ERR_UNSUPPORTED
    for graph in (True, False):
        b.render_checked(graph=graph)
        got = fa.read()
        for k in before:
            assert np.array_equal(got[k], before[k]), (graph, k)
    b.close()


def test_failed_light_upload_then_batch_renders_unlit(gpu_ctx):
    """An upload that fails after the old buffers were freed (the shadow
    maps of 1000 shadow-casting 4000^2 lights, 384 GB, cannot be allocated)
    leaves the scene without lights: batches set before it render the scene
    unlit, equal to the same scene with no lights rendered alone."""
    (sa, va, ga, fa), (sb, vb, gb, fb), b = _batch_of_two(gpu_ctx, 420)
    b.render_checked(graph=True)
    many = np.zeros(1000, dtype=abi.LIGHT)
    many["pos"][:, :3] = sa.lights["pos"][0, :3]
    many["col"][:, :3] = 1.0
    many["shadow"] = 1
    many["brightness"] = 1.0
    many["radius"] = 20000.0
    many["diffuse"] = 1.0
    with pytest.raises(runtime.OcrError) as ei:
        ga.upload_lights(many, 4000)
    assert ei.value.code == abi.OCR_ERR_OOM, ei.value
    unlit = scenes.Scene(sa.tris, sa.objs, np.zeros(0, dtype=abi.LIGHT), sa.textures, 256)
    ref = runtime.Frame(gpu_ctx, fa.width, fa.height).render_read(runtime.GpuScene(gpu_ctx, unlit), va)
    ref_b = runtime.Frame(gpu_ctx, fb.width, fb.height).render_read(gb, vb)
    for graph in (True, False):
        b.render_checked(graph=graph)
        got = fa.read()
        for k in ("depth", "tri_id", "rgba", "normals"):
            assert np.array_equal(got[k], ref[k]), (graph, k)
        assert np.array_equal(fb.read()["rgba"], ref_b["rgba"]), graph
    # a later upload restores the lights
    ga.upload_lights(sa.lights, sa.light_buffer_dim)
    b.render_checked(graph=True)
    alone = runtime.Frame(gpu_ctx, fa.width, fa.height).render_read(ga, va)
    assert np.array_equal(fa.read()["rgba"], alone["rgba"])
    b.close()


def test_frame_serves_render_or_batches(gpu_ctx):
    (sa, va, ga, fa), (sb, vb, gb, fb), b = _batch_of_two(gpu_ctx, 430)
    b.render_checked()
    # a batch frame: ocr_render and the passes that need the frame's own
    # cut list and counters refuse it
    for call in (lambda: fa.render(ga, va), lambda: fa.stats(), lambda: fa.count_texels(ga, va),
                 lambda: fa.motion_blur(ga, va, va["c_pos"][0], va["c_rot"][0], 0)):
        with pytest.raises(runtime.OcrError) as ei:
            call()
        assert ei.value.code == abi.OCR_ERR_INVALID
    # screen passes and reads still work on it
    fa.pseudo_aa()
    fa.read()
    # a frame rendered by ocr_render cannot join a batch
    fr = runtime.Frame(gpu_ctx, fb.width, fb.height)
    fr.render(gb, vb)
    b2 = runtime.Batch(gpu_ctx, 2)
    with pytest.raises(runtime.OcrError) as ei:
        b2.set_views([gb], [vb], [fr])
    assert ei.value.code == abi.OCR_ERR_INVALID
    # the batch frames can move to another batch
    b.close()
    b2.set_views([ga, gb], [va, vb], [fa, fb])
    b2.render_checked(graph=True)
    alone = runtime.Frame(gpu_ctx, fa.width, fa.height).render_read(ga, va)
    assert np.array_equal(fa.read()["rgba"], alone["rgba"])
    b2.close()


def test_server_frame_pool_within_budget(gpu_ctx):
    """Requests of alternating sizes, one at a time, under a budget of a
    little over two 640x480 frames: the pool frees idle frames of other sizes
    instead of growing past the budget, and every result still equals the
    view rendered alone."""
    sc, _ = scenes.random_scene(440, n_tris=200)
    g = runtime.GpuScene(gpu_ctx, sc)
    free, _ = gpu_ctx.free_hbm()
    budget = int(2.2 * 20 * 640 * 480)
    srv = runtime.Server(gpu_ctx, max_batch_views=2, max_inflight=1, max_batch_tris=1 << 16,
                         hbm_fraction=budget / free, read_mask=runtime.READ_RGBA)
    slot = srv.add_scene(g)
    sizes = [(640, 480), (320, 240), (480, 360), (600, 400), (640, 480), (200, 150), (560, 420)]
    alone = {}
    for i in range(2 * len(sizes)):
        w, h = sizes[i % len(sizes)]
        view = scenes.make_view(w, h, (0.0, 0.0, -10.0), hfov=60.0)
        rid = srv.submit(slot, view)
        assert rid is not None, (w, h)
        t_end = time.time() + 30
        done = []
        while not done and time.time() < t_end:
            done = srv.poll(wait_ms=50)
        assert done and done[0].status == 0, (w, h)
        info = srv.info()
        assert info["frame_bytes"] <= info["hbm_budget"], info
        if (w, h) not in alone:
            alone[(w, h)] = runtime.Frame(gpu_ctx, w, h).render_read(g, view)["rgba"]
        assert np.array_equal(srv.outputs(done[0])["rgba"], alone[(w, h)]), (w, h)
        srv.release(rid)
    assert srv.info()["n_evicted"] > 0
    srv.close()


def test_profiling_modes(gpu_ctx):
    """ocr_frame_set_profiling: 1 times every pass of every render, 2 the
    shading pass of every render; other values are refused."""
    sc, view = scenes.random_scene(77, n_tris=200)
    g = runtime.GpuScene(gpu_ctx, sc)
    fr = runtime.Frame(gpu_ctx, int(view["width"][0]), int(view["height"][0]))
    ref = fr.render_read(g, view)
    for mode, n_renders in ((1, 5), (2, 7)):
        fr.set_profiling(True, shade_only=mode == 2)
        for _ in range(n_renders):
            fr.render(g, view)
        st = fr.stats()
        print(f"profiling mode {mode}: {st['n_profiled']} renders timed", flush=True)
        assert st["n_profiled"] == n_renders
        assert st["ms_shade"] > 0.0
        if mode == 1:
            assert st["ms_depth"] > 0.0 and st["ms_prearrange"] > 0.0 and st["ms_total"] >= st["ms_shade"]
        else:
            assert st["ms_depth"] == 0.0 and st["ms_prearrange"] == 0.0
    fr.set_profiling(False)
    out = fr.read()
    assert np.array_equal(out["depth"], ref["depth"]) and np.array_equal(out["rgba"], ref["rgba"])
    for bad in (3, -1):
        with pytest.raises(runtime.OcrError) as e:
            runtime.check(fr.lib.ocr_frame_set_profiling(fr.h, bad))
        assert e.value.code == abi.OCR_ERR_INVALID
