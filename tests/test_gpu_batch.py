"""The batched form: views of uneven size over several resident scenes in one
render (one launch per pass), graph replays and autotuned tunings all give
each view the same buffers as rendering it alone, and as the oracle."""
import numpy as np
import pytest

from openclrenderer_amd import runtime, scenes

pytestmark = pytest.mark.gpu


def _views(rng, n):
    out = []
    for i in range(n):
        w = int(rng.integers(20, 220))
        h = int(rng.integers(16, 180))
        out.append((i % 3, w, h, tuple(rng.uniform(-0.25, 0.25, 3)), float(rng.uniform(50, 80))))
    return out


def _make(gpu_ctx, seed=0, n=12):
    rng = np.random.default_rng(seed)
    scs = [scenes.random_scene(100 + k, n_tris=40 + 30 * k)[0] for k in range(3)]
    gs = [runtime.GpuScene(gpu_ctx, s) for s in scs]
    views, frames, owners = [], [], []
    for si, w, h, rot, hfov in _views(rng, n):
        v = scenes.make_view(w, h, (0.0, 0.0, -10.0), c_rot=rot, hfov=hfov)
        views.append(v)
        frames.append(runtime.Frame(gpu_ctx, w, h))
        owners.append(si)
    return scs, gs, views, frames, owners


def test_batch_equals_single_views_and_oracle(gpu_ctx, oracle):
    scs, gs, views, frames, owners = _make(gpu_ctx)
    b = runtime.Batch(gpu_ctx, 64)
    b.set_views([gs[o] for o in owners], views, frames)
    b.render_checked()
    for i, (v, f, o) in enumerate(zip(views, frames, owners)):
        got = f.read()
        alone = runtime.Frame(gpu_ctx, f.width, f.height).render_read(gs[o], v)
        for k in ("depth", "tri_id", "rgba", "normals"):
            assert np.array_equal(got[k], alone[k]), (i, k)
        ref = oracle.render(scs[o], v)
        assert np.array_equal(got["depth"], ref["depth"]), i
        assert np.array_equal(got["tri_id"], ref["tri_id"]), i
        assert np.abs(got["rgba"].astype(int) - ref["rgba"].astype(int)).max() <= 1, i
    b.close()


def test_batch_graph_replay_and_autotune(gpu_ctx):
    scs, gs, views, frames, owners = _make(gpu_ctx, seed=3, n=9)
    b = runtime.Batch(gpu_ctx, 16)
    b.set_views([gs[o] for o in owners], views, frames)
    b.render_checked()
    first = [f.read() for f in frames]
    for _ in range(3):
        b.render_checked(graph=True)
        for f, a in zip(frames, first):
            got = f.read()
            for k in a:
                assert np.array_equal(got[k], a[k]), k
    blocks, tile, groups, _rt = b.autotune(reps=2)
    assert 1 <= blocks <= 32 and 0 <= tile <= 8 and groups in (1, 2, 4)
    b.render_checked(graph=True)
    for f, a in zip(frames, first):
        got = f.read()
        for k in a:
            assert np.array_equal(got[k], a[k]), k
    # raster_tile != 0: the tiled raster path (tiles of 2^(rt & 255) x 2^(rt >> 8))
    for bl, tl, gr, rt in ((2, 3, 1, 0), (16, 6, 2, 0), (4, 5, 4, 0), (8, 4, 3, 0), (8, 4, 1, 4 | 4 << 8),
                           (8, 5, 2, 5 | 4 << 8), (8, 4, 4, 5 | 5 << 8), (8, 4, 1, 3 | 5 << 8), (8, 4, 3, 6 | 5 << 8)):
        b.set_tuning(bl, tl, gr, rt)
        assert b.tuning()[3] == rt
        for graph in (False, True):
            b.render_checked(graph=graph)
            for f, a in zip(frames, first):
                got = f.read()
                for k in a:
                    assert np.array_equal(got[k], a[k]), (bl, tl, gr, rt, graph, k)
    with pytest.raises(runtime.OcrError):
        b.set_tuning(8, 4, 1, 6 | 6 << 8)   # 4096-pixel tiles: beyond the 2048-pixel limit
    b.close()


def test_batch_with_config1_and_shadows(gpu_ctx, oracle):
    sc1, v1 = scenes.config1(shadow=1)
    g1 = runtime.GpuScene(gpu_ctx, sc1)
    g1.shadow_pass()
    scr, vr = scenes.random_scene(5)
    gr = runtime.GpuScene(gpu_ctx, scr)
    f1 = runtime.Frame(gpu_ctx, 800, 600)
    fr = runtime.Frame(gpu_ctx, int(vr["width"][0]), int(vr["height"][0]))
    b = runtime.Batch(gpu_ctx, 4)
    b.set_views([g1, gr], [v1, vr], [f1, fr])
    b.render_checked(graph=True)
    maps = oracle.shadow_maps(sc1)
    ref = oracle.render(sc1, v1, maps=maps)
    got = f1.read()
    assert np.array_equal(got["depth"], ref["depth"])
    assert np.array_equal(got["tri_id"], ref["tri_id"])
    assert np.abs(got["rgba"].astype(int) - ref["rgba"].astype(int)).max() <= 1
    b.close()


def test_batch_follows_stream_ordered_updates(gpu_ctx, oracle):
    """A batch (plain and graph launches) after ocr_scene_update_objs_async
    renders the updated objects: its scene table follows the update."""
    scs, gs, views, frames, owners = _make(gpu_ctx, seed=5, n=6)
    b = runtime.Batch(gpu_ctx, 16)
    b.set_views([gs[o] for o in owners], views, frames)
    b.render_checked(graph=True)
    for step, graph in ((1, True), (2, False), (3, True)):
        for si, (s, g) in enumerate(zip(scs, gs)):
            objs = s.objs.copy()
            objs["world_pos"][:, 0] += 7.0 * step * (si + 1)
            g.update_objs_async(0, objs)
            s.objs = objs
        b.render_checked(graph=graph)
        for i, (v, f, o) in enumerate(zip(views, frames, owners)):
            got = f.read()
            ref = oracle.render(scs[o], v)
            assert np.array_equal(got["depth"], ref["depth"]), (step, i)
            assert np.array_equal(got["tri_id"], ref["tri_id"]), (step, i)
    b.close()


def test_batch_after_light_and_atlas_uploads(gpu_ctx):
    """Uploads that replace a scene's buffers after set_views — lights (one
    now casting shadows, which also switches the shading kernel) and the
    atlas — are followed by the next batch launch, graph and plain: every
    view equals the view rendered alone after the same uploads."""
    scs, gs, views, frames, owners = _make(gpu_ctx, seed=11, n=6)
    b = runtime.Batch(gpu_ctx, 8)
    b.set_views([gs[o] for o in owners], views, frames)
    b.render_checked(graph=True)
    lights = scs[0].lights.copy()
    lights["shadow"][0] = 1
    gs[0].upload_lights(lights, 256)
    gs[0].shadow_pass()
    gs[1].upload_textures([scenes.value_noise_texture(64, 99, lattice=4)])
    for graph in (True, False, True):
        b.render_checked(graph=graph)
        for i, (v, f, o) in enumerate(zip(views, frames, owners)):
            got = f.read()
            alone = runtime.Frame(gpu_ctx, f.width, f.height).render_read(gs[o], v)
            for k in ("depth", "tri_id", "rgba", "normals"):
                assert np.array_equal(got[k], alone[k]), (graph, i, k)
    # and back to no shadow-casting light
    lights["shadow"][0] = 0
    gs[0].upload_lights(lights, 256)
    b.render_checked(graph=True)
    for i, (v, f, o) in enumerate(zip(views, frames, owners)):
        alone = runtime.Frame(gpu_ctx, f.width, f.height).render_read(gs[o], v)
        assert np.array_equal(f.read()["rgba"], alone["rgba"]), i
    b.close()
