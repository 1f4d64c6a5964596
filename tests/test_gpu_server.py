"""The served form: requests batched with others give each request the same
depth, triangle ids and RGBA as rendering its view alone (and as the
oracle); admission refuses what the HBM budget cannot hold and admits again
once completions are released."""
import time

import numpy as np
import pytest

from openclrenderer_amd import abi, runtime, scenes

pytestmark = pytest.mark.gpu

ALL = runtime.READ_RGBA | runtime.READ_DEPTH | runtime.READ_IDKEY


def _wait_all(srv, ids, timeout_s=60.0):
    got = {}
    t_end = time.time() + timeout_s
    while len(got) < len(ids) and time.time() < t_end:
        for c in srv.poll(wait_ms=50):
            got[c.request_id] = (c, srv.outputs(c))
    assert len(got) == len(ids), f"{len(got)} of {len(ids)} requests completed"
    return got


def _scenes(gpu_ctx):
    sc1, v1 = scenes.config1(shadow=1)
    out = [(sc1, v1)]
    for k in range(2):
        out.append(scenes.random_scene(300 + k, n_tris=80 + 40 * k))
    gs = [runtime.GpuScene(gpu_ctx, s) for s, _ in out]
    return out, gs


def test_alone_vs_mix_and_oracle(gpu_ctx, oracle):
    sv, gs = _scenes(gpu_ctx)
    srv = runtime.Server(gpu_ctx, max_batch_views=8, max_inflight=2, max_batch_tris=1 << 20,
                         hbm_fraction=0.05, read_mask=ALL)
    slots = [srv.add_scene(g) for g in gs]
    rng = np.random.default_rng(7)
    reqs = []
    for i in range(14):
        k = i % 3
        if k == 0:
            w, h = 800, 600
            view = sv[0][1]
        else:
            w, h = int(rng.integers(24, 200)), int(rng.integers(16, 160))
            view = scenes.make_view(w, h, (0.0, 0.0, -10.0), c_rot=tuple(rng.uniform(-0.2, 0.2, 3)),
                                    hfov=float(rng.uniform(50, 80)))
        rid = srv.submit(slots[k], view)
        assert rid is not None
        reqs.append((rid, k, view))
    got = _wait_all(srv, [r[0] for r in reqs])
    assert max(c.batch_views for c, _ in got.values()) > 1, "no request shared a batch"
    maps = oracle.shadow_maps(sv[0][0])
    for rid, k, view in reqs:
        c, o = got[rid]
        assert c.status == 0
        alone = runtime.Frame(gpu_ctx, int(view["width"][0]), int(view["height"][0])).render_read(gs[k], view)
        tri = runtime.id_keys_to_tri(o["id_key"], o["depth"])
        assert np.array_equal(o["depth"], alone["depth"]), rid
        assert np.array_equal(tri, alone["tri_id"]), rid
        assert np.array_equal(o["rgba"], alone["rgba"]), rid
        if k == 0 and rid == reqs[0][0]:
            ref = oracle.render(sv[0][0], view, maps=maps)
            assert np.array_equal(o["depth"], ref["depth"])
            assert np.array_equal(tri, ref["tri_id"])
            assert np.abs(o["rgba"].astype(int) - ref["rgba"].astype(int)).max() <= 1
        srv.release(rid)
    info = srv.info()
    assert info["n_done"] == len(reqs) and info["hbm_held"] == 0
    srv.close()


def test_admission_by_hbm_budget(gpu_ctx):
    sc, view = scenes.random_scene(11, n_tris=60)
    g = runtime.GpuScene(gpu_ctx, sc)
    free, _ = gpu_ctx.free_hbm()
    w, h = int(view["width"][0]), int(view["height"][0])
    one = 20 * w * h
    srv = runtime.Server(gpu_ctx, max_batch_views=4, max_inflight=1, max_batch_tris=1 << 16,
                         hbm_fraction=min(1.0, 1.5 * one / free), read_mask=runtime.READ_RGBA)
    slot = srv.add_scene(g)
    assert srv.info()["hbm_budget"] < 2 * one
    r1 = srv.submit(slot, view)
    assert r1 is not None
    assert srv.submit(slot, view) is None          # budget holds one frame
    _wait_all(srv, [r1])
    assert srv.submit(slot, view) is None          # completed but not released
    srv.release(r1)
    r2 = srv.submit(slot, view)
    assert r2 is not None
    _wait_all(srv, [r2])
    srv.release(r2)
    assert srv.info()["n_refused"] == 2
    with pytest.raises(runtime.OcrError) as ei:
        srv.release(r2)
    assert ei.value.code == abi.OCR_ERR_INVALID
    srv.close()


def test_one_request_per_batch_matches(gpu_ctx):
    sv, gs = _scenes(gpu_ctx)
    srv = runtime.Server(gpu_ctx, max_batch_views=1, max_inflight=1, max_batch_tris=1 << 20,
                         hbm_fraction=0.05, read_mask=runtime.READ_RGBA)
    slots = [srv.add_scene(g) for g in gs]
    ids = [srv.submit(slots[k % 3], sv[k % 3][1]) for k in range(6)]
    got = _wait_all(srv, ids)
    for i, rid in enumerate(ids):
        c, o = got[rid]
        assert c.batch_views == 1
        k = i % 3
        alone = runtime.Frame(gpu_ctx, c.width, c.height).render_read(gs[k], sv[k][1])
        assert np.array_equal(o["rgba"], alone["rgba"])
        srv.release(rid)
    srv.close()


def test_scene_added_while_serving(gpu_ctx):
    """A scene created and registered while earlier requests are in flight
    serves its own requests, and every result equals the request rendered
    alone."""
    sv, gs = _scenes(gpu_ctx)
    srv = runtime.Server(gpu_ctx, max_batch_views=4, max_inflight=2, max_batch_tris=1 << 20,
                         hbm_fraction=0.05, read_mask=runtime.READ_RGBA)
    slot0 = srv.add_scene(gs[1])
    first = [srv.submit(slot0, sv[1][1]) for _ in range(12)]
    assert all(r is not None for r in first)
    sc_new, view_new = scenes.random_scene(321, n_tris=150)
    g_new = runtime.GpuScene(gpu_ctx, sc_new)
    slot1 = srv.add_scene(g_new)
    second = [srv.submit(slot1, view_new) for _ in range(6)]
    assert all(r is not None for r in second)
    got = _wait_all(srv, first + second)
    alone0 = runtime.Frame(gpu_ctx, int(sv[1][1]["width"][0]), int(sv[1][1]["height"][0])).render_read(gs[1], sv[1][1])
    alone1 = runtime.Frame(gpu_ctx, int(view_new["width"][0]), int(view_new["height"][0])).render_read(g_new, view_new)
    for rid in first:
        assert np.array_equal(got[rid][1]["rgba"], alone0["rgba"])
        srv.release(rid)
    for rid in second:
        assert got[rid][0].scene_slot == slot1
        assert np.array_equal(got[rid][1]["rgba"], alone1["rgba"])
        srv.release(rid)
    srv.close()


def test_updates_while_serving(gpu_ctx):
    """The application thread updates a served scene's objects (stream
    ordered, alternating between two states) while the server's dispatcher
    renders requests for it: every result is the scene in one state or the
    other, never a mixture (a batch reads one geometry version)."""
    sc, view = scenes.random_scene(330, n_tris=300, width=96, height=72)
    objs_a = sc.objs.copy()
    objs_b = sc.objs.copy()
    objs_b["world_pos"][:, 0] += np.array([35.0, -25.0, 15.0, -45.0], np.float32)[: len(objs_b)]
    objs_b["world_pos"][:, 1] -= 20.0
    refs = []
    for objs in (objs_a, objs_b):
        s2 = scenes.Scene(sc.tris, objs, sc.lights, sc.textures, sc.light_buffer_dim)
        refs.append(runtime.Frame(gpu_ctx, 96, 72).render_read(runtime.GpuScene(gpu_ctx, s2), view))
    assert not np.array_equal(refs[0]["depth"], refs[1]["depth"])
    g = runtime.GpuScene(gpu_ctx, sc)
    srv = runtime.Server(gpu_ctx, max_batch_views=4, max_inflight=2, max_batch_tris=1 << 20,
                         hbm_fraction=0.05, read_mask=runtime.READ_RGBA | runtime.READ_DEPTH)
    slot = srv.add_scene(g)
    ids = []
    for i in range(48):
        ids.append(srv.submit(slot, view))
        g.update_objs_async(0, objs_b if i % 2 == 0 else objs_a)
    assert all(r is not None for r in ids)
    got = _wait_all(srv, ids)
    seen = [0, 0]
    for rid in ids:
        o = got[rid][1]
        match = [np.array_equal(o["depth"], r["depth"]) and np.array_equal(o["rgba"], r["rgba"]) for r in refs]
        assert any(match), rid
        seen[match.index(True)] += 1
        srv.release(rid)
    srv.close()
    print("requests per state", seen)
