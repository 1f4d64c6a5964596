"""The shadow launch's three ways of rasterising a piece give the same maps
as the oracle (prearrange_realtime_shadowing + kernel1_realtime_shadowing,
cl2.cl:4419-4636 and 5129-5246), bit for bit:

* pieces above the local threshold listed for the chip-wide raster launch,
  the rest rasterised by the prearrange launch's wavefronts;
* everything listed (threshold 0) or everything local (threshold above any
  piece);
* a list too small for the listed pieces: allocations past a partition's
  capacity fail, their wavefronts rasterise those pieces themselves and the
  raster launch reads only each partition's valid prefix (the list is split
  into 64 partitions; capacities 0, 1, 16, 1000 and 6400 records);
* the raster launch over 64 partitions whose chunks a wavefront takes in
  turn (sample indices restart in every partition), on the config-3
  generator at several thresholds;
* more shadow-casting lights than one launch takes (18 lights: two groups),
  with a static light among them.
"""
import numpy as np
import pytest

from openclrenderer_amd import runtime, scenes

pytestmark = pytest.mark.gpu
UINT_MAX = 0xFFFFFFFF


@pytest.fixture(scope="module")
def near_scene():
    """Config-1 geometry with the shadow-casting light among the cubes, so
    its faces hold pieces of one sample up to tens of thousands."""
    sc, view = scenes.config1(shadow=1)
    sc.lights["pos"][0, :3] = sc.objs["world_pos"][1, :3] + np.array([30.0, 40.0, -20.0], np.float32)
    return sc, view


@pytest.fixture(scope="module")
def near_maps(near_scene, oracle):
    return oracle.shadow_maps(near_scene[0], 16)


@pytest.mark.parametrize("local,cap", [("16", None), ("32", None), ("0", None), ("4000000000", None), ("0", "0"),
                                       ("0", "1"), ("0", "16"), ("8", "1000"), ("4", "6400")])
def test_shadow_maps_every_path(gpu_ctx, near_scene, near_maps, monkeypatch, local, cap):
    sc, _ = near_scene
    monkeypatch.setenv("OCR_SHADOW_LOCAL_SAMPLES", local)
    if cap is None:
        monkeypatch.delenv("OCR_SHADOW_LIST_RECORDS", raising=False)
    else:
        monkeypatch.setenv("OCR_SHADOW_LIST_RECORDS", cap)
    print(f"local {local} cap {cap}", flush=True)
    g = runtime.GpuScene(gpu_ctx, sc)
    g.shadow_pass()
    g.shadow_pass()    # the second pass reuses the list and counters
    dyn, _ = g.read_shadow_maps()
    odyn, _ = near_maps
    assert np.array_equal(dyn, odyn), f"{int((dyn != odyn).sum())} texels differ"
    assert (odyn != UINT_MAX).sum() > 100000


@pytest.fixture(scope="module")
def cfg3_small(oracle):
    sc, view = scenes.config3(n_inst=100, target_tris=200_000, width=960, height=540)
    return sc, oracle.shadow_maps(sc, 16)[0]


@pytest.mark.parametrize("local", ["8", "32", "64"])
def test_config3_partitions(gpu_ctx, cfg3_small, monkeypatch, local):
    sc, odyn = cfg3_small
    monkeypatch.setenv("OCR_SHADOW_LOCAL_SAMPLES", local)
    print(f"config-3 generator, 200k triangles, local {local}", flush=True)
    g = runtime.GpuScene(gpu_ctx, sc)
    g.shadow_pass()
    dyn, _ = g.read_shadow_maps()
    assert np.array_equal(dyn, odyn), f"{int((dyn != odyn).sum())} texels differ"


def test_eighteen_lights_two_groups(gpu_ctx, oracle):
    sc, view = scenes.config1(shadow=1)
    rng = np.random.default_rng(3)
    extra = [scenes.make_light(tuple(rng.uniform(-800, 800, 3)), shadow=1, is_static=int(i == 5))
             for i in range(17)]
    sc.lights = np.concatenate([sc.lights] + extra)
    assert int(np.sum(sc.lights["shadow"] == 1)) == 18
    g = runtime.GpuScene(gpu_ctx, sc)
    g.shadow_pass()
    dyn, stat = g.read_shadow_maps()
    odyn, ostat = oracle.shadow_maps(sc, 16)
    assert np.array_equal(dyn, odyn)
    assert np.array_equal(stat, ostat)
    fr = runtime.Frame(gpu_ctx, int(view["width"][0]), int(view["height"][0]))
    out = fr.render_read(g, view)
    ref = oracle.render(sc, view, maps=(odyn, ostat), nthreads=16)
    assert np.array_equal(out["depth"], ref["depth"])
    assert np.array_equal(out["tri_id"], ref["tri_id"])
    assert np.abs(out["rgba"].astype(int) - ref["rgba"].astype(int)).max() <= 1
