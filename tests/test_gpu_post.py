"""The screen-space passes on the GPU against the oracle's restatements, on
frames the path rendered: blend_screens_with_depth
(ocr_frame_blend_with_depth: bit-exact RGBA8, the destination depth
unchanged; sizes include one whose pixel count is not a multiple of 4, the
kernel's tail), do_pseudo_aa (bit-exact) and do_motion_blur (RGBA8 within
one unit, the recorded poses bit-exact, two passes of both parities)."""
import numpy as np
import pytest

from openclrenderer_amd import runtime, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h,seed", [(37, 23, 1), (640, 480, 2), (1920, 1080, 3)])
def test_blend_with_depth_matches_oracle(gpu_ctx, oracle, w, h, seed):
    sa, va = scenes.random_scene(500 + seed, n_tris=400, width=w, height=h)
    sb, vb = scenes.random_scene(600 + seed, n_tris=400, width=w, height=h)
    ga, gb = runtime.GpuScene(gpu_ctx, sa), runtime.GpuScene(gpu_ctx, sb)
    fa, fb = runtime.Frame(gpu_ctx, w, h), runtime.Frame(gpu_ctx, w, h)
    src = fa.render_read(ga, va)
    dst = fb.render_read(gb, vb)
    fb.blend_from(fa)
    out = fb.read()
    want = oracle.blend_with_depth(src["depth"], dst["depth"], src["rgba"], dst["rgba"])
    assert np.array_equal(out["rgba"], want.reshape(out["rgba"].shape))
    assert np.array_equal(out["depth"], dst["depth"])
    blended = dst["depth"].reshape(-1) >= src["depth"].reshape(-1)
    assert blended.any() and (~blended).any()


def test_blend_rejects_mismatched_frames(gpu_ctx):
    fa, fb = runtime.Frame(gpu_ctx, 64, 48), runtime.Frame(gpu_ctx, 64, 40)
    with pytest.raises(runtime.OcrError):
        fb.blend_from(fa)
    with pytest.raises(runtime.OcrError):
        fa.blend_from(fa)


@pytest.mark.parametrize("w,h,seed", [(37, 23, 4), (640, 480, 5), (1920, 1080, 6)])
def test_pseudo_aa_matches_oracle(gpu_ctx, oracle, w, h, seed):
    sc, view = scenes.random_scene(700 + seed, n_tris=600, width=w, height=h)
    g = runtime.GpuScene(gpu_ctx, sc)
    fr = runtime.Frame(gpu_ctx, w, h)
    before = fr.render_read(g, view)
    fr.pseudo_aa()
    after = fr.read()
    want = oracle.pseudo_aa(before["depth"], before["normals"], before["rgba"])
    assert np.array_equal(after["rgba"], want.reshape(after["rgba"].shape))
    assert np.array_equal(after["depth"], before["depth"])
    if w >= 640:
        assert (after["rgba"] != before["rgba"]).any()
    # a second pass reads the smoothed colours
    fr.pseudo_aa()
    again = fr.read()
    assert np.array_equal(again["rgba"], oracle.pseudo_aa(after["depth"], after["normals"], after["rgba"]).reshape(
        again["rgba"].shape))


def _moving_scene(seed, w, h):
    """A random scene whose objects were elsewhere last frame: pose slot 1
    displaced by up to 15 units and turned a little, slot 2 by other amounts."""
    sc, view = scenes.random_scene(900 + seed, n_tris=600, width=w, height=h)
    rng = np.random.default_rng(seed)
    for k, spread in ((1, 15.0), (2, 6.0)):
        sc.objs[f"old_world_pos_{k}"][:, :3] = sc.objs["world_pos"][:, :3] + rng.uniform(-spread, spread, (len(sc.objs), 3))
        q = sc.objs["world_rot_quat"] + rng.normal(scale=0.05, size=(len(sc.objs), 4))
        sc.objs[f"old_world_rot_quat_{k}"] = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    return sc, view


def _pose_fields(objs):
    return np.stack([objs["old_world_pos_1"], objs["old_world_pos_2"], objs["old_world_rot_quat_1"],
                     objs["old_world_rot_quat_2"]], axis=1)


@pytest.mark.parametrize("w,h,seed,cc,strength", [(37, 23, 1, 1.0, 1.0), (640, 480, 2, 0.5, 1.0),
                                                   (1920, 1080, 3, 1.0, 3.0), (1920, 1080, 4, 0.0, 0.5)])
def test_motion_blur_matches_oracle(gpu_ctx, oracle, w, h, seed, cc, strength):
    sc, view = _moving_scene(seed, w, h)
    g = runtime.GpuScene(gpu_ctx, sc)
    assert np.array_equal(g.read_motion(), _pose_fields(sc.objs))
    fr = runtime.Frame(gpu_ctx, w, h)
    before = fr.render_read(g, view)
    c_pos_old = view["c_pos"][0] + np.array([12.0, -7.0, 20.0, 0.0], np.float32)
    c_rot_old = view["c_rot"][0] + np.array([0.01, -0.02, 0.015, 0.0], np.float32)
    fr.motion_blur(g, view, c_pos_old, c_rot_old, 0, strength, cc)
    after = fr.read()
    want, wobjs = oracle.motion_blur(view, before, sc.tris, sc.objs, c_pos_old, c_rot_old, 0, strength, cc)
    diff = np.abs(after["rgba"].astype(np.int16) - want.astype(np.int16))
    assert diff.max() <= 1, int(diff.max())
    assert np.array_equal(g.read_motion(), _pose_fields(wobjs))
    assert np.array_equal(after["depth"], before["depth"])
    if w >= 640:
        assert (after["rgba"] != before["rgba"]).any()
    # the next parity reads the pose this pass recorded and records the other slot
    fr.motion_blur(g, view, c_pos_old, c_rot_old, 1, strength, cc)
    again = fr.read()
    want2, wobjs2 = oracle.motion_blur(view, after, sc.tris, wobjs, c_pos_old, c_rot_old, 1, strength, cc)
    diff = np.abs(again["rgba"].astype(np.int16) - want2.astype(np.int16))
    assert diff.max() <= 1, int(diff.max())
    assert np.array_equal(g.read_motion(), _pose_fields(wobjs2))


def test_motion_blur_argument_checks(gpu_ctx):
    sc, view = scenes.random_scene(950, n_tris=50, width=64, height=48)
    g = runtime.GpuScene(gpu_ctx, sc)
    fr = runtime.Frame(gpu_ctx, 64, 48)
    with pytest.raises(runtime.OcrError):          # nothing rendered yet
        fr.motion_blur(g, view, (0, 0, 0), (0, 0, 0), 0)
    fr.render_read(g, view)
    _, other = scenes.random_scene(951, n_tris=50, width=64, height=40)
    with pytest.raises(runtime.OcrError):          # view size differs from the frame's
        fr.motion_blur(g, other, (0, 0, 0), (0, 0, 0), 0)
