"""The reference's draw API (engine / object_context / objects_container /
texture_context / light, include/ocr_engine.hpp) driven by a main.cpp-style
program renders the same frame as the C ABI and the oracle."""
import os
import subprocess

import numpy as np
import pytest

from openclrenderer_amd import abi, scenes
from openclrenderer_amd import build as b

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shadow,mode", [(0, "sync"), (1, "sync"), (1, "async"), (0, "motion")])
def test_facade_demo_matches_oracle(oracle, tmp_path, shadow, mode):
    """mode async: half the containers are built and drawn first, the rest
    are added by an asynchronous object_context::build while frames of the
    previous scene are drawn, and flip() makes the full scene current.
    mode motion: swap_buffers records the camera, the camera moves, one more
    frame is drawn and engine::do_motion_blur blurs it (objects still, so the
    motion is the camera's alone)."""
    sc, view = scenes.config1(shadow=shadow)
    scene_bin = tmp_path / "scene.bin"
    out_bin = tmp_path / "out.bin"
    scenes.dump_for_facade(sc, view, 60.0, str(scene_bin))
    assert os.path.exists(b.FACADE_DEMO), "build() first"
    r = subprocess.run([b.FACADE_DEMO, str(scene_bin), str(out_bin), "2", mode], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    if mode == "async":
        assert "async build:" in r.stdout, r.stdout
    raw = out_bin.read_bytes()
    used = np.frombuffer(raw[:64], dtype=abi.VIEW).copy()
    W, H = int(used["width"][0]), int(used["height"][0])
    n = W * H
    depth = np.frombuffer(raw[64:64 + 4 * n], np.uint32).reshape(H, W)
    tri = np.frombuffer(raw[64 + 4 * n:64 + 8 * n], np.uint32).reshape(H, W)
    rgba = np.frombuffer(raw[64 + 8 * n:64 + 12 * n], np.uint8).reshape(H, W, 4)
    assert abs(float(used["fov_const"][0]) - float(view["fov_const"][0])) < 1e-3
    ref = oracle.render(sc, used)
    assert np.array_equal(depth, ref["depth"])
    assert np.array_equal(tri, ref["tri_id"])
    if mode == "motion":
        assert "motion blur: frame_id 3" in r.stdout, r.stdout
        objs = sc.objs.copy()            # generate_gpu_object_descriptor: old poses = current
        for k in (1, 2):
            objs[f"old_world_pos_{k}"] = objs["world_pos"]
            objs[f"old_world_rot_quat_{k}"] = objs["world_rot_quat"]
        blurred, _ = oracle.motion_blur(used, ref, sc.tris, objs, view["c_pos"][0], view["c_rot"][0], 3)
        assert np.abs(rgba.astype(int) - blurred.astype(int)).max() <= 1
        assert (blurred != ref["rgba"]).any()
        return
    assert np.abs(rgba.astype(int) - ref["rgba"].astype(int)).max() <= 1
