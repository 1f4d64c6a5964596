"""Pins of the oracle's do_motion_blur (cl2.cl:6715-6858): each case puts one
flat object across a synthetic frame and moves the camera or the object by a
known amount, so that the screen motion of every pixel has a closed form
from projection geometry (a translation of t units at depth Z moves a pixel
by t * fov / Z; a quarter turn about the view axis maps a screen offset
(dx, dy) to (-dy, dx)).  The expected pixel is the blur definition
evaluated in float64 on that vector, with scipy's linear interpolation as
the sampler, so a wrong vector (sign, scale, camera or pose slot), sample
count, start point or step fails by many units of the 8-bit output."""
import math

import numpy as np
import pytest

from openclrenderer_amd import abi

ndimage = pytest.importorskip("scipy.ndimage")

W, H = 63, 48                     # odd width: pixel offsets from W/2 are half-integers
FOV = 100.0
D_RAW = 12271336                  # quantised depth: Z = D_RAW / 2^32 * 350000 ~ 1000 units
Z = D_RAW / 2.0**32 * 350000.0
RNG = np.random.default_rng(4242)
QUAT_ID = (0.0, 0.0, 0.0, 1.0)


def _view(c_pos=(0, 0, 0), c_rot=(0, 0, 0)):
    v = np.zeros(1, abi.VIEW)
    v["width"], v["height"], v["fov_const"] = W, H, FOV
    v["c_pos"][0, :3] = c_pos
    v["c_rot"][0, :3] = c_rot
    return v


def _scene(wp=(0, 0, 0), q=QUAT_ID, old1=None, old2=None):
    tris = np.zeros(1, abi.TRIANGLE)
    objs = np.zeros(2, abi.OBJ_DESC)               # object 1 is never drawn
    for o in range(2):
        objs[o]["world_pos"][:3] = wp
        objs[o]["world_rot_quat"] = q
        for k in (1, 2):
            objs[o][f"old_world_pos_{k}"][:3] = wp
            objs[o][f"old_world_rot_quat_{k}"] = q
    for k, pose in ((1, old1), (2, old2)):
        if pose is not None:
            objs[0][f"old_world_pos_{k}"][:3] = pose[0]
            objs[0][f"old_world_rot_quat_{k}"] = pose[1]
    return tris, objs


def _frame(img=None):
    if img is None:
        img = RNG.integers(0, 256, (H, W, 4), dtype=np.uint8)
    return {"depth": np.full((H, W), D_RAW, np.uint32), "tri_id": np.zeros((H, W), np.uint32), "rgba": img}


def _expected(img, x, y, vx, vy):
    """The blur at (x, y) along (vx, vy): n = trunc(max(|vx|, |vy|) + 1)
    samples (at most 50, the vector scaled to match), from (x, y) - v / 2 in
    steps of v / n, in-frame samples averaged."""
    n = int(max(abs(vx), abs(vy)) + 1.0)
    if n > 50:
        vx, vy, n = vx / n * 50.0, vy / n * 50.0, 50
    f = img.astype(np.float64) / 255.0
    xs = x - vx / 2.0 + np.arange(n) * (vx / n)
    ys = y - vy / 2.0 + np.arange(n) * (vy / n)
    keep = (xs >= 0) & (xs < W) & (ys >= 0) & (ys < H)
    out = np.zeros(4)
    for c in range(4):
        out[c] = ndimage.map_coordinates(f[..., c], [ys[keep], xs[keep]], order=1, mode="nearest").mean()
    return np.floor(np.clip(out, 0.0, 1.0) * 255.0 + 0.5)


def _check(out, img, vec, pixels):
    for (x, y) in pixels:
        vx, vy = vec(x, y)
        exp = _expected(img, x, y, vx, vy)
        got = out[y, x].astype(np.float64)
        assert np.abs(got - exp).max() <= 1, (x, y, vx, vy, got, exp)


PIXELS = [(x, y) for y in (5, 17, 24, 40) for x in (3, 20, 31, 45, 60)]


def test_sampler_is_bilinear_with_edge_clamp(oracle):
    img = RNG.integers(0, 256, (H, W, 4), dtype=np.uint8)
    f = img.astype(np.float64) / 255.0
    for _ in range(200):
        cx, cy = RNG.uniform(-0.4, W - 0.6), RNG.uniform(-0.4, H - 0.6)
        got = oracle.sample_linear(img, np.float32(cx + 0.5), np.float32(cy + 0.5))
        # the coordinate the sampler sees, after the two float32 roundings
        tx = float(np.float32(np.float32(cx + 0.5) - np.float32(0.5)))
        ty = float(np.float32(np.float32(cy + 0.5) - np.float32(0.5)))
        exp = [ndimage.map_coordinates(f[..., c], [[ty], [tx]], order=1, mode="nearest")[0] for c in range(4)]
        assert np.abs(got - np.array(exp)).max() < 1e-6
    # texel centres read the texel
    for (x, y) in PIXELS:
        assert np.array_equal(oracle.sample_linear(img, x + 0.5, y + 0.5), (img[y, x] / np.float32(255)).astype(np.float32))


def test_no_motion_is_identity_and_records_pose(oracle):
    tris, objs = _scene(wp=(5, -3, 2), q=(0.1, 0.2, 0.3, 0.9))
    fr = _frame()
    for frame_id in (6, 7):
        out, ob = oracle.motion_blur(_view(), fr, tris, objs, (0, 0, 0), (0, 0, 0), frame_id)
        # a rounding-sized vector: one sample at the pixel itself, except that
        # in row / column 0 it can fall at -epsilon, outside the frame, where
        # no sample is taken and the pixel becomes 0
        assert np.array_equal(out[1:, 1:], fr["rgba"][1:, 1:])
        edge = np.concatenate([out[0], out[1:, 0]]), np.concatenate([fr["rgba"][0], fr["rgba"][1:, 0]])
        assert all(np.array_equal(a, b) or not a.any() for a, b in zip(*edge))
        assert np.array_equal(ob[1], objs[1])               # not drawn: untouched
    # the recorded slot holds the current pose (pos xyz only; w untouched)
    o = objs.copy()
    o[0]["old_world_pos_1"] = (9, 9, 9, 7)
    o[0]["old_world_rot_quat_1"] = (0, 0, 0, 1)
    _, ob = oracle.motion_blur(_view(), fr, tris, o, (0, 0, 0), (0, 0, 0), 4)   # even: reads 1, writes 2
    assert np.array_equal(ob[0]["old_world_pos_1"], o[0]["old_world_pos_1"])
    assert np.array_equal(ob[0]["old_world_pos_2"][:3], o[0]["world_pos"][:3])
    assert ob[0]["old_world_pos_2"][3] == o[0]["old_world_pos_2"][3]
    assert np.array_equal(ob[0]["old_world_rot_quat_2"], o[0]["world_rot_quat"])


def test_camera_translation(oracle):
    dxc = 73.0                       # the camera moved +73 units in x: 7.3 px of motion
    tris, objs = _scene()
    fr = _frame()
    out, _ = oracle.motion_blur(_view((dxc, 0, 0)), fr, tris, objs, (0, 0, 0), (0, 0, 0), 2)
    _check(out, fr["rgba"], lambda x, y: (-dxc * FOV / Z, 0.0), PIXELS)
    # the same motion in y
    out, _ = oracle.motion_blur(_view((0, dxc, 0)), fr, tris, objs, (0, 0, 0), (0, 0, 0), 2)
    _check(out, fr["rgba"], lambda x, y: (0.0, -dxc * FOV / Z), PIXELS)


@pytest.mark.parametrize("cc", [0.0, 0.5, 1.0])
def test_object_and_camera_weights(oracle, cc):
    """Object moved +ox (pose recorded in slot 1, read by even frame ids)
    and camera moved +dxc: the vector is (ox - cc * dxc) * fov / Z."""
    ox, dxc = 46.0, 113.0
    tris, objs = _scene(wp=(ox, 0, 0), old1=((0, 0, 0), QUAT_ID), old2=((1e4, 1e4, 0), QUAT_ID))
    fr = _frame()
    out, _ = oracle.motion_blur(_view((dxc, 0, 0)), fr, tris, objs, (0, 0, 0), (0, 0, 0), 8,
                                camera_contribution=cc)
    _check(out, fr["rgba"], lambda x, y: ((ox - cc * dxc) * FOV / Z, 0.0), PIXELS)
    # odd frame ids read slot 2: swap the slots and expect the same image
    tris, objs = _scene(wp=(ox, 0, 0), old1=((1e4, 1e4, 0), QUAT_ID), old2=((0, 0, 0), QUAT_ID))
    out2, _ = oracle.motion_blur(_view((dxc, 0, 0)), fr, tris, objs, (0, 0, 0), (0, 0, 0), 9,
                                 camera_contribution=cc)
    assert np.array_equal(out, out2)


def test_quarter_turn_about_the_view_axis(oracle):
    """The object (centred on the view axis) was turned by -90 degrees about
    z last frame: a screen offset (dx, dy) from the centre came from
    (-dy, dx), so the vector is (dx + dy, dy - dx)."""
    s = math.sqrt(0.5)
    tris, objs = _scene(wp=(0, 0, Z), old1=((0, 0, Z), (0, 0, s, s)))
    fr = _frame()
    out, _ = oracle.motion_blur(_view(), fr, tris, objs, (0, 0, 0), (0, 0, 0), 0)
    px = [(x, y) for (x, y) in PIXELS if abs(x - W / 2) + abs(y - H / 2) < 40]

    def vec(x, y):
        dx, dy = x - W / 2, y - H / 2
        return dx + dy, dy - dx
    _check(out, fr["rgba"], vec, px)


def test_camera_yaw_at_the_centre(oracle):
    """A yaw of d between frames moves the centre pixel by fov * tan(d)
    along x and not at all along y."""
    d = 0.06
    tris, objs = _scene()
    fr = _frame()
    out, _ = oracle.motion_blur(_view(), fr, tris, objs, (0, 0, 0), (0, d, 0), 0)
    x, y = 31, 24                     # 0.5 px left of the centre column, on the centre row
    shift = FOV * math.tan(d)
    exp = [_expected(fr["rgba"], x, y, sgn * shift, 0.0) for sgn in (1, -1)]
    got = out[y, x].astype(np.float64)
    assert min(np.abs(got - e).max() for e in exp) <= 1
    assert np.abs(got - fr["rgba"][y, x]).max() > 1          # the pixel was blurred


def test_strength_and_the_50_sample_bound(oracle):
    dxc = 413.0                      # 41.3 px, x3 = 123.9 px: scaled back to 50 samples
    tris, objs = _scene()
    fr = _frame()
    out, _ = oracle.motion_blur(_view((dxc, 0, 0)), fr, tris, objs, (0, 0, 0), (0, 0, 0), 2, strength=3.0)
    _check(out, fr["rgba"], lambda x, y: (-3.0 * dxc * FOV / Z, 0.0), PIXELS)
    out, _ = oracle.motion_blur(_view((dxc, 0, 0)), fr, tris, objs, (0, 0, 0), (0, 0, 0), 2, strength=0.25)
    _check(out, fr["rgba"], lambda x, y: (-0.25 * dxc * FOV / Z, 0.0), PIXELS)


def test_behind_the_previous_near_plane_keeps_the_pixel(oracle):
    tris, objs = _scene()
    fr = _frame()
    # the previous camera stood 990 units further forward: z = 10 < 20 there
    out, _ = oracle.motion_blur(_view(), fr, tris, objs, (0, 0, 990), (0, 0, 0), 0)
    assert np.array_equal(out, fr["rgba"])


def test_background_and_foreign_objects_untouched(oracle):
    tris, objs = _scene()
    fr = _frame()
    fr["depth"][:, :20] = 0xFFFFFFFF
    fr["tri_id"][:, :20] = 0xFFFFFFFF
    out, ob = oracle.motion_blur(_view((73, 0, 0)), fr, tris, objs, (0, 0, 0), (0, 0, 0), 0)
    assert np.array_equal(out[:, :20], fr["rgba"][:, :20])
    assert not np.array_equal(out[:, 20:], fr["rgba"][:, 20:])
    assert np.array_equal(ob[1], objs[1])
