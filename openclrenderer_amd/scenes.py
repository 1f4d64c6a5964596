"""Procedural scenes for the five workloads of the path.

World scale: 1 scene unit = 100 reference units, so geometry sits inside the
reference's depth_icutoff = 20 / depth_far = 350000 / SHADOWBIAS = 50 design
range (cl2.cl:17-23, ocl.h:234).  Meshes are stored in object space and placed
by their object descriptor (world_pos, world_rot_quat, scale), exactly as the
reference's objects_container / object_context do (object_context.cpp:228-339).
The triangle id of a triangle is its index in the concatenated array
(object_context.cpp:398-453).

RNG: SplitMix64 seeded with ``20261202 + config_index``; every generated array
is a pure function of the seed.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

SEED_BASE = 20261202
UNIT = 100.0


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014)."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def normal(self):
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def numpy_seed(self):
        return self.next_u64() & 0xFFFFFFFF


def fov_const_from_hfov(hfov_degrees, width):
    """engine.cpp:119-133 + :474-477: the FOV constant the reference bakes into
    its kernels: float arithmetic, then std::to_string (6 decimals) and parsed
    back as a float literal."""
    f32 = np.float32
    fov_radians = f32(float(f32(f32(hfov_degrees) / f32(360.0)) * f32(2)) * math.pi)
    triangle_angle = f32(fov_radians / f32(2))
    fov = f32(f32(f32(width) / f32(2)) / np.tan(triangle_angle))
    return float(np.float32(float("%f" % float(fov))))


# ------------------------------------------------------------------ meshes
def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v, f


_ICO_CACHE = {}


def icosphere(subdiv):
    """Unit icosphere, 20*4^subdiv triangles, (T,3,3) positions with outward
    winding: cross(p1-p0, p2-p0) points away from the centre."""
    if subdiv in _ICO_CACHE:
        return _ICO_CACHE[subdiv]
    v, f = _icosahedron()
    verts = [tuple(x) for x in v]
    faces = f.tolist()
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                p = np.add(verts[a], verts[b]) * 0.5
                p = p / np.linalg.norm(p)
                verts.append(tuple(p))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        faces = nf
    V = np.array(verts, dtype=np.float64)
    F = np.array(faces, dtype=np.int64)
    tri = V[F]
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    flip = np.einsum("ij,ij->i", n, tri.mean(axis=1)) < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    _ICO_CACHE[subdiv] = tri
    return tri


def sphere_mesh(subdiv):
    """Positions, radial normals and spherical UVs of a unit icosphere."""
    pos = icosphere(subdiv)
    nrm = pos.copy()
    u = 0.5 + np.arctan2(pos[..., 2], pos[..., 0]) / (2 * math.pi)
    v = 0.5 - np.arcsin(np.clip(pos[..., 1], -1, 1)) / math.pi
    vt = np.stack([u, v], axis=-1)
    return pos, nrm, vt


_CUBE_CACHE = {}


def cube_mesh(n):
    """Unit cube (edge 1, centred), each face an n x n grid of quads split in
    two triangles (12 n^2 triangles), face normals, per-face UVs 0..1."""
    if n in _CUBE_CACHE:
        return _CUBE_CACHE[n]
    tris, nrms, vts = [], [], []
    axes = [(0, 1, 2), (0, 1, 2)]
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            a1, a2 = [a for a in range(3) if a != axis]
            g = np.linspace(-0.5, 0.5, n + 1)
            uu = np.linspace(0.0, 1.0, n + 1)
            for i in range(n):
                for j in range(n):
                    quad = []
                    quv = []
                    for (di, dj) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = [0.0, 0.0, 0.0]
                        p[axis] = 0.5 * sgn
                        p[a1] = g[i + di]
                        p[a2] = g[j + dj]
                        quad.append(p)
                        quv.append((uu[i + di], uu[j + dj]))
                    for idx in ((0, 1, 2), (0, 2, 3)):
                        t = np.array([quad[k] for k in idx])
                        tv = np.array([quv[k] for k in idx])
                        nn = np.cross(t[1] - t[0], t[2] - t[0])
                        if nn[axis] * sgn < 0:
                            t = t[[0, 2, 1]]
                            tv = tv[[0, 2, 1]]
                        tris.append(t)
                        vts.append(tv)
                        fn = [0.0, 0.0, 0.0]
                        fn[axis] = sgn
                        nrms.append([fn, fn, fn])
    del axes
    out = (np.array(tris), np.array(nrms, dtype=np.float64), np.array(vts))
    _CUBE_CACHE[n] = out
    return out


# ------------------------------------------------------------------ scene
@dataclass
class Texture:
    width: int
    height: int
    rgba: np.ndarray          # (H, W, 4) uint8, row 0 first


@dataclass
class Scene:
    tris: np.ndarray          # abi.TRIANGLE
    objs: np.ndarray          # abi.OBJ_DESC
    lights: np.ndarray        # abi.LIGHT
    textures: list = field(default_factory=list)
    light_buffer_dim: int = 1024
    name: str = ""

    @property
    def ntri(self):
        return int(self.tris.shape[0])


def make_view(width, height, c_pos, c_rot=(0.0, 0.0, 0.0), hfov=60.0, linear=1,
              clear=(0.0, 0.0, 0.0, 0.0)):
    v = np.zeros(1, dtype=abi.VIEW)
    v["width"] = width
    v["height"] = height
    v["fov_const"] = fov_const_from_hfov(hfov, width)
    v["use_linear_rendering"] = linear
    v["c_pos"][0, :3] = c_pos
    v["c_rot"][0, :3] = c_rot
    v["clear_col"][0] = clear
    return v


def look_at_rot(c_pos, target):
    """Euler angles (rx, ry, 0) for which rot() (cl2.cl:220-236) maps
    `target` onto the +z axis in front of the camera."""
    rel = np.asarray(target, dtype=np.float64) - np.asarray(c_pos, dtype=np.float64)
    ry = math.atan2(rel[0], rel[2])
    d = math.hypot(rel[0], rel[2])
    rx = math.atan2(-rel[1], d)
    return (rx, ry, 0.0)


def value_noise_texture(size, seed, lattice=8):
    """RGBA8 value noise: an independent random lattice per colour channel,
    bilinearly interpolated, alpha = 255."""
    rng = np.random.default_rng(seed)
    out = np.zeros((size, size, 4), dtype=np.uint8)
    coords = np.arange(size) * (lattice / size)
    i0 = np.floor(coords).astype(np.int64)
    fr = coords - i0
    i1 = (i0 + 1) % lattice
    for ch in range(3):
        lat = rng.random((lattice, lattice))
        a = lat[i0][:, i0] * (1 - fr)[None, :] + lat[i0][:, i1] * fr[None, :]
        b = lat[i1][:, i0] * (1 - fr)[None, :] + lat[i1][:, i1] * fr[None, :]
        val = a * (1 - fr)[:, None] + b * fr[:, None]
        out[..., ch] = np.clip(np.round(val * 255.0), 0, 255).astype(np.uint8)
    out[..., 3] = 255
    return Texture(size, size, out)


def _quat_from_rng(rng):
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    return (math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
            math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
            math.sqrt(u1) * math.sin(2 * math.pi * u3),
            math.sqrt(u1) * math.cos(2 * math.pi * u3))


def _fill_tris(pos, nrm, vt, object_id, out):
    out["v"]["pos"][:, :, :3] = pos
    out["v"]["normal"][:, :, :3] = nrm
    out["v"]["vt"] = vt
    out["v"]["object_id"][:, 0] = object_id
    return out


def _desc(world_pos, quat, scale, tid=0, specular=0.9, spec_mult=1.0, diffuse=1.0, flags=0):
    d = np.zeros(1, dtype=abi.OBJ_DESC)
    d["world_pos"][0, :3] = world_pos
    d["world_rot_quat"][0] = quat
    d["old_world_pos_1"] = d["world_pos"]
    d["old_world_pos_2"] = d["world_pos"]
    d["old_world_rot_quat_1"] = d["world_rot_quat"]
    d["old_world_rot_quat_2"] = d["world_rot_quat"]
    d["scale"] = scale
    d["tid"] = tid
    d["rid"] = 0xFFFFFFFF
    d["ssid"] = 0xFFFFFFFF
    d["specular"] = specular           # object.cpp:72-74 defaults
    d["spec_mult"] = spec_mult
    d["diffuse"] = diffuse
    d["feature_flag"] = flags
    return d


def make_light(pos, col=(1.0, 1.0, 1.0), brightness=1.0, radius=20000.0, shadow=0,
               is_static=0, diffuse=1.0):
    l = np.zeros(1, dtype=abi.LIGHT)
    l["pos"][0, :3] = pos
    l["col"][0, :3] = col
    l["shadow"] = shadow
    l["brightness"] = brightness
    l["radius"] = radius
    l["diffuse"] = diffuse
    l["is_static"] = is_static
    return l


# ------------------------------------------------------------------ configs
def config1(shadow=0, seed_offset=0):
    """800x600 plain frame: icosphere subdiv 3 (1280 tris, radius 1 unit) at
    the origin plus a 10x10 floor of unit cubes (edge 1 unit, spacing 2 units,
    12 tris each) at y = +1.5 units spanning z in [-6, 12] units, so some cubes
    sit behind and across the near plane.  One 256^2 value-noise texture, one
    point light at (3, 3, -3) units, radius 20000.  Camera 5 units back at
    (0, 0, -5) looking +z, hfov 60 degrees."""
    rng = SplitMix64(SEED_BASE + 1 + seed_offset)
    tris, objs = [], []
    pos, nrm, vt = sphere_mesh(3)
    t = np.zeros(pos.shape[0], dtype=abi.TRIANGLE)
    tris.append(_fill_tris(pos, nrm, vt, 0, t))
    objs.append(_desc((0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0), UNIT))
    cpos, cnrm, cvt = cube_mesh(1)
    oid = 1
    for gz in range(10):
        for gx in range(10):
            t = np.zeros(cpos.shape[0], dtype=abi.TRIANGLE)
            tris.append(_fill_tris(cpos, cnrm, cvt, oid, t))
            wp = ((gx - 4.5) * 2 * UNIT, 1.5 * UNIT, (gz * 2 - 6) * UNIT)
            q = _quat_from_rng(rng)
            objs.append(_desc(wp, q, UNIT))
            oid += 1
    tex = value_noise_texture(256, rng.numpy_seed())
    lights = make_light((3 * UNIT, 3 * UNIT, -3 * UNIT), shadow=shadow)
    sc = Scene(np.concatenate(tris), np.concatenate(objs), lights, [tex], 1024, "cfg1")
    view = make_view(800, 600, (0.0, 0.0, -5 * UNIT))
    return sc, view


def _instance_scene(rng, n_inst, target_tris, box_units, cam_pos, fov, cam_target=(0, 0, 0),
                    median_px=4.0, sigma=2.0, ntex=1, n_lights=1, shadow_lights=0,
                    name=""):
    """Instanced spheres and subdivided cubes at seeded uniform positions in a
    box_units^3 box, per-instance subdivision chosen so the triangle total is
    within 0.5% of target_tris, and per-instance scale chosen so the expected
    screen-space triangle area follows a log-normal (median median_px px^2,
    clipped to [0.1, 5000] px^2) at the instance's distance from cam_pos."""
    half = 0.5 * box_units * UNIT
    kinds = []
    per = target_tris / n_inst
    sphere_opts = [(s, 20 * 4 ** s) for s in range(0, 7)]
    cube_opts = [(n, 12 * n * n) for n in range(1, 64)]
    total = 0
    for i in range(n_inst):
        is_sphere = (i % 2 == 0)
        remaining = target_tris - total
        left = n_inst - i
        want = remaining / left
        # jitter the per-instance budget so instances differ
        want *= math.exp(0.6 * rng.normal())
        want = max(want, 20.0)
        opts = sphere_opts if is_sphere else cube_opts
        best = min(opts, key=lambda o: abs(math.log(o[1] / want)))
        kinds.append(("s" if is_sphere else "c", best[0], best[1]))
        total += best[1]
    # trim / pad with cube resolution to land within 0.5% of the target
    guard = 0
    while abs(total - target_tris) > 0.005 * target_tris and guard < 100000:
        guard += 1
        j = int(rng.uniform() * n_inst)
        k, p, c = kinds[j]
        if k != "c":
            continue
        if total > target_tris and p > 1:
            np_ = p - 1
        elif total < target_tris and p < 63:
            np_ = p + 1
        else:
            continue
        total += 12 * np_ * np_ - c
        kinds[j] = ("c", np_, 12 * np_ * np_)
    del per
    tris, objs = [], []
    cam = np.asarray(cam_pos, dtype=np.float64)
    for oid, (k, p, c) in enumerate(kinds):
        wp = np.array([rng.uniform() * 2 - 1, rng.uniform() * 2 - 1, rng.uniform() * 2 - 1]) * half
        wp = wp + np.asarray(cam_target, dtype=np.float64)
        dist = max(np.linalg.norm(wp - cam), 1.0)
        a_px = math.exp(math.log(median_px) + sigma * rng.normal())
        a_px = min(max(a_px, 0.1), 5000.0)
        a_world = a_px * (dist / fov) ** 2
        if k == "s":
            pos, nrm, vt = sphere_mesh(p)
            surf = 4 * math.pi
        else:
            pos, nrm, vt = cube_mesh(p)
            surf = 6.0
        scale = math.sqrt(a_world * c / surf)
        t = np.zeros(c, dtype=abi.TRIANGLE)
        tris.append(_fill_tris(pos, nrm, vt, oid, t))
        tid = int(rng.uniform() * ntex) if ntex > 1 else 0
        objs.append(_desc(tuple(wp), _quat_from_rng(rng), scale, tid=tid))
    lights = []
    for li in range(n_lights):
        lp = np.array([rng.uniform() * 2 - 1, rng.uniform() * 2 - 1, rng.uniform() * 2 - 1]) * half * 1.2
        lp = lp + np.asarray(cam_target, dtype=np.float64)
        col = (0.4 + 0.6 * rng.uniform(), 0.4 + 0.6 * rng.uniform(), 0.4 + 0.6 * rng.uniform())
        lights.append(make_light(tuple(lp), col=col, shadow=1 if li < shadow_lights else 0))
    return np.concatenate(tris), np.concatenate(objs), np.concatenate(lights)


def config2(n_inst=400, target_tris=1_000_000, width=1920, height=1080):
    """1920x1080, 1M triangles: 400 instances (alternating icospheres and
    subdivided cubes) uniform in a 100^3-unit box centred on the origin,
    camera 150 units back on -z looking +z (hfov 60), one 2048^2 value-noise
    texture (full mip chain in the atlas), 4 point lights, no shadow maps."""
    rng = SplitMix64(SEED_BASE + 2)
    cam = (0.0, 0.0, -150 * UNIT)
    fov = fov_const_from_hfov(60.0, width)
    tris, objs, lights = _instance_scene(rng, n_inst, target_tris, 100, cam, fov, n_lights=4,
                                         name="cfg2")
    tex = value_noise_texture(2048, rng.numpy_seed(), lattice=32)
    sc = Scene(tris, objs, lights, [tex], 1024, "cfg2")
    view = make_view(width, height, cam)
    return sc, view


def config3(n_inst=1600, target_tris=4_000_000, width=3840, height=2160, n_lights=16):
    """3840x2160, 4M triangles (4x the config-2 instances), 16 shadow-casting
    lights with 6 x 1024^2 cube maps each, 8 material textures of 1024^2."""
    rng = SplitMix64(SEED_BASE + 3)
    cam = (0.0, 0.0, -150 * UNIT)
    fov = fov_const_from_hfov(60.0, width)
    tris, objs, lights = _instance_scene(rng, n_inst, target_tris, 100, cam, fov, ntex=8,
                                         n_lights=n_lights, shadow_lights=n_lights, name="cfg3")
    texs = [value_noise_texture(1024, rng.numpy_seed(), lattice=16) for _ in range(8)]
    sc = Scene(tris, objs, lights, texs, 1024, "cfg3")
    view = make_view(width, height, cam)
    return sc, view


def small_scene(seed, n_inst, target_tris, box_units=10.0, n_lights=1, shadow_lights=0,
                ntex=1, tex_size=256, name="scene"):
    """A resident scene for the batched and served workloads: the config-2
    generator in a box_units^3 box centred on the origin."""
    rng = SplitMix64(seed)
    cam = (0.0, 0.0, -2.0 * box_units * UNIT)
    fov = fov_const_from_hfov(60.0, 1280)
    tris, objs, lights = _instance_scene(rng, n_inst, target_tris, box_units, cam, fov,
                                         ntex=ntex, n_lights=n_lights,
                                         shadow_lights=shadow_lights, name=name)
    texs = [value_noise_texture(tex_size, rng.numpy_seed()) for _ in range(ntex)]
    return Scene(tris, objs, lights, texs, 1024, name)


def orbit_view(rng, width, height, radius_units, centre=(0.0, 0.0, 0.0)):
    """A camera on a sphere of the given radius around `centre`, looking at it."""
    while True:
        d = np.array([rng.normal(), rng.normal(), rng.normal()])
        n = np.linalg.norm(d)
        if n > 1e-6:
            break
    c = np.asarray(centre, dtype=np.float64) + d / n * radius_units * UNIT
    return make_view(width, height, tuple(c), look_at_rot(c, centre))


def config4_views(n_views=64, n_scenes=8, seed=SEED_BASE + 4):
    """64 views, W = round(320 + 2240 s), H = round(240 + 1200 s), s ~ U[0,1],
    each on one of n_scenes scenes, camera radius U[3, 20] units."""
    rng = SplitMix64(seed)
    out = []
    for i in range(n_views):
        s = rng.uniform()
        w = int(round(320 + 2240 * s))
        h = int(round(240 + 1200 * s))
        scene_idx = i % n_scenes
        r = 3 + 17 * rng.uniform()
        out.append((scene_idx, orbit_view(rng, w, h, r)))
    return out


def config4_scene_sizes(n_scenes=8, seed=SEED_BASE + 4):
    rng = SplitMix64(seed ^ 0x5EED)
    return [int(round(math.exp(math.log(50_000) + (math.log(2_000_000) - math.log(50_000)) * rng.uniform())))
            for _ in range(n_scenes)]


def config4_workload():
    """The batched workload: 8 resident scenes of log-uniform 50k-2M
    triangles (4 lights, one 1024^2 texture each) and the 64 views over them
    ((scene index, view) pairs)."""
    sizes = config4_scene_sizes()
    scs = [small_scene(SEED_BASE + 40 + i, max(20, min(400, n // 5000)), n, n_lights=4, ntex=1, tex_size=1024,
                       name=f"cfg4-{i}") for i, n in enumerate(sizes)]
    return scs, config4_views()


def config5_workload(seed=None):
    """The served workload: 12 resident scenes of log-uniform 50k-4M
    triangles, each with 1-16 lights that all cast shadows, and the seeded
    generator the request stream continues from."""
    rng = SplitMix64(SEED_BASE + 5 if seed is None else seed)
    scs = []
    for i in range(12):
        n = int(round(math.exp(math.log(50_000) + (math.log(4_000_000) - math.log(50_000)) * rng.uniform())))
        nl = 1 + int(rng.uniform() * 16)
        scs.append(small_scene(SEED_BASE + 50 + i, max(20, min(400, n // 5000)), n, n_lights=nl,
                               shadow_lights=nl, ntex=1, tex_size=1024, name=f"cfg5-{i}"))
    return scs, rng


CONFIG5_RESOLUTIONS = [(640, 480), (1280, 720), (1920, 1080), (3840, 2160)]
CONFIG5_RES_P = [0.4, 0.3, 0.2, 0.1]


def config5_requests(rng, n, rate, n_scenes=12):
    """n requests of the served stream: Poisson arrival times at `rate` per
    second, a scene index, and a view at a resolution drawn with
    CONFIG5_RES_P on a seeded orbit of radius 3-20 units."""
    cum = np.cumsum(CONFIG5_RES_P)
    t, out = 0.0, []
    for _ in range(n):
        t += -math.log(max(rng.uniform(), 1e-12)) / rate
        u = rng.uniform()
        w, h = CONFIG5_RESOLUTIONS[int(np.searchsorted(cum, u, side="right"))] if u < 1.0 else CONFIG5_RESOLUTIONS[-1]
        k = int(rng.uniform() * n_scenes)
        out.append((t, k, orbit_view(rng, w, h, 3 + 17 * rng.uniform())))
    return out


def random_scene(seed, n_tris=60, width=64, height=48, n_objs=4, near_cross=True,
                 duplicate_ties=True, vertex_colours=True, two_sided=True, shadow=0):
    """Small seeded scenes for parity tests: random triangles in front of a
    camera at the origin looking +z, some straddling the near plane, some
    duplicated in a later object (coplanar depth ties), some objects two-sided
    and some triangles vertex-coloured."""
    rng = np.random.default_rng(seed)
    objs = []
    for o in range(n_objs):
        flags = abi.FEATURE_TWO_SIDED if (two_sided and o % 2 == 1) else 0
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        objs.append(_desc((rng.uniform(-50, 50), rng.uniform(-50, 50), rng.uniform(200, 600)),
                          tuple(q), float(rng.uniform(0.5, 2.0)), flags=flags))
    tris = np.zeros(n_tris, dtype=abi.TRIANGLE)
    for t in range(n_tris):
        oid = int(rng.integers(0, n_objs))
        c = rng.uniform(-120, 120, 3)
        if near_cross and t % 7 == 0:
            c[2] = -rng.uniform(150, 450)   # pulls a vertex towards / behind the near plane
        for v in range(3):
            p = c + rng.normal(scale=rng.uniform(5, 80), size=3)
            tris["v"]["pos"][t, v, :3] = p
            n = rng.normal(size=3)
            tris["v"]["normal"][t, v, :3] = n / np.linalg.norm(n)
            tris["v"]["vt"][t, v] = rng.uniform(-0.5, 1.5, 2)
        tris["v"]["object_id"][t, 0] = oid
        if vertex_colours and t % 5 == 3:
            tris["v"]["vertex_col"][t] = rng.integers(1, 2 ** 32 - 1, 3, dtype=np.uint64).astype(np.uint32)
    if duplicate_ties and n_tris >= 4:
        for t in range(0, n_tris // 4, 3):
            src = n_tris - 1 - t
            tris[t]["v"]["pos"] = tris[src]["v"]["pos"]
            tris["v"]["object_id"][t, 0] = tris["v"]["object_id"][src, 0]
    objs = np.concatenate(objs)
    lights = np.concatenate([make_light((rng.uniform(-300, 300), rng.uniform(-300, 300), rng.uniform(-100, 100)),
                                        shadow=shadow, radius=20000.0),
                             make_light((0.0, -400.0, 300.0), col=(0.3, 0.5, 1.0), radius=900.0)])
    tex = value_noise_texture(64, int(rng.integers(0, 2 ** 31)), lattice=4)
    sc = Scene(tris, objs, lights, [tex], 64, f"random{seed}")
    view = make_view(width, height, (0.0, 0.0, -10.0), c_rot=tuple(rng.uniform(-0.2, 0.2, 3)), hfov=70.0)
    return sc, view


def dump_for_facade(sc, view, hfov, path):
    """Writes a scene for ocr_facade_demo: one container per object, the
    objects' triangles in object order (object ids must be contiguous and
    ascending so the triangle ids are unchanged), textures, lights, view and
    horizontal FOV."""
    import struct
    oids = sc.tris["v"]["object_id"][:, 0]
    if np.any(np.diff(oids.astype(np.int64)) < 0):
        raise ValueError("triangles must be grouped by ascending object id")
    with open(path, "wb") as f:
        f.write(b"OCRSCN01")
        f.write(struct.pack("<I", sc.objs.shape[0]))
        for o in range(sc.objs.shape[0]):
            sel = np.ascontiguousarray(sc.tris[oids == o])
            f.write(struct.pack("<III", 1, sel.shape[0], int(sc.objs["tid"][o])))
            f.write(np.ascontiguousarray(sc.objs[o:o + 1]).tobytes())
            f.write(sel.tobytes())
        f.write(struct.pack("<I", len(sc.textures)))
        for t in sc.textures:
            f.write(struct.pack("<II", t.width, t.height))
            f.write(np.ascontiguousarray(t.rgba, dtype=np.uint8).tobytes())
        f.write(struct.pack("<I", sc.lights.shape[0]))
        f.write(np.ascontiguousarray(sc.lights).tobytes())
        f.write(np.ascontiguousarray(view).tobytes())
        f.write(struct.pack("<f", hfov))
