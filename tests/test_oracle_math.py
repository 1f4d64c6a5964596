"""Pins for the oracle's scalar and geometric building blocks.

Each check compares the oracle with something the mathematics fixes
independently of the oracle's own formula: rotation matrices and quaternion
algebra in float64, exact integer edge functions, textbook functions (acos,
log2, the sRGB curves), inverse operations, and published test values.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle.binding import SinCos, camera_sincos

RNG = np.random.default_rng(1234)
f32 = np.float32


def fp(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def rot_matrix(ang):
    """The forward camera matrix written in the reference comment
    (cl2.cl:208-212): rows (c3c2, s3c2, -s2), (c3s1s2 - s3c1, s3s1s2 + c3c1,
    s1c2), (c3c1s2 + s3s1, s3c1s2 - c3s1, c1c2), 1 = x, 2 = y, 3 = z."""
    s1, s2, s3 = (math.sin(a) for a in ang)
    c1, c2, c3 = (math.cos(a) for a in ang)
    return np.array([[c3 * c2, s3 * c2, -s2],
                     [c3 * s1 * s2 - s3 * c1, s3 * s1 * s2 + c3 * c1, s1 * c2],
                     [c3 * c1 * s2 + s3 * s1, s3 * c1 * s2 - c3 * s1, c1 * c2]])


def quat_matrix(q):
    x, y, z, w = np.asarray(q, dtype=np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_rot_matches_rotation_matrix(oracle):
    for _ in range(200):
        ang = RNG.uniform(-math.pi, math.pi, 3).astype(np.float32)
        p = RNG.uniform(-1000, 1000, 3).astype(np.float32)
        c = RNG.uniform(-1000, 1000, 3).astype(np.float32)
        sc = camera_sincos(ang)
        pa, pp = fp(p)
        ca, cp = fp(c)
        oa, op = fp(np.zeros(3))
        oracle.lib.oracle_rot(pp, cp, C.byref(sc), op)
        want = rot_matrix(ang.astype(np.float64)) @ (p.astype(np.float64) - c)
        assert np.allclose(oa, want, rtol=1e-5, atol=2e-3), (oa, want)


def test_back_rot_inverts_rot(oracle):
    for _ in range(200):
        ang = RNG.uniform(-math.pi, math.pi, 3).astype(np.float32)
        p = RNG.uniform(-1000, 1000, 3).astype(np.float32)
        sc = camera_sincos(ang)
        pa, pp = fp(p)
        za, zp = fp(np.zeros(3))
        ra, rp = fp(np.zeros(3))
        oracle.lib.oracle_rot(pp, zp, C.byref(sc), rp)
        ba, bp = fp(np.zeros(3))
        oracle.lib.oracle_back_rot(rp, C.byref(sc), bp)
        assert np.allclose(ba, p, atol=1e-3)
        # back_rot is the transpose (cl2.cl:247-255)
        want = rot_matrix(ang.astype(np.float64)).T @ ra.astype(np.float64)
        assert np.allclose(ba, want, rtol=1e-5, atol=1e-3)


def test_rot_quat_matches_quaternion_algebra(oracle):
    for _ in range(200):
        q = RNG.normal(size=4).astype(np.float32) * RNG.uniform(0.1, 10)
        p = RNG.uniform(-100, 100, 3).astype(np.float32)
        r = oracle.call_vec("oracle_rot_quat", p, q)
        want = quat_matrix(q) @ p.astype(np.float64)
        assert np.allclose(r, want, rtol=1e-5, atol=1e-4)
        assert abs(np.linalg.norm(r) - np.linalg.norm(p)) < 1e-4 * max(1, np.linalg.norm(p))
        back = oracle.call_vec("oracle_back_rot_quat", r, q)
        assert np.allclose(back, p, atol=1e-3)


def _clip(oracle, pts):
    pa, pp = fp(np.asarray(pts, dtype=np.float32).reshape(9))
    oa, op = fp(np.zeros(18))
    n = oracle.lib.oracle_clip(pp, op)
    return n, oa.reshape(2, 3, 3)[:n]


@pytest.mark.parametrize("behind", [0, 1, 2, 3])
def test_near_clip_cases(oracle, behind):
    """generate_new_triangles (cl2.cl:577-663): 0/1/2/3 vertices at z <= 20
    give 1/2/1/0 triangles; new vertices lie on the original edges at z = 20."""
    for _ in range(100):
        z = np.concatenate([RNG.uniform(-500, 19, behind), RNG.uniform(25, 5000, 3 - behind)])
        RNG.shuffle(z)
        pts = np.stack([RNG.uniform(-500, 500, 3), RNG.uniform(-500, 500, 3), z], axis=1).astype(np.float32)
        n, tris = _clip(oracle, pts)
        assert n == {0: 1, 1: 2, 2: 1, 3: 0}[behind]
        for t in tris:
            for v in t:
                assert v[2] >= 20 - 1e-3
                if not any(np.allclose(v, p) for p in pts):
                    assert abs(v[2] - 20) < 1e-2
                    # on an edge between a kept and a clipped vertex
                    ok = False
                    for a in range(3):
                        for b in range(3):
                            if a != b and pts[a][2] > 20 >= pts[b][2]:
                                tpar = (20 - pts[b][2]) / (pts[a][2] - pts[b][2])
                                e = pts[b] + tpar * (pts[a] - pts[b])
                                ok |= np.allclose(v, e, rtol=1e-4, atol=1e-2)
                    assert ok
        if behind == 0:
            assert np.array_equal(tris[0], pts)
    # far plane: z > 350000 also counts as behind
    pts = np.array([[0, 0, 400000], [1, 0, 100], [0, 1, 100]], np.float32)
    assert _clip(oracle, pts)[0] == 2


def test_far_clip_vertices_count_as_behind(oracle):
    pts = np.array([[0, 0, 400000], [1, 0, 400000], [0, 1, 100]], np.float32)
    assert _clip(oracle, pts)[0] == 1


def test_point_in_tri_is_exact_edge_function(oracle):
    for _ in range(300):
        v = RNG.integers(-40, 40, size=(3, 2))
        xs, xp = fp(v[:, 0])
        ys, yp = fp(v[:, 1])
        for _ in range(30):
            p = RNG.integers(-45, 45, size=2)
            got = oracle.lib.oracle_point_in_tri(float(p[0]), float(p[1]), xp, yp)
            (x0, y0), (x1, y1), (x2, y2) = v.tolist()
            A = 0.5 * (-y1 * x2 + y0 * (-x1 + x2) + x0 * (y1 - y2) + x1 * y2)
            sign = -1 if A < 0 else 1
            s = (y0 * x2 - x0 * y2 + (y2 - y0) * p[0] + (x0 - x2) * p[1]) * sign
            t = (x0 * y1 - y0 * x1 + (y0 - y1) * p[0] + (x1 - x0) * p[1]) * sign
            want = s >= 0 and t >= 0 and (s + t) < 2.0001 * A * sign
            assert bool(got) == bool(want)


def _walk_all(oracle, cut, W, H, op_size=500):
    ca, cp = fp(np.asarray(cut, np.float32).reshape(9))
    mm_a, mm_p = fp(np.zeros(4))
    oracle.lib.oracle_calc_min_max(cp, W, H, mm_p)
    area = (mm_a[1] - mm_a[0]) * (mm_a[3] - mm_a[2])
    nfr = int(math.ceil(np.float32(area) / np.float32(op_size)))
    cov = {}
    xs = np.zeros(1024, np.int32)
    ys = np.zeros(1024, np.int32)
    ds = np.zeros(1024, np.uint32)
    for ch in range(nfr):
        n = oracle.lib.oracle_walk(cp, W, H, ch, op_size, xs.ctypes.data, ys.ctypes.data, ds.ctypes.data, 1024)
        for i in range(min(n, 1024)):
            cov[(int(xs[i]), int(ys[i]))] = int(ds[i])
    return mm_a, cov


def test_fragment_walk_covers_bbox_exactly(oracle):
    """The union of a cut triangle's 500-pixel fragments (kernel1 walk,
    cl2.cl:5046-5126) is exactly the set of integer pixels of the half-open
    box [min_x, max_x) x [min_y, max_y) that pass the edge test, and each
    depth is UINT_MAX*z/350000 of the perspective-correct plane through the
    rounded vertices."""
    W, H = 300, 200
    for trial in range(150):
        c = RNG.uniform([5, 5], [W - 5, H - 5], size=(3, 2))
        z = RNG.uniform(30, 3000, size=3)
        cut = np.concatenate([c, z[:, None]], axis=1).astype(np.float32)
        mm, cov = _walk_all(oracle, cut, W, H)
        xr, yr = np.round(cut[:, 0]).astype(int), np.round(cut[:, 1]).astype(int)
        verts = list(zip(xr.tolist(), yr.tolist()))
        want = set()
        for y in range(int(mm[2]), int(mm[3])):
            for x in range(int(mm[0]), int(mm[1])):
                (x0, y0), (x1, y1), (x2, y2) = verts
                A = 0.5 * (-y1 * x2 + y0 * (-x1 + x2) + x0 * (y1 - y2) + x1 * y2)
                sign = -1 if A < 0 else 1
                s = (y0 * x2 - x0 * y2 + (y2 - y0) * x + (x0 - x2) * y) * sign
                t = (x0 * y1 - y0 * x1 + (y0 - y1) * x + (x1 - x0) * y) * sign
                if s >= 0 and t >= 0 and (s + t) < 2.0001 * A * sign:
                    want.add((x, y))
        assert set(cov) == want, trial
        M = np.array([[xr[i], yr[i], 1.0] for i in range(3)])
        if abs(np.linalg.det(M)) > 0.5:
            # plane of 1/(z/350000) through the rounded vertices, float64
            f = 350000.0 / z.astype(np.float32).astype(np.float64)
            coef = np.linalg.solve(M, f)
            for (x, y), d in list(cov.items())[:50]:
                fd = coef[0] * x + coef[1] * y + coef[2]
                # float32 evaluation of A, B, C and of the plane: the error
                # bound scales with the magnitudes that cancel
                mag = (abs(coef[0] * x) + abs(coef[1] * y) + abs(coef[2])
                       + abs(coef[0] * xr[0]) + abs(coef[1] * yr[0]) + abs(f).max())
                rel = 64 * 2.0 ** -24 * mag / abs(fd)
                assert abs(d - 4294967296.0 / fd) <= rel * 4294967296.0 / fd + 2


def test_calc_min_max(oracle):
    cut = np.array([[10.4, 20.6, 1], [30.5, 5.2, 1], [-3.7, 40.49, 1]], np.float32)
    ca, cp = fp(cut.reshape(9))
    mm, mp = fp(np.zeros(4))
    oracle.lib.oracle_calc_min_max(cp, 25.0, 30.0, mp)
    # round half away from zero, min - 1, clamp to [0, W-1] x [0, H-1]
    assert list(mm) == [0.0, 24.0, 4.0, 29.0]


def test_xorshift_published_value(oracle):
    # Marsaglia, "Xorshift RNGs" (2003): xor32 with (13, 17, 5) from
    # y = 2463534242 returns 723471715.
    assert oracle.lib.oracle_xorshift(2463534242) == 723471715


def test_wang_hash_is_a_bijection_with_avalanche(oracle):
    xs = RNG.choice(2 ** 32, size=1 << 16, replace=False).astype(np.uint64)
    hs = {oracle.lib.oracle_wang_hash(int(x)) for x in xs}
    assert len(hs) == len(xs)
    flips = []
    for x in xs[:2000]:
        b = 1 << int(RNG.integers(0, 32))
        flips.append(bin(oracle.lib.oracle_wang_hash(int(x)) ^ oracle.lib.oracle_wang_hash(int(x) ^ b)).count("1"))
    assert 12 < np.mean(flips) < 20


def test_rational_acos_against_acos(oracle):
    assert oracle.lib.oracle_rational_acos(0.0) == f32(f32(3.1415927) / f32(2))
    for x in np.linspace(-1, 1, 4001):
        assert abs(oracle.lib.oracle_rational_acos(float(x)) - math.acos(x)) < 0.02


def test_log2_approx_against_log2(oracle):
    for x in np.exp(np.linspace(math.log(1e-4), math.log(1e5), 4001)):
        assert abs(oracle.lib.oracle_log2_approx(float(x)) - math.log2(x)) < 0.006


def _g(oracle, fn, x):
    i = (C.c_float * 4)(x, x, x, 0.5)
    o = (C.c_float * 4)()
    getattr(oracle.lib, fn)(i, o)
    assert o[3] == 0.5
    return o[0]


def test_gamma_approximations_follow_srgb(oracle):
    def s2l(c):
        return c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4

    def l2s(c):
        return 12.92 * c if c <= 0.0031308 else 1.055 * c ** (1 / 2.4) - 0.055

    assert _g(oracle, "oracle_gamma_transform", 0.0) == 0.0
    assert abs(_g(oracle, "oracle_gamma_transform", 1.0) - 1.0) < 1e-6
    for x in np.linspace(0, 1, 501):
        assert abs(_g(oracle, "oracle_gamma_transform", float(x)) - s2l(x)) < 0.002
        if x > 0.01:
            assert abs(_g(oracle, "oracle_gamma_inverse", float(x)) - l2s(x)) < 0.01


def test_cubeface_dominant_axis(oracle):
    cases = {(1, 0, 0): 5, (-1, 0, 0): 4, (0, 1, 0): 3, (0, -1, 0): 1, (0, 0, 1): 0, (0, 0, -1): 2}
    la, lp = fp(np.array([10, 20, 30]))
    for d, face in cases.items():
        for scale in (0.5, 7.0, 300.0):
            p = np.array([10, 20, 30]) + np.array(d) * scale + RNG.uniform(-0.1, 0.1, 3)
            pa, pp = fp(p)
            assert oracle.lib.oracle_ret_cubeface(pp, lp) == face


def test_exp_within_one_ulp(oracle):
    """exp in binary32 (Cody-Waite reduction + fitted polynomial): within one
    ulp of the correctly rounded exp over its whole range, exact at 0, 0
    below the subnormal range, inf above FLT_MAX's logarithm, NaN kept."""
    xs = np.concatenate([RNG.uniform(-104, 88.7, 40000), RNG.uniform(-1, 1, 10000),
                         RNG.uniform(-1e-6, 1e-6, 2000)]).astype(np.float32)
    f = oracle.lib.oracle_exp
    got = np.array([f(float(x)) for x in xs], np.float32)
    want = np.exp(xs.astype(np.float64)).astype(np.float32)
    ulps = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    # about 8% of results are one ulp from the correctly rounded exp (the
    # binary32 polynomial's own rounding); subnormal results (x < -87.3) are
    # rounded twice (p, then the scaling)
    normal = xs > -87.3
    assert ulps.max() <= 1 and np.mean(ulps[normal]) < 0.15
    assert f(0.0) == 1.0 and f(-0.0) == 1.0
    assert f(-105.0) == 0.0 and f(-1e30) == 0.0
    assert f(89.0) == np.inf and math.isnan(f(float("nan")))


@pytest.mark.parametrize("fn,ref,lo,hi", [
    ("oracle_tan", np.tan, 0.0, 1.56),
    ("oracle_powr5", lambda x: x ** 5, 0.0, 1.0),
    ("oracle_pow4", lambda x: x ** 4, -1.0, 1.0),
])
def test_binary64_functions_are_correctly_rounded(oracle, fn, ref, lo, hi):
    xs = RNG.uniform(lo, hi, 20000).astype(np.float32)
    f = getattr(oracle.lib, fn)
    got = np.array([f(float(x)) for x in xs], np.float32)
    want = ref(xs.astype(np.float64)).astype(np.float32)
    ulps = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1 and np.mean(ulps) < 1e-3


def test_barycentric_closed_forms(oracle):
    for _ in range(100):
        a, b, c = (RNG.uniform(-100, 100, 3).astype(np.float32) for _ in range(3))
        for p, want in ((a, (1, 0, 0)), (b, (0, 1, 0)), (c, (0, 0, 1)),
                        ((a + b + c) / 3, (1 / 3, 1 / 3, 1 / 3))):
            u, v, w = C.c_float(), C.c_float(), C.c_float()
            arrs = [fp(x) for x in (p, a, b, c)]
            oracle.lib.oracle_get_barycentric(*[x[1] for x in arrs], C.byref(u), C.byref(v), C.byref(w))
            assert np.allclose([u.value, v.value, w.value], want, atol=2e-3)


def test_encode_normal_round_trips_through_decode(oracle):
    """decode_normal (cl2.cl:5632-5646) inverts encode_normal.  Near the
    poles the 16-bit code of z = (vx^2 + vy^2) * 2 - 1 loses x and y (their
    error grows like 1/hypot(x, y)), so normals within 0.1 of the z axis
    are left out; the draw has its own seed, so the cases do not depend on
    which tests ran before in the process."""
    rng = np.random.default_rng(5632)
    for _ in range(500):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        if math.hypot(n[0], n[1]) < 0.1:
            continue
        na, npp = fp(n)
        e = oracle.lib.oracle_encode_normal(npp)
        sx, sy = e & 0xFFFF, e >> 16
        vx = (sx / 65535.0) * 2 - 1
        vy = (sy / 65535.0) * 2 - 1
        z = (vx * vx + vy * vy) * 2 - 1
        l = math.hypot(vx, vy)
        r = math.sqrt(max(1 - z * z, 0))
        dec = np.array([vx / l * r, vy / l * r, z])
        assert np.allclose(dec, n, atol=2e-3)


def test_texture_mod(oracle):
    # texture_mod (cl2.cl:1457-1468): [0,1) unchanged, >= 1 mirrored, < 0 wrapped
    for x, want in ((0.25, 0.25), (1.25, 0.75), (2.5, 0.5), (-0.25, 0.25), (-1.75, 0.75)):
        o = oracle.call_vec("oracle_texture_mod", np.array([x, x]), out_len=2)
        assert abs(o[0] - want) < 1e-6
