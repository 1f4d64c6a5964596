"""Pins of the oracle's blend_screens_with_depth (cl2.cl:7842-7875,
alpha_blend cl2.cl:1721-1728): closed forms and bounds fixed by the
definition, independent of the oracle's own arithmetic."""
import numpy as np

RNG = np.random.default_rng(77)


def _blend(oracle, sd, dd, sr, dr):
    return oracle.blend_with_depth(sd, dd, sr, dr)


def test_opaque_source_replaces_and_transparent_keeps(oracle):
    n = 256
    c = np.arange(256, dtype=np.uint8)
    sr = np.stack([c, c[::-1], c, np.full(n, 255, np.uint8)], axis=1)
    dr = RNG.integers(0, 256, (n, 4), dtype=np.uint8)
    depth = np.full(n, 1000, np.uint32)
    out = _blend(oracle, depth, depth, sr, dr)          # equal depths blend
    assert np.array_equal(out, sr)                      # alpha 1: exactly the source
    sr0 = sr.copy()
    sr0[:, 3] = 0
    out = _blend(oracle, depth, depth, sr0, dr)
    assert np.array_equal(out[:, :3], dr[:, :3]) and not out[:, 3].any()   # alpha 0: dst colour, alpha 0


def test_depth_rule(oracle):
    n = 1000
    sd = RNG.integers(0, 2**32 - 1, n, dtype=np.uint32)
    dd = sd.copy()
    dd[: n // 3] -= 1                                   # destination nearer: untouched
    dd[n // 3: 2 * n // 3] += 1                         # source nearer: blended
    sr = RNG.integers(0, 256, (n, 4), dtype=np.uint8)
    sr[:, 3] = 255
    dr = RNG.integers(0, 256, (n, 4), dtype=np.uint8)
    out = _blend(oracle, sd, dd, sr, dr)
    keep = dd < sd
    assert np.array_equal(out[keep], dr[keep])
    assert np.array_equal(out[~keep], sr[~keep])


def test_partial_alpha_within_one_unit_of_exact_blend(oracle):
    n = 200000
    sr = RNG.integers(0, 256, (n, 4), dtype=np.uint8)
    dr = RNG.integers(0, 256, (n, 4), dtype=np.uint8)
    depth = np.zeros(n, np.uint32)
    out = _blend(oracle, depth, depth, sr, dr).astype(np.int64)
    a = sr[:, 3:4].astype(np.float64) / 255.0
    exact = np.rint((sr[:, :3] / 255.0 * a + dr[:, :3] / 255.0 * (1 - a)) * 255.0)
    assert np.abs(out[:, :3] - exact).max() <= 1
    assert np.array_equal(out[:, 3], sr[:, 3].astype(np.int64))


# ---- do_pseudo_aa (cl2.cl:6438-6657)
DEPTH = 1_000_000_000                     # 81.5 units
FAR = DEPTH + 5_000_000                   # 407 units further: beyond the 100-unit bound


def _enc(oracle, n):
    import ctypes as C
    a = np.ascontiguousarray(n, np.float32)
    v = oracle.lib.oracle_encode_normal(a.ctypes.data_as(C.POINTER(C.c_float)))
    return np.array([v & 0xFFFF, v >> 16], np.uint16)


def _flat(oracle, h=6, w=7):
    depth = np.full((h, w), DEPTH, np.uint32)
    nrm = np.tile(_enc(oracle, [0.0, 0.0, 1.0]), (h, w, 1))
    rgba = RNG.integers(0, 256, (h, w, 4), dtype=np.uint8)
    return depth, nrm, rgba


def test_pseudo_aa_flat_and_straight_edges_unchanged(oracle):
    depth, nrm, rgba = _flat(oracle)
    assert np.array_equal(oracle.pseudo_aa(depth, nrm, rgba), rgba)
    depth[:, 4:] = FAR                    # a vertical step: one x neighbour, no y neighbour
    assert np.array_equal(oracle.pseudo_aa(depth, nrm, rgba), rgba)
    depth[:, :] = DEPTH
    depth[3:, :] = FAR                    # a horizontal step
    assert np.array_equal(oracle.pseudo_aa(depth, nrm, rgba), rgba)


def _corner_case(oracle, by_normal):
    depth, nrm, rgba = _flat(oracle, 5, 5)
    a, b = np.array([40, 200, 90], np.uint8), np.array([230, 10, 120], np.uint8)
    rgba[..., :3] = a
    other = [(2, 3), (1, 2), (1, 3)]      # (y, x): right, up, up-right of the centre
    for y, x in other:
        rgba[y, x, :3] = b
        if by_normal:
            nrm[y, x] = _enc(oracle, [np.sin(np.radians(40)), 0.0, np.cos(np.radians(40))])
        else:
            depth[y, x] = FAR
    out = oracle.pseudo_aa(depth, nrm, rgba)
    exact = np.rint((0.65 * a / 255.0 + 0.35 * b / 255.0) * 255.0)
    assert np.abs(out[2, 2, :3].astype(int) - exact).max() <= 1
    assert out[2, 2, 3] == 255
    # border pixels are never touched
    assert np.array_equal(out[0], rgba[0]) and np.array_equal(out[:, 0], rgba[:, 0])


def test_pseudo_aa_corner_by_depth(oracle):
    _corner_case(oracle, by_normal=False)


def test_pseudo_aa_corner_by_normal(oracle):
    _corner_case(oracle, by_normal=True)


def test_pseudo_aa_skips_background(oracle):
    depth, nrm, rgba = _flat(oracle, 5, 5)
    depth[2, 3] = depth[1, 2] = depth[1, 3] = FAR
    depth[2, 2] = 0xFFFFFFFF              # background centre: left alone
    assert np.array_equal(oracle.pseudo_aa(depth, nrm, rgba)[2, 2], rgba[2, 2])
