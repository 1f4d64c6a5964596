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
