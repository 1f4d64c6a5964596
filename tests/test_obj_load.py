"""Scene ingest: the OBJ + MTL loader (include/ocr_obj.h, obj_load.cpp:181-568)
against the oracle's restatement, on files written here that exercise each of
the reference's parse rules, and on the reference's own OBJ assets when its
tree is present; the PNG decoder against images whose pixels are known."""
import ctypes as C
import os
import struct
import zlib

import numpy as np
import pytest

from openclrenderer_amd import abi
from oracle import obj_load as ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"


class Part(C.Structure):
    _fields_ = [("first_tri", C.c_uint32), ("ntri", C.c_uint32), ("texture_kind", C.c_int32),
                ("colour", C.c_uint8 * 4), ("texture_cache", C.c_char * 512), ("texture_path", C.c_char * 512),
                ("bump_path", C.c_char * 512), ("group_name", C.c_char * 512)]


@pytest.fixture(scope="module")
def eng():
    from openclrenderer_amd import build as b
    if not os.path.exists(b.FACADE_LIB):
        b.build()
    lib = C.CDLL(b.FACADE_LIB)
    vp = C.c_void_p
    lib.ocr_obj_load.argtypes = [C.c_char_p, C.c_float, C.POINTER(vp)]
    lib.ocr_obj_ntri.argtypes = [vp]
    lib.ocr_obj_ntri.restype = C.c_uint32
    lib.ocr_obj_nparts.argtypes = [vp]
    lib.ocr_obj_nparts.restype = C.c_uint32
    lib.ocr_obj_get.argtypes = [vp, vp, vp]
    lib.ocr_obj_free.argtypes = [vp]
    lib.ocr_png_load.argtypes = [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 C.POINTER(C.POINTER(C.c_uint8))]
    lib.ocr_png_free.argtypes = [C.POINTER(C.c_uint8)]
    return lib


def load_native(lib, path, scale=1.0):
    m = C.c_void_p()
    rc = lib.ocr_obj_load(path.encode(), scale, C.byref(m))
    if rc != 0:
        return rc, None, None
    n, k = lib.ocr_obj_ntri(m), lib.ocr_obj_nparts(m)
    tris = np.zeros(n, dtype=abi.TRIANGLE)
    parts = (Part * max(k, 1))()
    lib.ocr_obj_get(m, tris.ctypes.data, C.addressof(parts))
    lib.ocr_obj_free(m)
    return 0, tris, [parts[i] for i in range(k)]


def assert_same(lib, path, scale=1.0):
    rc, tris, parts = load_native(lib, path, scale)
    assert rc == 0
    want, wparts = ref.load(path, scale)
    v = tris["v"]
    assert np.array_equal(v["pos"].view(np.uint32), want["pos"].view(np.uint32))
    assert np.array_equal(v["normal"].view(np.uint32), want["normal"].view(np.uint32))
    assert np.array_equal(v["vt"].view(np.uint32), want["vt"].view(np.uint32))
    assert not v["object_id"].any() and not v["vertex_col"].any()
    assert len(parts) == len(wparts)
    for p, w in zip(parts, wparts):
        assert (p.first_tri, p.ntri, p.texture_kind) == (w["first_tri"], w["ntri"], w["texture_kind"])
        assert tuple(p.colour) == tuple(w["colour"])
        for key in ("texture_cache", "texture_path", "bump_path", "group_name"):
            assert getattr(p, key).decode("latin-1") == w[key][:511], key
    return tris, parts


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode("latin-1"))
    return str(p)


CUBEISH = """# header
mtllib t.mtl
v -1.000000 -1.000000 1.000000
v  1.5 2.25  -3.125
v 1e3 2.0 -0.5
v 4 5 6
vt 0.25 0.75
vt 1 0
vn 0.0 1.0 0.0
vn 0 0 1
f 1/1/1 2/2/2 3/1/1
usemtl Red
g red_group
f 1/1/1 2/2/2 3/1/2
f 2/2/2 3/1/1 4/2/2
usemtl Bumpy
g second
f 4/2/2 3/1/1 1/1/1
usemtl Missing
g third
f 1/1/1 4/2/2 2/2/1
g fourth
usemtl Tex\r
f 3/1/1 2/2/2 1/1/1
"""

CUBEISH_MTL = """newmtl Tex\r
map_Kd sub/dir/tex.png
newmtl RedDouble
Kd 0.5 0.25 1.0
newmtl Red
Kd 0.8 0.2 0.2
map_Bump red_bump.png
newmtl Bumpy
Kd 0.1 0.2 0.3
map_Bump bump.png
newmtl Missing
"""


def test_parse_rules_match_oracle(eng, tmp_path):
    obj = write(tmp_path, "t.obj", CUBEISH)
    write(tmp_path, "t.mtl", CUBEISH_MTL)
    (tmp_path / "tex.png").write_bytes(_png(np.zeros((2, 2, 4), np.uint8) + 7, 6))
    tris, parts = assert_same(eng, obj, scale=2.5)
    # the face before the first usemtl belongs to no part
    assert len(tris) == 6 and parts[0].first_tri == 1 and parts[0].ntri == 2
    # "Red" matches "newmtl RedDouble" first (name prefix): RedDouble's Kd;
    # the bump search goes on through "newmtl Red", which matches too
    assert parts[0].texture_kind == 1 and tuple(parts[0].colour) == (127, 63, 255, 255)
    assert parts[0].bump_path.decode().endswith("/red_bump.png")
    # a block with Kd and map_Bump: Kd * 255 truncated, the bump beside the file
    assert tuple(parts[1].colour) == (25, 51, 76, 255) and parts[1].bump_path.decode().endswith("/bump.png")
    assert parts[1].group_name == b"second"
    assert parts[2].texture_kind == 2 and tuple(parts[2].colour) == (255, 0, 255, 255)
    # "Tex\r": sub/dir/tex.png is absent, its base name beside the .obj is used
    assert parts[3].texture_kind == 0 and parts[3].texture_path.decode().endswith("/tex.png")
    assert parts[3].texture_cache.decode().endswith("/sub/dir/tex.png")


def test_double_spaces_and_atoi_lines(eng, tmp_path):
    obj = write(tmp_path, "a.obj", "v  1 2 3\nv 4 5 6\nv 7e1 8 9\nvt 1 1\nvn 0 0 1\nusemtl m\n"
                                   "f 1/1/1 2/1/1 3/1/1\n")
    tris, _ = assert_same(eng, obj)
    # no '.' on the line: atoi, and the double space repeats the first field
    assert tris["v"]["pos"][0, 0, :3].tolist() == [1.0, 1.0, 2.0]
    assert tris["v"]["pos"][0, 2, :3].tolist() == [7.0, 8.0, 9.0]


def test_no_usemtl_gives_no_parts_and_bad_indices_fail(eng, tmp_path):
    obj = write(tmp_path, "b.obj", "v 0.0 0.0 0.0\nvt 0.0 0.0\nvn 0.0 0.0 1.0\nf 1/1/1 1/1/1 1/1/1\n")
    _, tris, parts = load_native(eng, obj)
    assert len(tris) == 1 and parts == []
    bad = write(tmp_path, "c.obj", "v 0.0 0.0 0.0\nvt 0.0 0.0\nvn 0.0 0.0 1.0\nusemtl x\nf 1/1/1 2/1/1 1/1/1\n")
    assert load_native(eng, bad)[0] != 0
    assert load_native(eng, str(tmp_path / "absent.obj"))[0] != 0


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference tree not present")
@pytest.mark.parametrize("rel", ["objects/cube.obj", "objects/square.obj", "objects/high_cylinder_forward.obj",
                                 "objects/high_cylinder_forward_low.obj", "Game/Res/tex_cube.obj",
                                 "Game/Res/tex_cube_2.obj", "Game/Res/tex_cube_3.obj"])
def test_reference_assets_match_oracle(eng, rel):
    tris, parts = assert_same(eng, os.path.join(REFERENCE, rel), scale=100.0)
    assert len(tris) > 0 and len(parts) > 0


def _png(rgba, ctype, filters=None, palette=None, trns=None):
    h, w, _ = rgba.shape
    if ctype == 0:
        px = rgba[..., :1]
    elif ctype == 4:
        px = np.concatenate([rgba[..., :1], rgba[..., 3:4]], axis=2)
    elif ctype == 2:
        px = rgba[..., :3]
    elif ctype == 6:
        px = rgba
    else:
        px = palette
    ch = px.shape[2]
    rows = px.reshape(h, w * ch).astype(np.int64)
    raw = bytearray()
    for y in range(h):
        ft = 0 if filters is None else filters[y % len(filters)]
        cur = rows[y]
        up = rows[y - 1] if y else np.zeros_like(cur)
        left = np.concatenate([np.zeros(ch, np.int64), cur[:-ch]])
        ul = np.concatenate([np.zeros(ch, np.int64), up[:-ch]])
        if ft == 0:
            pred = np.zeros_like(cur)
        elif ft == 1:
            pred = left
        elif ft == 2:
            pred = up
        elif ft == 3:
            pred = (left + up) >> 1
        else:
            p = left + up - ul
            pa, pb, pc = np.abs(p - left), np.abs(p - up), np.abs(p - ul)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, up, ul))
        raw.append(ft)
        raw += ((cur - pred) & 0xFF).astype(np.uint8).tobytes()

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)

    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0))
    if ctype == 3:
        out += chunk(b"PLTE", trns[0].tobytes())
        if trns[1] is not None:
            out += chunk(b"tRNS", trns[1].tobytes())
    return out + chunk(b"IDAT", zlib.compress(bytes(raw))) + chunk(b"IEND", b"")


def _decode(lib, path):
    w, h, px = C.c_uint32(), C.c_uint32(), C.POINTER(C.c_uint8)()
    assert lib.ocr_png_load(path.encode(), C.byref(w), C.byref(h), C.byref(px)) == 0
    out = np.ctypeslib.as_array(px, shape=(h.value, w.value, 4)).copy()
    lib.ocr_png_free(px)
    return out


@pytest.mark.parametrize("ctype", [0, 2, 3, 4, 6])
def test_png_decodes_known_pixels(eng, tmp_path, ctype):
    rng = np.random.default_rng(ctype)
    h, w = 37, 53
    rgba = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    palette = trns = None
    if ctype == 0:
        rgba[..., 1] = rgba[..., 2] = rgba[..., 0]
        rgba[..., 3] = 255
    elif ctype == 4:
        rgba[..., 1] = rgba[..., 2] = rgba[..., 0]
    elif ctype == 2:
        rgba[..., 3] = 255
    elif ctype == 3:
        pal = rng.integers(0, 256, (40, 3), dtype=np.uint8)
        alpha = rng.integers(0, 256, 25, dtype=np.uint8)     # entries past 25 are opaque
        idx = rng.integers(0, 40, (h, w), dtype=np.uint8)
        rgba[..., :3] = pal[idx]
        rgba[..., 3] = np.where(idx < 25, alpha[np.minimum(idx, 24)], 255)
        palette, trns = idx[..., None], (pal, alpha)
    path = str(tmp_path / f"p{ctype}.png")
    open(path, "wb").write(_png(rgba, ctype, filters=[0, 1, 2, 3, 4], palette=palette, trns=trns))
    got = _decode(eng, path)
    assert np.array_equal(got, rgba)
    assert np.array_equal(ref.png_decode(path), rgba)


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference tree not present")
@pytest.mark.parametrize("rel", ["objects/red.png", "objects/square_mtl.png"])
def test_reference_textures_decode_like_oracle(eng, rel):
    path = os.path.join(REFERENCE, rel)
    assert np.array_equal(_decode(eng, path), ref.png_decode(path))


def test_committed_model_fixture(eng):
    """tests/golden/obj_model (make_obj_fixture.py): the native loader equals
    the oracle's reading, the four parts take a PNG map, a Kd colour, the
    PNG again (same cache key) and magenta, and both PNG decoders give the
    generator's pixels although every PNG filter type occurs."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_obj_fixture import checker_rgba
    model = os.path.join(ROOT, "tests", "golden", "obj_model", "model.obj")
    tris, parts = assert_same(eng, model, 100.0)
    assert tris.shape[0] == 21
    kinds = [(p.group_name.decode(), p.texture_kind, tuple(p.colour)) for p in parts]
    assert kinds == [("crate", 0, (0, 0, 0, 0)), ("roof", 1, (204, 76, 25, 255)), ("floor", 0, (0, 0, 0, 0)),
                     ("ghost", 2, (255, 0, 255, 255))]
    assert parts[0].texture_cache == parts[2].texture_cache
    png = os.path.join(ROOT, "tests", "golden", "obj_model", "checker.png")
    assert np.array_equal(ref.png_decode(png), checker_rgba())
    w, h, px = C.c_uint32(), C.c_uint32(), C.POINTER(C.c_uint8)()
    assert eng.ocr_png_load(png.encode(), C.byref(w), C.byref(h), C.byref(px)) == 0
    got = np.ctypeslib.as_array(px, shape=(h.value, w.value, 4)).copy()
    eng.ocr_png_free(px)
    assert np.array_equal(got, checker_rgba())
