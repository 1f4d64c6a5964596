"""Writes the OBJ + MTL + PNG model under tests/golden/obj_model/ (a
project-written fixture for scene ingest, obj_load.cpp:181-568):

* group `crate`, material `crate` (map_Kd checker.png): a cube of edge 2 at
  x = -1.6 with per-face normals and 0..1 UVs;
* group `roof`, material `paint` (Kd only): a square pyramid at x = +1.6;
* group `floor`, material `crate` again (the cached texture): a quad at
  y = +1.2 (the camera's down) spanning x in [-4, 4], z in [-3, 3];
* group `ghost`, material `ghost` (map_Kd of a missing file, no Kd): one
  triangle, magenta.

checker.png is 64x64 RGBA8: 8x8 checks of two colours with a horizontal
ramp in blue and a vertical ramp in alpha, its rows filtered with PNG filter
types 0-4 in turn (row y uses type y % 5).

Run: python tests/golden/make_obj_fixture.py
"""
import os
import struct
import zlib

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "obj_model")


def checker_rgba():
    y, x = np.mgrid[0:64, 0:64]
    check = ((x // 8 + y // 8) % 2).astype(bool)
    img = np.zeros((64, 64, 4), np.uint8)
    img[..., 0] = np.where(check, 230, 40)
    img[..., 1] = np.where(check, 200, 90)
    img[..., 2] = (x * 4).astype(np.uint8)
    img[..., 3] = (255 - y).astype(np.uint8)
    return img


def png_bytes(img):
    h, w, _ = img.shape
    raw = img.reshape(h, w * 4).astype(np.int64)
    out = bytearray()
    for yy in range(h):
        ft = yy % 5
        row, up = raw[yy], raw[yy - 1] if yy else np.zeros(w * 4, np.int64)
        out.append(ft)
        for i in range(w * 4):
            a = row[i - 4] if i >= 4 else 0
            b = up[i]
            c = up[i - 4] if i >= 4 else 0
            if ft == 0:
                pred = 0
            elif ft == 1:
                pred = a
            elif ft == 2:
                pred = b
            elif ft == 3:
                pred = (a + b) >> 1
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            out.append(int(row[i] - pred) & 0xFF)

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(bytes(out), 9)) + chunk(b"IEND", b""))


class ObjWriter:
    def __init__(self):
        self.v, self.vt, self.vn, self.lines = [], [], [], []

    def _idx(self, lst, item):
        lst.append(item)
        return len(lst)

    def tri(self, p, uv, n):
        ni = self._idx(self.vn, n)
        f = []
        for k in range(3):
            vi = self._idx(self.v, p[k])
            ti = self._idx(self.vt, uv[k])
            f.append(f"{vi}/{ti}/{ni}")
        self.lines.append("f " + " ".join(f))

    def quad(self, a, b, c, d, n):
        self.tri((a, b, c), ((0, 0), (1, 0), (1, 1)), n)
        self.tri((a, c, d), ((0, 0), (1, 1), (0, 1)), n)

    def text(self):
        s = ["# project-written fixture: make_obj_fixture.py", "mtllib model.mtl"]
        s += ["v %.4f %.4f %.4f" % tuple(p) for p in self.v]
        s += ["vt %.4f %.4f" % tuple(t) for t in self.vt]
        s += ["vn %.4f %.4f %.4f" % tuple(n) for n in self.vn]
        return "\n".join(s + self.lines) + "\n"


def cube(w, cx):
    h = 1.0
    c = np.array([cx, 0.0, 0.0])
    corners = {k: c + np.array(k) * h for k in
               [(sx, sy, sz) for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]}
    faces = [((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (0, 0, -1)),
             ((1, -1, 1), (-1, -1, 1), (-1, 1, 1), (1, 1, 1), (0, 0, 1)),
             ((-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1), (-1, 0, 0)),
             ((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1), (1, 0, 0)),
             ((-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, -1, -1), (0, -1, 0)),
             ((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1), (0, 1, 0))]
    for a, b, cc, d, n in faces:
        w.quad(corners[a], corners[b], corners[cc], corners[d], n)


def pyramid(w, cx):
    apex = np.array([cx, -1.2, 0.0])
    base = [np.array([cx + sx, 1.0, sz]) for sx, sz in ((-1, -1), (1, -1), (1, 1), (-1, 1))]
    for i in range(4):
        a, b = base[i], base[(i + 1) % 4]
        n = np.cross(b - a, apex - a)
        n = n / np.linalg.norm(n)
        w.tri((a, b, apex), ((0, 1), (1, 1), (0.5, 0)), n)
    w.quad(base[3], base[2], base[1], base[0], (0, 1, 0))


def main():
    os.makedirs(HERE, exist_ok=True)
    w = ObjWriter()
    w.lines += ["g crate", "usemtl crate"]
    cube(w, -1.6)
    w.lines += ["g roof", "usemtl paint"]
    pyramid(w, 1.6)
    w.lines += ["g floor", "usemtl crate"]
    w.quad((-4, 1.2, -3), (4, 1.2, -3), (4, 1.2, 3), (-4, 1.2, 3), (0, -1, 0))
    w.lines += ["g ghost", "usemtl ghost"]
    w.tri(((-0.5, -1.8, -1.5), (0.5, -1.8, -1.5), (0.0, -2.6, -1.5)), ((0, 0), (1, 0), (0.5, 1)), (0, 0, -1))
    with open(os.path.join(HERE, "model.obj"), "w") as f:
        f.write(w.text())
    with open(os.path.join(HERE, "model.mtl"), "w") as f:
        f.write("# project-written fixture: make_obj_fixture.py\n"
                "newmtl crate\nKd 0.5 0.5 0.5\nmap_Kd checker.png\n\n"
                "newmtl paint\nKd 0.8 0.3 0.1\n\n"
                "newmtl ghost\nmap_Kd nothere.png\n")
    with open(os.path.join(HERE, "checker.png"), "wb") as f:
        f.write(png_bytes(checker_rgba()))


if __name__ == "__main__":
    main()
