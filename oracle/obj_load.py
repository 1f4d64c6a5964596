"""Restatement of the reference's OBJ + MTL loader (obj_load, obj_load.cpp:
181-568) and of PNG decoding, for the tests of the loader in
openclrenderer_amd/csrc/obj_load.cpp.  Test infrastructure: nothing in the
product imports it.

Every rule of the reference is kept, quirks included (each cited below):
C atoi/atof prefix parsing, fields read from the character after the
previous space with int positions (a missing space restarts at the line's
first character), usemtl parts, prefix matching of newmtl names, map_Kd
lookup running past the next material, Kd colours truncated to bytes.
atof is C strtod on decimal input; hexadecimal and inf/nan spellings are not
restated (no OBJ writer emits them).
"""
import os
import re
import struct
import zlib

import numpy as np

_FLOAT = re.compile(r"[ \t\n\v\f\r]*([+-]?(?:\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?))")
_INT = re.compile(r"[ \t\n\v\f\r]*([+-]?\d+)")


def atof(s):
    m = _FLOAT.match(s)
    return float(m.group(1)) if m else 0.0


def atoi(s):
    m = _INT.match(s)
    return int(m.group(1)) if m else 0


def _find(s, ch, start):
    """std::string::find into an int: -1 when absent; a start of -1 + 1
    searches from 0."""
    return s.find(ch, max(start, 0))


def decompose_attribute(s, n):
    """obj_load.cpp:144-165."""
    has_dot = "." in s
    prev = _find(s, " ", 0)
    out = []
    for _ in range(n):
        nxt = _find(s, " ", prev + 1)
        tail = s[prev + 1:]
        out.append(np.float32(atof(tail)) if has_dot else np.float32(atoi(tail)))
        prev = nxt
    return out


def decompose_face(s):
    """obj_load.cpp:168-186: 0-based v, vt, vn of three corners."""
    start = 2
    v, vt, vn = [], [], []
    for _ in range(3):
        s1 = _find(s, "/", start)
        s2 = _find(s, "/", s1 + 1)
        s3 = _find(s, " ", s2 + 1)
        v.append(atoi(s[start:]) - 1)
        vt.append(atoi(s[s1 + 1:]) - 1)
        vn.append(atoi(s[s2 + 1:]) - 1)
        start = s3 + 1
    return v, vt, vn


def read_lines(path):
    """getline while good(): a trailing newline gives a last empty line."""
    with open(path, "rb") as f:
        data = f.read().decode("latin-1")
    lines = data.split("\n")
    return lines


def _after_last_space(s):
    return s[s.rfind(" ") + 1:]


def is_samemtl(line, name):
    """obj_load.cpp:26-29."""
    return line[:7] == "newmtl " and line[line.rfind(" ") + 1:][:len(name)] == name


def retrieve_diffuse(mtl, name):
    """obj_load.cpp:32-51."""
    found = False
    for line in mtl:
        if is_samemtl(line, name):
            found = True
        if found and line[:7] == "map_Kd ":
            return _after_last_space(line)
    return ""


def retrieve_bumpmap(mtl, name):
    """obj_load.cpp:54-79."""
    found = False
    for line in mtl:
        if is_samemtl(line, name):
            found = True
            continue
        if found and line[:9] == "map_Bump ":
            return _after_last_space(line)
        if found and line[:7] == "newmtl ":
            return "None"
    return "None"


def get_kd(mtl, name):
    """obj_load.cpp:83-119; None when absent."""
    found = False
    for line in mtl:
        if is_samemtl(line, name):
            found = True
            continue
        if not found or line[:2] != "Kd":
            continue
        parts = line.split(" ")
        if parts and parts[-1] == "":
            parts = parts[:-1]          # getline on ' ' drops a trailing empty field
        if len(parts) <= 3:
            return None
        return [np.float32(atof(p)) for p in parts[1:4]]
    return None


def _to_u8(v):
    v = float(v)
    if not v > 0:
        return 0
    return 255 if v >= 255 else int(v)


def _cpp_to_string(f):
    return "%f" % float(f)


def load(path, scale=1.0):
    """(triangles, parts): triangles as an (n, 3, 3) array per field dict,
    parts as dicts mirroring ocr_obj_part."""
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    lines = read_lines(path)
    dot = path.rfind(".")
    mtlname = (path[:dot] if dot >= 0 else path) + ".mtl"
    slash = path.rfind("/")
    d = "." if slash < 0 else path[:slash]
    mtl = read_lines(mtlname) if os.path.exists(mtlname) else []
    vl, vtl, vnl, fl = [], [], [], []
    use_pos, use_name, group_pos, group_name = [], [], [], []
    for line in lines:
        c = (line + "\0\0\0")[:3]
        if c[0] == "f" and c[1] == " ":
            fl.append(decompose_face(line))
        elif c[0] == "v" and c[1] == " ":
            vl.append(decompose_attribute(line, 3))
        elif c[0] == "v" and c[1] == "t" and c[2] == " ":
            vtl.append(decompose_attribute(line, 2))
        elif c[0] == "v" and c[1] == "n" and c[2] == " ":
            vnl.append(decompose_attribute(line, 3))
        elif c == "use":
            use_pos.append(len(fl))
            use_name.append(_after_last_space(line))
        elif line[:2] == "g ":
            group_pos.append(len(fl))
            group_name.append(_after_last_space(line))
    n = len(fl)
    pos = np.zeros((n, 3, 4), np.float32)
    nrm = np.zeros((n, 3, 4), np.float32)
    vt = np.zeros((n, 3, 2), np.float32)
    s32 = np.float32(scale)
    for i, (v, t, nn) in enumerate(fl):
        for j in range(3):
            for idx, lst in ((v[j], vl), (t[j], vtl), (nn[j], vnl)):
                if not 0 <= idx < len(lst):
                    raise ValueError(f"face {i}: index out of range")
            pos[i, j, :3] = np.array(vl[v[j]], np.float32) * s32
            pos[i, j, 3] = np.float32(0) * s32
            vt[i, j] = vtl[t[j]]
            nrm[i, j, :3] = vnl[nn[j]]
    using_groups = len(use_pos) == len(group_pos)
    use_pos.append(n)
    parts = []
    for i in range(len(use_pos) - 1):
        tname = retrieve_diffuse(mtl, use_name[i])
        bname = retrieve_bumpmap(mtl, use_name[i])
        full = d + "/" + tname
        cache = full
        exists = os.path.isfile(full) or os.path.isdir(full)
        if not exists and tname != "":
            base = tname[max(tname.rfind("/"), tname.rfind("\\")) + 1:]
            full = d + "/" + base
            exists = os.path.isfile(full) or os.path.isdir(full)
        p = {"first_tri": use_pos[i], "ntri": use_pos[i + 1] - use_pos[i], "texture_cache": "",
             "texture_path": "", "bump_path": "", "group_name": group_name[i] if using_groups else ""}
        if tname != "" and exists:
            p.update(texture_kind=0, colour=(0, 0, 0, 0), texture_cache=cache, texture_path=full)
        else:
            kd = get_kd(mtl, use_name[i])
            if kd is not None:
                r, g, b = (np.float32(k) * np.float32(255.0) for k in kd)
                p.update(texture_kind=1, colour=(_to_u8(r), _to_u8(g), _to_u8(b), 255),
                         texture_cache="CACHE" + "".join(_cpp_to_string(x) for x in (r, g, b, 255.0)))
            else:
                p.update(texture_kind=2, colour=(255, 0, 255, 255))
        if bname != "None":
            p["bump_path"] = d + "/" + bname
        parts.append(p)
    return {"pos": pos, "normal": nrm, "vt": vt}, parts


def png_decode(path):
    """8-bit non-interlaced PNG -> (h, w, 4) uint8, top row first (PNG
    specification: chunk layout, zlib stream, filter types 0-4)."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    p, idat, plte, trns = 8, b"", b"", b""
    while p + 12 <= len(data):
        n, = struct.unpack(">I", data[p:p + 4])
        kind, body = data[p + 4:p + 8], data[p + 8:p + 8 + n]
        if kind == b"IHDR":
            w, h, depth, ctype, _, _, lace = struct.unpack(">IIBBBBB", body)
        elif kind == b"PLTE":
            plte = body
        elif kind == b"tRNS":
            trns = body
        elif kind == b"IDAT":
            idat += body
        elif kind == b"IEND":
            break
        p += 12 + n
    assert depth == 8 and lace == 0
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    raw = zlib.decompress(idat)
    stride = w * ch
    out = np.zeros((h, stride), np.int64)
    for y in range(h):
        ft = raw[y * (stride + 1)]
        src = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int64)
        up = out[y - 1] if y else np.zeros(stride, np.int64)
        row = out[y]
        for i in range(stride):
            a = row[i - ch] if i >= ch else 0
            b = up[i]
            c = up[i - ch] if i >= ch else 0
            if ft == 0:
                pred = 0
            elif ft == 1:
                pred = a
            elif ft == 2:
                pred = b
            elif ft == 3:
                pred = (a + b) >> 1
            else:
                pp = a + b - c
                pa, pb, pc = abs(pp - a), abs(pp - b), abs(pp - c)
                pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            row[i] = (src[i] + pred) & 0xFF
    px = out.astype(np.uint8).reshape(h, w, ch)
    rgba = np.zeros((h, w, 4), np.uint8)
    if ctype == 0:
        rgba[..., :3] = px[..., :1]
        rgba[..., 3] = 255
    elif ctype == 4:
        rgba[..., :3] = px[..., :1]
        rgba[..., 3] = px[..., 1]
    elif ctype == 2:
        rgba[..., :3] = px
        rgba[..., 3] = 255
    elif ctype == 6:
        rgba[:] = px
    else:
        pal = np.frombuffer(plte, np.uint8).reshape(-1, 3)
        alpha = np.full(len(pal), 255, np.uint8)
        alpha[:len(trns)] = np.frombuffer(trns, np.uint8)[:len(pal)]
        rgba[..., :3] = pal[px[..., 0]]
        rgba[..., 3] = alpha[px[..., 0]]
    return rgba
