"""The host-side C and C++ under sanitizers, on the CPU:

* the oracle (oracle/ocr_oracle.c) built with AddressSanitizer and
  UndefinedBehaviorSanitizer (oracle/Makefile `asan`), running the oracle
  test files, and shown to be the build in use by a planted overflow that
  the sanitizer must report;
* the facade (facade.cpp), the OBJ / MTL / PNG readers (obj_load.cpp) and
  the main.cpp-style demo with ASan + UBSan over a host-only stand-in of the
  C ABI (tests/sanitize/abi_stub.cpp): config-1 frames in the sync, async
  and motion-blur flows, the committed OBJ model, and a corpus of malformed
  OBJ, MTL and PNG files;
* the served form's request layer (server.hip: dispatcher and completion
  threads, frame and pinned-buffer pools, HBM admission, batch splits) and
  the facade's asynchronous rebuild under ThreadSanitizer, the GPU calls
  replaced by host-only stand-ins (tests/sanitize/server_stub.hip).

Any sanitizer report fails the test (the first report aborts the run)."""
import os
import random
import struct
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "openclrenderer_amd", "csrc")
SAN_DIR = os.path.join(ROOT, "tests", "sanitize")
INC = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANGXX = "/opt/rocm/llvm/bin/clang++"
ASAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
MARKERS = ("ERROR: AddressSanitizer", "runtime error:", "WARNING: ThreadSanitizer", "ERROR: LeakSanitizer")


def _gcc_runtime(name):
    return subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()


def _run(cmd, env=None, timeout=600):
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run(cmd, capture_output=True, text=True, env=e, timeout=timeout)
    return r.returncode, r.stdout + r.stderr


def _clean(rc, out, what):
    bad = [m for m in MARKERS if m in out]
    assert not bad, f"{what}: sanitizer report\n{out[-6000:]}"
    assert rc == 0, f"{what}: exit {rc}\n{out[-4000:]}"


@pytest.fixture(scope="module")
def bins(tmp_path_factory):
    d = tmp_path_factory.mktemp("san")
    srcs = [os.path.join(CSRC, "facade_demo.cpp"), os.path.join(CSRC, "facade.cpp"), os.path.join(CSRC, "obj_load.cpp"),
            os.path.join(SAN_DIR, "abi_stub.cpp")]
    out = {}
    for tag, flags in (("facade_asan", ASAN), ("facade_tsan", ["-O1", "-g", "-fsanitize=thread"])):
        rc, log = _run(["g++", "-std=c++17", *flags, "-I", INC, *srcs, "-lz", "-pthread", "-o", str(d / tag)])
        assert rc == 0, log
        out[tag] = str(d / tag)
    rc, log = _run(["g++", "-std=c++17", *ASAN, "-I", INC, os.path.join(SAN_DIR, "ingest_main.cpp"),
                    os.path.join(CSRC, "obj_load.cpp"), "-lz", "-o", str(d / "ingest_asan")])
    assert rc == 0, log
    out["ingest_asan"] = str(d / "ingest_asan")
    # server.hip's host code with ThreadSanitizer (host-only compile: no
    # device code, the sanitizer on the host side only)
    objs = []
    for src in (os.path.join(CSRC, "server.hip"), os.path.join(SAN_DIR, "server_stub.hip"),
                os.path.join(SAN_DIR, "server_tsan_main.hip")):
        o = str(d / (os.path.basename(src) + ".o"))
        rc, log = _run([HIPCC, "-x", "hip", "--offload-host-only", "-std=c++17", "-O1", "-g", "-fPIC",
                        "-Xarch_host", "-fsanitize=thread", "-c", src, "-o", o])
        assert rc == 0, log
        objs.append(o)
    rc, log = _run([CLANGXX, "-fsanitize=thread", *objs, "-pthread", "-o", str(d / "server_tsan")])
    assert rc == 0, log
    out["server_tsan"] = str(d / "server_tsan")
    out["dir"] = d
    return out


@pytest.fixture(scope="module")
def asan_oracle():
    rc, log = _run(["make", "-C", os.path.join(ROOT, "oracle"), "asan"])
    assert rc == 0, log
    return os.path.join(ROOT, "oracle", "_build", "libocr_oracle_asan.so")


def _asan_env(lib):
    return {"OCR_ORACLE_LIB": lib, "LD_PRELOAD": ":".join(filter(None, [_gcc_runtime("libasan.so"),
                                                                           os.environ.get("LD_PRELOAD", "")])),
            "ASAN_OPTIONS": "detect_leaks=0", "UBSAN_OPTIONS": "print_stacktrace=1"}


def test_oracle_under_asan_ubsan(asan_oracle):
    files = [os.path.join(ROOT, "tests", f) for f in ("test_oracle_math.py", "test_oracle_pins.py",
                                                     "test_oracle_frames.py", "test_oracle_post.py",
                                                     "test_oracle_motion_blur.py")]
    rc, out = _run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *files],
                   env=_asan_env(asan_oracle), timeout=900)
    _clean(rc, out, "oracle tests")


def test_asan_oracle_reports_a_planted_overflow(asan_oracle):
    """An output array 16 elements short of what the call writes: the
    sanitizer build must stop at the first write past it."""
    code = ("import ctypes as C, numpy as np\n"
            "from oracle.binding import Oracle\n"
            "o = Oracle()\n"
            "o.lib.oracle_exp_n.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]\n"
            "x = np.zeros(64, np.float32); y = np.zeros(48, np.float32)\n"
            "o.lib.oracle_exp_n(x.ctypes.data, y.ctypes.data, 64)\n")
    env = _asan_env(asan_oracle)
    env["PYTHONPATH"] = ROOT
    rc, out = _run([sys.executable, "-c", code], env=env, timeout=120)
    assert rc != 0 and "heap-buffer-overflow" in out, out[-3000:]


def _scene_files(d):
    sys.path.insert(0, ROOT)
    from openclrenderer_amd import scenes
    sc, view = scenes.config1(shadow=1)
    s1 = str(d / "cfg1.bin")
    scenes.dump_for_facade(sc, view, 60.0, s1)
    lights = scenes.make_light((250.0, -400.0, -350.0), shadow=1)
    v2 = scenes.make_view(320, 240, (0.0, -120.0, -700.0))
    s2 = str(d / "lights.bin")
    with open(s2, "wb") as f:
        f.write(b"OCRSCN01" + struct.pack("<II", 0, 0) + struct.pack("<I", 1) + lights.tobytes() + v2.tobytes() +
                struct.pack("<f", 60.0))
    return s1, s2


def test_facade_under_asan_ubsan(bins):
    d = bins["dir"]
    s1, s2 = _scene_files(d)
    model = os.path.join(ROOT, "tests", "golden", "obj_model", "model.obj")
    for args in ([s1, str(d / "o.bin"), "2"], [s1, str(d / "o.bin"), "2", "async"],
                 [s1, str(d / "o.bin"), "2", "motion"], [s1, str(d / "o.bin"), "2", "tiled"],
                 [s2, str(d / "o.bin"), "2", "obj", model, "100"]):
        rc, out = _run([bins["facade_asan"], *args], timeout=300)
        _clean(rc, out, " ".join(args[3:]) or "sync")


def _malformed_corpus(d):
    rng = random.Random(7)
    model_dir = os.path.join(ROOT, "tests", "golden", "obj_model")
    obj = open(os.path.join(model_dir, "model.obj"), "rb").read()
    png = open(os.path.join(model_dir, "checker.png"), "rb").read()
    files = []

    def put(name, data):
        p = d / name
        p.write_bytes(data)
        files.append(str(p))

    texts = ["", "f 1/1/1 2/2/2 3/3/3\n", "v 1 2\nvt 0\nvn 0 0\nf 1/1/1 1/1/1 1/1/1\n",
             "v 0 0 0\nvt 0 0\nvn 0 0 1\nf 1//1 1//1 1//1\n", "v 0 0 0\nf 1 1 1\n",
             "v 0 0 0\nvt 0 0\nvn 0 0 1\nf -1/-1/-1 -1/-1/-1 -1/-1/-1\n",
             "v 0 0 0\nvt 0 0\nvn 0 0 1\nf 99999999999/1/1 1/1/1 1/1/1\n",
             "mtllib nothere.mtl\nusemtl\nv 0 0 0\nvt 0 0\nvn 0 0 1\nf 1/1/1 1/1/1 1/1/1\n",
             "v\nvt\nvn\nf\nusemtl x\ng\n", "v " + "9" * 10000 + " 1 1\n",
             "v 0 0 0\nvt 0 0\nvn 0 0 1\nusemtl a\nf 1/1/1 1/1/1 1/1/1\nusemtl b\n"]
    for i, t in enumerate(texts):
        put(f"t{i}.obj", t.encode())
    # an MTL with a map and a colour without arguments, next to its OBJ
    put("m.mtl", b"newmtl a\nmap_Kd\nKd\nmap_Bump\nnewmtl\n")
    put("m.obj", b"mtllib m.mtl\nv 0 0 0\nvt 0 0\nvn 0 0 1\nusemtl a\nf 1/1/1 1/1/1 1/1/1\n")
    put("garbage.obj", bytes(rng.randrange(256) for _ in range(4096)))
    for k in range(20):
        put(f"cut{k}.obj", obj[:rng.randrange(len(obj))])
    # PNG: truncations, flipped bytes, lying headers
    for k in range(25):
        put(f"cut{k}.png", png[:rng.randrange(len(png))])
    for k in range(20):
        b = bytearray(png)
        i = rng.randrange(33, len(b) - 12)
        b[i] ^= 1 << rng.randrange(8)
        put(f"flip{k}.png", bytes(b))

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + b"\0\0\0\0"

    import zlib
    sig = b"\x89PNG\r\n\x1a\n"
    put("huge.png", sig + chunk(b"IHDR", struct.pack(">IIBBBBB", 30000, 30000, 8, 6, 0, 0, 0)) +
        chunk(b"IDAT", zlib.compress(b"\0" * 100)) + chunk(b"IEND", b""))
    put("nopal.png", sig + chunk(b"IHDR", struct.pack(">IIBBBBB", 4, 4, 8, 3, 0, 0, 0)) +
        chunk(b"IDAT", zlib.compress(b"\0\1\2\3\4" * 4)) + chunk(b"IEND", b""))
    put("palidx.png", sig + chunk(b"IHDR", struct.pack(">IIBBBBB", 4, 4, 8, 3, 0, 0, 0)) +
        chunk(b"PLTE", b"\xff\0\0\0\xff\0") + chunk(b"IDAT", zlib.compress(b"\0\0\1\7\xff" * 4)) +
        chunk(b"IEND", b""))
    put("badfilter.png", sig + chunk(b"IHDR", struct.pack(">IIBBBBB", 2, 2, 8, 6, 0, 0, 0)) +
        chunk(b"IDAT", zlib.compress(b"\x09" + b"\0" * 8 + b"\x09" + b"\0" * 8)) + chunk(b"IEND", b""))
    put("longchunk.png", sig + struct.pack(">I", 0xFFFFFFF0) + b"IHDR" + b"\0" * 20)
    put("empty.png", b"")
    put("notpng.png", b"GIF89a" + b"\0" * 64)
    return files


def test_obj_and_png_readers_on_malformed_files(bins):
    d = bins["dir"] / "corpus"
    d.mkdir()
    files = _malformed_corpus(d)
    rc, out = _run([bins["ingest_asan"], *files], timeout=300)
    _clean(rc, out, "ingest corpus")
    lines = [l for l in out.splitlines() if l.startswith(str(d))]
    assert len(lines) == len(files)
    status = {l.split()[0]: int(l.split()[2]) for l in lines}
    assert status[str(d / "huge.png")] != 0 and status[str(d / "palidx.png")] != 0
    assert status[str(d / "badfilter.png")] != 0 and status[str(d / "longchunk.png")] != 0
    assert status[str(d / "t6.obj")] != 0          # index past the vertex list


def test_server_request_layer_under_tsan(bins):
    rc, out = _run([bins["server_tsan"]], env={"TSAN_OPTIONS": "halt_on_error=1"}, timeout=300)
    _clean(rc, out, "server")
    fields = dict(zip(out.split()[0::2], out.split()[1::2]))
    assert fields["done"] == "800" and fields["bad"] == "0"
    assert int(fields["refused"]) > 0 and int(fields["splits"]) > 0 and int(fields["evicted"]) > 0, out


def test_facade_async_build_under_tsan(bins):
    d = bins["dir"]
    s1, _ = _scene_files(d)
    rc, out = _run([bins["facade_tsan"], s1, str(d / "o.bin"), "3", "async"],
                   env={"TSAN_OPTIONS": "halt_on_error=1"}, timeout=300)
    _clean(rc, out, "facade async")
    assert "async build:" in out
