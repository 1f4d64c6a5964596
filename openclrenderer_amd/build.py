"""Builds libocr.so (the HIP path, gfx950) in-tree with hipcc.

Usage: python -m openclrenderer_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libocr.so")
CALIB_LIB = os.path.join(HERE, "libocr_calib.so")
FACADE_LIB = os.path.join(HERE, "libocr_engine.so")
FACADE_DEMO = os.path.join(HERE, "ocr_facade_demo")
SOURCES = ["runtime.hip", "raster.hip", "tiles.hip", "shade.hip", "atlas.hip", "batch.hip", "server.hip", "post.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("OCR_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no implicit fused multiply-add (results must match the
# arithmetic contract bit for bit); correctly rounded f32 division and sqrt;
# packed-fp32 instructions disabled.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-pthread",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return r


def build(verbose=False, jobs=4):
    os.makedirs(OBJDIR, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    while procs:
        _wait(procs.pop(0))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-pthread", "-o", tmp], verbose)
    os.replace(tmp, LIB)
    # the C++ draw API over the C ABI, and its demo program
    inc = os.path.join(os.path.dirname(HERE), "include")
    # host-only C++ (g++), linked against the C ABI and the HIP runtime API
    hip = ["-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    tmp = FACADE_LIB + ".tmp"
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", inc, os.path.join(CSRC, "facade.cpp"),
          os.path.join(CSRC, "obj_load.cpp"), "-L", HERE, "-locr", *hip, "-lz", "-Wl,-rpath,$ORIGIN", "-o", tmp],
         verbose)
    os.replace(tmp, FACADE_LIB)
    tmp = FACADE_DEMO + ".tmp"
    _run([CXX, "-O2", "-std=c++17", "-Wall", "-I", inc, os.path.join(CSRC, "facade_demo.cpp"),
          "-L", HERE, "-locr_engine", "-locr", *hip, "-Wl,-rpath,$ORIGIN", "-o", tmp], verbose)
    os.replace(tmp, FACADE_DEMO)
    # the calibration copy kernel used by profiling runs (separate library)
    tmp = CALIB_LIB + ".tmp"
    _run([HIPCC, *FLAGS, "-shared", os.path.join(CSRC, "calib.hip"), "-o", tmp], verbose)
    os.replace(tmp, CALIB_LIB)
    return LIB


def _wait(item):
    cmd, p = item
    out, _ = p.communicate()
    if p.returncode != 0:
        sys.stderr.write(out)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build_oracle(verbose=False):
    root = os.path.dirname(HERE)
    _run(["make", "-s", "-C", os.path.join(root, "oracle")], verbose)
    return os.path.join(root, "oracle", "_build", "libocr_oracle.so")


if __name__ == "__main__":
    print(build(verbose=True))
