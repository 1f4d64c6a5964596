"""The C-ABI library loads and exports every entry point include/ocr.h
declares, and the record layouts match the reference's host structs.  No
compute calls: these run without a GPU."""
import ctypes
import os
import re

import numpy as np

from openclrenderer_amd import abi, runtime
from openclrenderer_amd import build as b

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ocr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(ocr_[a-z_0-9]+)\s*\(", text)
    return sorted(set(names))


def test_header_lists_match_binding():
    assert declared_functions() == sorted(runtime.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(b.LIB):
        b.build()
    lib = ctypes.CDLL(b.LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_record_sizes_match_reference_structs():
    # cl2.cl:131-151 (vertex 48 B, triangle 144 B), :99-124 (descriptor 144 B
    # with cl_float4 alignment), :77-87 (light 56 B -> 64 B aligned)
    assert abi.VERTEX.itemsize == 48
    assert abi.TRIANGLE.itemsize == 144
    assert abi.OBJ_DESC.itemsize == 144
    assert abi.LIGHT.itemsize == 64
    assert abi.OBJ_DESC.fields["feature_flag"][1] == 132
    assert abi.OBJ_DESC.fields["scale"][1] == 96
    assert abi.LIGHT.fields["is_static"][1] == 52


def test_header_constants():
    text = open(os.path.join(ROOT, "include", "ocr.h")).read()
    assert "#define OCR_DEPTH_FAR        350000.0f" in text
    assert "#define OCR_OP_SIZE          500" in text
    assert "#define OCR_OP_SIZE_LIGHT    300" in text
    assert "#define OCR_BUF_ERROR        20" in text
    assert np.uint32(abi.NO_TRIANGLE) == 0xFFFFFFFF


def test_server_records_match_header(tmp_path):
    """The ctypes mirrors of the served-form records have the C layout the
    header gives (sizes and field offsets, measured by compiling the header)."""
    import subprocess
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "ocr.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(ocr_server_config),"
        " sizeof(ocr_completion), sizeof(ocr_server_info), offsetof(ocr_completion, host),"
        " offsetof(ocr_completion, status), offsetof(ocr_server_config, hbm_fraction),"
        " sizeof(ocr_frame_stats)); return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(runtime.ServerConfig), ctypes.sizeof(runtime.Completion),
            ctypes.sizeof(runtime.ServerInfo), runtime.Completion.host.offset, runtime.Completion.status.offset,
            runtime.ServerConfig.hbm_fraction.offset, ctypes.sizeof(runtime.FrameStats)]
    assert got == want
    assert got[:3] == [24, 64, 72]


def test_engine_library_exports_scene_ingest():
    """include/ocr_obj.h (OBJ/MTL + PNG ingest) lives in libocr_engine.so."""
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "ocr_obj.h")).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(ocr_[a-z_0-9]+)\s*\(", text)))
    assert names == ["ocr_obj_free", "ocr_obj_get", "ocr_obj_load", "ocr_obj_nparts", "ocr_obj_ntri",
                     "ocr_png_free", "ocr_png_load"]
    if not os.path.exists(b.FACADE_LIB):
        b.build()
    lib = ctypes.CDLL(b.FACADE_LIB)
    assert not [n for n in names if not hasattr(lib, n)]
