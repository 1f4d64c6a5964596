"""Device-side pins of the short reciprocal sequences (csrc/ocr_math.h),
each checked for every input on the GPU:

* rcp_rn: the hardware reciprocal estimate refined by one Newton step equals
  the IEEE quotient 1/b for every binary32 b with |b| in [2^-126, 2^126]
  (the kernels use it for divisors in [2^-60, 2^60] and for the depth
  quotient, whose other inputs depth_quant handles explicitly);
* inv_sqrt_cr: equals 1.0f / sqrtf(d) for every binary32 d >= 0, +inf and
  NaN included;
* exp_f equals the oracle's binary32 exp bit for bit on 4 M inputs spread
  over its range and on its special values;
* sqrt_cr (the hardware root corrected once by the residual times half the
  hardware reciprocal root, with the full sequence for inputs outside
  [2^-100, FLT_MAX]): equals sqrtf(x) for every binary32 pattern, negative
  numbers, zeros, infinities and NaNs included.
"""
import ctypes as C
import os

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def calib(gpu_ctx):
    lib = C.CDLL(os.path.join(ROOT, "openclrenderer_amd", "libocr_calib.so"))
    lib.ocr_calib_rcp_check.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.POINTER(C.c_ulonglong),
                                        C.POINTER(C.c_uint32)]
    return lib


def _check(lib, mode, lo, hi):
    bad, first = C.c_ulonglong(), C.c_uint32()
    assert lib.ocr_calib_rcp_check(mode, C.c_uint32(lo), C.c_uint32(hi), C.byref(bad), C.byref(first)) == 0
    print(f"mode {mode} [{lo:08x}, {hi:08x}]: {bad.value} mismatches, first {first.value:08x}", flush=True)
    return bad.value, first.value


def test_rcp_rn_matches_ieee_reciprocal(calib):
    for lo, hi in ((0x00800000, 0x7E800000), (0x80800000, 0xFE800000)):
        bad, first = _check(calib, 0, lo, hi)
        assert bad == 0, f"first mismatch at bits {first:08x}"


def test_inv_sqrt_cr_matches_ieee(calib):
    # +0 .. +inf, then the positive NaNs
    for lo, hi in ((0x00000000, 0x7F800000), (0x7F800001, 0x7FFFFFFF)):
        bad, first = _check(calib, 1, lo, hi)
        assert bad == 0, f"first mismatch at bits {first:08x}"


def test_sqrt_cr_matches_ieee(calib):
    for lo, hi in ((0x00000000, 0x7FFFFFFF), (0x80000000, 0xFFFFFFFF)):
        bad, first = _check(calib, 2, lo, hi)
        assert bad == 0, f"first mismatch at bits {first:08x}"


def test_exp_matches_oracle(calib, oracle):
    import numpy as np
    calib.ocr_calib_eval.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_uint32]
    oracle.lib.oracle_exp_n.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-110, 92, 1 << 22), rng.uniform(-1, 1, 1 << 16),
                        np.array([0.0, -0.0, -104.0, -103.97, 88.72, 88.75, 88.76, np.inf, -np.inf, np.nan])
                        ]).astype(np.float32)
    gpu = np.zeros_like(x)
    ref = np.zeros_like(x)
    assert calib.ocr_calib_eval(0, x.ctypes.data, gpu.ctypes.data, x.size) == 0
    oracle.lib.oracle_exp_n(x.ctypes.data, ref.ctypes.data, x.size)
    same = (gpu.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(gpu) & np.isnan(ref))
    assert same.all(), f"{int((~same).sum())} mismatches, first at x = {x[~same][0]!r}"
