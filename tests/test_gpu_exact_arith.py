"""Device-side pins of the short reciprocal sequences (csrc/ocr_math.h),
each checked for every input on the GPU:

* rcp_rn: the hardware reciprocal estimate refined by one Newton step equals
  the IEEE quotient 1/b for every binary32 b with |b| in [2^-126, 2^126]
  (the kernels use it for divisors in [2^-60, 2^60] and for the depth
  quotient, whose other inputs depth_quant handles explicitly);
* inv_sqrt_cr: equals 1.0f / sqrtf(d) for every binary32 d >= 0, +inf and
  NaN included.
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
