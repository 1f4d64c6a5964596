"""The shading pass replaces some correctly rounded divisions with a
three-instruction form (div_rcp: q = a*y, r = fma(-q, b, a), fma(r, y, q)
with y = RN(1/b)) and skips tan where the bias saturates.  These checks fix
that both changes leave every result bit-identical:

* constant divisors 9, 125, 2.5, 255, 100 (shading) and 350000, 300, 500
  (prearrange: z / depth_far, area / op_size): a = 0 and every finite
  a >= 2^-100, exhaustively (the values divided are integers, depths of at
  least 20 or sums of weights in [0, 9]);
* divisors known only at run time (fov, light radius, the 1/z plane value):
  Markstein's correction step, exercised on random pairs over exponents
  -60..60 and exhaustively over every dividend for sixteen divisor
  significands (extremes, badly rounded reciprocals, seeded random ones);
* bytes over 255 in two operations (unorm8), for all 256 bytes;
* tan(a) >= 1 for every binary32 a from float(pi/4) rounded up to
  rational_acos(0.05), so the bias clamp gives SHADOWBIAS there.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def div_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("div") / "exact_division_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "exact_division_check.c"),
                    "-lm", "-ldl"], check=True)
    return exe


def test_constant_divisors_exhaustive(div_check):
    procs = [subprocess.Popen([div_check, "const", b], stdout=subprocess.PIPE, text=True)
             for b in ("9", "125", "2.5", "255", "350000", "300", "500", "65535", "100")]
    outs = [(p.communicate(timeout=600)[0], p.returncode) for p in procs]
    for out, rc in outs:
        sys.stdout.write(out)
        assert rc == 0 and " 0 mismatches" in out, out


def test_runtime_divisors_random_pairs(div_check):
    r = subprocess.run([div_check, "random", "200000000"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and " 0 mismatches" in r.stdout, r.stdout


# Divisor significands checked exhaustively over every dividend: the
# largest and smallest significands, ones whose reciprocal rounds far from
# 1/b, and seeded random ones.  A divisor's exponent only scales the
# quotient, so each covers its whole significand class.  (Offline, 56 more
# random significands from random.seed(2026) gave 0 mismatches as well.)
RUNTIME_SIGNIFICANDS = ("0x1.fffffep+0", "0x1.000002p+0", "0x1.555556p+0", "0x1.800000p+0", "0x1.fffffcp+0",
                        "0x1.c71c72p+0", "0x1.99999ap+0", "0x1.6a09e6p+0", "0x1.2f6a3cp+0", "0x1.b3e0d2p+0",
                        "0x1.07c1f0p+0", "0x1.e7d5a8p+0", "0x1.4c2b96p+0", "0x1.91d3e4p+0", "0x1.3a0f5ep+0",
                        "0x1.d8e47ap+0")


def test_runtime_divisor_significands_exhaustive(div_check):
    procs = [subprocess.Popen([div_check, "const", b], stdout=subprocess.PIPE, text=True)
             for b in RUNTIME_SIGNIFICANDS]
    for p in procs:
        out = p.communicate(timeout=900)[0]
        assert p.returncode == 0 and " 0 mismatches" in out, out


def test_tan_saturates_from_quarter_pi(div_check, oracle):
    from oracle.binding import LIB_PATH
    r = subprocess.run([div_check, "tan", LIB_PATH], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout


def test_byte_over_255_two_term_form():
    """unorm8 (the device's b / 255 for a byte b): fma(b, hi, RN(b * lo))
    with hi = RN(1/255) and lo = RN(1/255 - hi) equals RN(b / 255) for all
    256 bytes, in exact rational arithmetic (each step rounded once)."""
    from fractions import Fraction

    import numpy as np

    def rn32(x):
        f = np.float32(float(x))
        cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
        # nearest, ties to the even significand
        return min(cands, key=lambda v: (abs(Fraction(float(v)) - x), int(v.view(np.uint32)) & 1))

    hi = rn32(Fraction(1, 255))
    lo = rn32(Fraction(1, 255) - Fraction(float(hi)))
    assert float(hi) == float.fromhex("0x1.010102p-8") and float(lo) == float.fromhex("-0x1.fdfdfep-33")
    for b in range(256):
        p = rn32(Fraction(b) * Fraction(float(lo)))
        q = rn32(Fraction(b) * Fraction(float(hi)) + Fraction(float(p)))
        assert q == rn32(Fraction(b, 255)), b
