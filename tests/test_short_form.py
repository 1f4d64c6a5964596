"""Exact rewrites of the reference's arithmetic, each checked on the CPU:

* the raster passes' short form of the sample walk (raster.hip:
  short_form_ok / sweep_samples) gives the same pixel, row and skip
  decision as the reference's walk (kernel1, cl2.cl:5000-5127) for every
  sample of every triangle it certifies: checked by
  tests/short_form_check.c over widths 1-520 and sampled widths to 4096,
  first rows 0-160 and sampled rows below 2^22, every sample count the short form admits, for
  the camera (op 500) and light (op 300) fragment sizes;
* the stored depth to distance in one product (ocr_math.h depth_of),
  exhaustively (tests/depth_scale_check.c)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_short_form_equals_walk(tmp_path):
    exe = str(tmp_path / "short_form_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "short_form_check.c"), "-lm"],
                   check=True)
    procs = [subprocess.Popen([exe, op, "520", "160", "20"], stdout=subprocess.PIPE, text=True)
             for op in ("500", "300")]
    for p in procs:
        out = p.communicate(timeout=600)[0]
        print(out)
        assert p.returncode == 0 and " 0 mismatches" in out, out
        assert int(out.split("certified ")[1].split(",")[0]) > 1000000


def test_depth_scale_one_product(tmp_path):
    """RN(RN(u) * (350000 * 2^-32)) equals the reference's
    (RN(u) / 2^32) * 350000 for every value RN(u) takes (0 and every float in
    [1, 2^32]): tests/depth_scale_check.c, exhaustive."""
    exe = str(tmp_path / "depth_scale_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "depth_scale_check.c"), "-lm"],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and " 0 mismatches" in r.stdout, r.stdout
