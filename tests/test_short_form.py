"""The raster passes' short form of the sample walk (raster.hip:
short_form_ok / sweep_samples) gives the same pixel, row and skip decision
as the reference's walk (kernel1, cl2.cl:5000-5127) for every sample of
every triangle it certifies: checked by tests/short_form_check.c on the
CPU over widths 1-520 and sampled widths to 4096, first rows 0-160 and
sampled rows below 2^22, every sample count the short form admits, for the
camera (op 500) and light (op 300) fragment sizes."""
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
