"""The tiled raster path visits, per tile, exactly the pixels the
reference's fragment walk visits there (kernel1, cl2.cl:5000-5127; kernel2,
cl2.cl:5419-5545): tests/tile_walk_check.c restates the walk by its cursor
steps over every fragment of a triangle and compares it with both of the
path's rules (the box clipped to the tile for triangles walk_is_rect
certifies, the walk's own samples of the tile's rows for the others) on
every width up to 120 at first rows 0-80 (where the row fma lags) and
heights up to 40, plus random boxes up to 5119 x 2879, for the camera (op
500) and light (op 300) fragment sizes and five tile shapes."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_tile_rules_equal_walk(tmp_path):
    exe = str(tmp_path / "tile_walk_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "tile_walk_check.c"), "-lm"],
                   check=True)
    procs = [subprocess.Popen([exe, op, "120", "80", "40", "40"], stdout=subprocess.PIPE, text=True)
             for op in ("500", "300")]
    for p in procs:
        out = p.communicate(timeout=600)[0]
        print(out)
        assert p.returncode == 0 and " 0 mismatches" in out, out
        rect = int(out.split("rect ")[1].split(",")[0])
        general = int(out.split("general ")[1].split(" ")[0])
        wrong = int(out.split("would be wrong for ")[1].split(")")[0])
        # both rules are exercised, and the certificate is what keeps the
        # box rule from the triangles where it would be wrong
        assert rect > 50000 and general > 100 and wrong > 50
