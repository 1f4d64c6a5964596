"""Writes tests/golden/config1_oracle.json: sha256 of the oracle's config-1
depth, triangle-id, RGBA8 and normal buffers plus its work-list counts.

The reference publishes no outputs for this path and cannot run here, so this
fixture is a regression anchor for the oracle itself: it changes only when the
oracle's arithmetic changes, which must then be justified against the closed
form and brute-force pins in tests/test_oracle_*.py.
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from openclrenderer_amd import scenes  # noqa: E402
from oracle.binding import Oracle  # noqa: E402

sc, view = scenes.config1()
out = Oracle().render(sc, view)
rec = {k: hashlib.sha256(out[k].tobytes()).hexdigest() for k in ("depth", "tri_id", "rgba", "normals")}
rec["n_frag"] = out["stats"]["n_frag"]
rec["n_cut"] = out["stats"]["n_cut"]
rec["covered_pixels"] = int((out["depth"] != 0xFFFFFFFF).sum())
json.dump(rec, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "config1_oracle.json"), "w"), indent=1)
print(rec)
