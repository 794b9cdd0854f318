"""Writes tests/golden/reference_scenes.json from the reference's scene files.

Run in a checkout that has the reference at /root/reference (its scenes/*.ll are the only
inputs read).  For each scene it records, from the independent interpreter in
tests/llinterp.py (which shares no code with either tape builder under test):
  - the number of body instructions and the opcode histogram the tape must have,
  - the register-use count of every SSA value,
  - f(x, y, z, t) in float64 at 48 seeded points, with the distance to the nearest
    floor/mod/atan2 discontinuity (points too close to one are not compared).

    python tests/golden/make_reference_fixtures.py
"""
from __future__ import annotations

import glob
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import llinterp  # noqa: E402

SCENES = "/root/reference/scenes"


def main():
    out = {}
    rng = np.random.default_rng(20240611)
    for path in sorted(glob.glob(os.path.join(SCENES, "*.ll"))):
        name = os.path.splitext(os.path.basename(path))[0]
        prog = llinterp.Program(open(path).read())
        pts = np.concatenate([rng.uniform(-1.5, 1.5, (48, 3)), rng.uniform(0.0, 1.0, (48, 1))], axis=1)
        pts = pts.astype(np.float32).astype(np.float64)
        vals, margins = [], []
        for p in pts:
            v, m = prog.eval(*p)
            vals.append(v)
            margins.append(m if np.isfinite(m) else 1e30)
        out[name] = {
            "body": len(prog.ins),
            "histogram": prog.histogram(),
            "uses": {str(k): v for k, v in sorted(prog.uses().items())},
            "points": pts.tolist(),
            "values": vals,
            "margins": margins,
        }
    with open(os.path.join(HERE, "reference_scenes.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(f"wrote {len(out)} scenes")


if __name__ == "__main__":
    main()
