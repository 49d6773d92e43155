"""Writes tests/golden/setup_fixtures.json from oracle/_ref (the reference's setup code compiled from
its own sources): digests of masks, node classes and boundary links, inlet reconstructions of
fixed inputs, and validation messages. Run where the reference tree was available to build
oracle/_ref:

    make -C oracle/ref && python tests/golden/make_setup_fixtures.py
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import refcheck  # noqa: E402
from test_setup_pins import bad_configs, c1, c2, digest, setup_cases  # noqa: E402


def main():
    ref = refcheck.load()
    out = {"source": "oracle/_ref (reference setup code compiled from the reference sources)"}
    solid = refcheck.circle_mask(400, 160, 80.0, 80.0, 16.0)
    s2, d2 = refcheck.sphere_mask(256, 128, 128, (64.0, 64.0, 64.0), 16.0)
    sb, db = refcheck.box_mask(12, 10, 9, (2.0, 3.5, 1.0), (7.0, 6.0, 4.0))
    out["masks"] = {"c1_circle": digest(solid), "c2_sphere": digest(s2, d2), "box_12_10_9": digest(sb, db)}
    cases = {}
    for name, (c, m) in setup_cases().items():
        cell, d, q = refcheck.boundary_links3(c, m)
        cases[name] = {"links": digest(cell, d, q), "classes": digest(refcheck.classify3(c, m)), "count": len(cell)}
    out["cases"] = cases
    c, m = c2()
    cell, d, q = refcheck.boundary_links3(c, m)
    out["c2_links"] = digest(cell, d, q)
    rng = np.random.default_rng(2024)
    inlet3 = []
    inlet2 = []
    for _ in range(8):
        f = 1.0 / 19 + 1e-3 * rng.standard_normal(19)
        u = [0.05 * float(rng.random()), 0.01 * float(rng.standard_normal()), 0.01 * float(rng.standard_normal())]
        g = f.copy()
        ref.ref3d_inlet_west(refcheck.ptr(g), *u)
        inlet3.append({"in": f.tolist(), "u": u, "out": g.tolist()})
        f2 = 1.0 / 9 + 1e-3 * rng.standard_normal(9)
        g2 = f2.copy()
        ref.ref2d_zou_he_west(refcheck.ptr(g2), u[0], u[1])
        inlet2.append({"in": f2.tolist(), "u": u[:2], "out": g2.tolist()})
    out["inlet"] = {"d3q19": inlet3, "d2q9": inlet2}
    out["validation"] = [refcheck.validate3(c, m) for c, m in bad_configs()]
    with open(os.path.join(HERE, "setup_fixtures.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", os.path.join(HERE, "setup_fixtures.json"))


if __name__ == "__main__":
    main()
