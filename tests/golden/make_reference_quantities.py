"""Writes reference_quantities.json: the reference's reference_quantities
(src/geom/reference_quantities.cpp, compiled into oracle/_ref by oracle/ref/Makefile) on the
primitive bodies the runner builds, at two raster resolutions. Run from the repository root
with oracle/_ref built: python tests/golden/make_reference_quantities.py"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import refcheck  # noqa: E402
from orrery_amd import experiment as ex, runner  # noqa: E402

CASES = {
    "sphere_d1.3": (ex.AeroShape.sphere, {"diameter_m": 1.3}),
    "capsule_d3.9": (ex.AeroShape.capsule, {"diameter_m": 3.9}),
    "cone_d1_l2": (ex.AeroShape.cone, {"diameter_m": 1.0, "length_m": 2.0}),
    "cylinder_d0.5_l2": (ex.AeroShape.cylinder, {"diameter_m": 0.5, "length_m": 2.0}),
    "cube_2": (ex.AeroShape.cube, {"side_m": 2.0}),
}
RESOLUTIONS = (64, runner.REFERENCE_RASTER)


def mesh_of(shape, dims):
    r = ex.GeometryRef(shape=shape)
    for k, v in dims.items():
        setattr(r.dimensions, k, v)
    return runner.prepare_geometry(r).mesh


def main():
    out = {}
    for key, (shape, dims) in CASES.items():
        m = mesh_of(shape, dims)
        for res in RESOLUTIONS:
            out[f"{key}@{res}"] = refcheck.reference_quantities(m.vertices, m.triangles, res)
    path = os.path.join(HERE, "reference_quantities.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(f"wrote {path}: {len(out)} cases")


if __name__ == "__main__":
    main()
