"""Writes tests/golden/geometry_fixtures.json from oracle/_ref (the reference's geometry code compiled
from its own sources): sizes and digests of the primitive tessellations, and of the voxel masks,
winding numbers and signed-distance bands of placed bodies. Run where the reference tree was
available to build oracle/_ref:

    make -C oracle/ref && python tests/golden/make_geometry_fixtures.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import refcheck  # noqa: E402
from test_geometry_pins import mesh_cases, voxel_cases  # noqa: E402
from test_setup_pins import digest  # noqa: E402


def main():
    out = {"source": "oracle/_ref (reference geometry code compiled from the reference sources)", "meshes": {},
           "voxels": {}}
    for name, (_, ref) in mesh_cases().items():
        v, t = ref()
        out["meshes"][name] = {"size": [len(v), len(t)], "digest": digest(v, t)}
    for name, (v, t, spec) in voxel_cases().items():
        solid, w = refcheck.voxelize(v, t, spec.origin, spec.spacing, spec.nx, spec.ny, spec.nz)
        d = refcheck.signed_distance_band(v, t, spec.origin, spec.spacing, spec.nx, spec.ny, spec.nz, solid)
        band = 3.0 * spec.spacing
        out["voxels"][name] = {"grid": [spec.nx, spec.ny, spec.nz],
                               "digests": {"solid": int(solid.sum()), "mask": digest(solid), "winding": digest(w),
                                           "distance": digest(d), "in_band": int((abs(d) < band).sum())}}
    path = os.path.join(HERE, "geometry_fixtures.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
