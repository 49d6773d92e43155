"""The device voxeliser against the host one (and so the reference): masks and narrow-band
distances bit for bit on the served bodies and the wing (the committed golden digests of the
reference's voxelisation), winding numbers within a few ulps of atan2, and a lattice-sized body."""
import math
import time

import numpy as np
import pytest

from orrery_amd import geometry as g
from orrery_amd.serve import body_mesh
from test_geometry_pins import golden, voxel_cases
from test_setup_pins import digest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(voxel_cases()))
def test_device_voxels_and_distance_are_the_host_ones(name):
    v, t, spec = voxel_cases()[name]
    m = g.TriangleMesh(v, t)
    host = g.voxelize(m, spec)
    dev = g.voxelize(m, spec, device=True)
    dw = np.abs(dev.winding - host.winding)
    print(f"{name}: {spec.cells} cells, {dev.rechecked} decided on the host, max |w_dev - w_host| {dw.max():.2e}",
          flush=True)
    assert np.array_equal(dev.solid, host.solid)
    assert digest(dev.solid) == golden()["voxels"][name]["digests"]["mask"]
    assert dw.max() < 1e-12
    hd = g.signed_distance_band(m, host, 3.0)
    dd = g.signed_distance_band(m, dev, 3.0, device=True)
    assert np.array_equal(dd.distance, hd.distance) and np.array_equal(dd.in_band, hd.in_band)
    assert digest(dd.distance) == golden()["voxels"][name]["digests"]["distance"]


def test_lattice_sized_capsule():
    """The 512 x 256 x 256 workload's body: a capsule of diameter 64 cells pitched 20 degrees,
    voxelised over its box on both sides."""
    mesh = body_mesh("capsule", 64.0).transformed(g.rotation_y(math.radians(20.0)), (0.0, 0.0, 0.0),
                                                  (128.0, 127.75, 127.75))
    lo, hi = mesh.bounds()
    first = [int(math.floor(lo[a])) - 4 for a in range(3)]
    count = [int(math.ceil(hi[a])) + 4 - first[a] + 1 for a in range(3)]
    spec = g.GridSpec(tuple(first[a] - 0.5 for a in range(3)), 1.0, *count)
    t0 = time.perf_counter()
    host = g.voxelize(mesh, spec)
    hd = g.signed_distance_band(mesh, host, 3.0)
    t1 = time.perf_counter()
    dev = g.voxelize(mesh, spec, device=True)
    dd = g.signed_distance_band(mesh, dev, 3.0, device=True)
    t2 = time.perf_counter()
    print(f"capsule box {count}: host {t1 - t0:.3f} s, device {t2 - t1:.3f} s, {dev.rechecked} decided on the host",
          flush=True)
    assert np.array_equal(dev.solid, host.solid)
    assert np.array_equal(dd.distance, hd.distance)
