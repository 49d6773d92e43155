"""Physical validation of the 3D adapter on the device, after the reference's
tests/aero_validation/lbm3d/test_sphere.cpp and body_case.h: the drag of a sphere (the
64-segment tessellation, voxelised by the adapter, interpolated wall, gradual start, outlet
sponge, free-slip sides) against the standard drag curve of Schiller and Naumann (1933),
Cd = 24 / Re (1 + 0.15 Re^0.687), within the reference's acceptance of 15 percent at Re 10, 100,
300 and 1000, with the reference's resolutions and domains; the steady cases converge."""
import math

import pytest

from orrery_amd import geometry as g
from orrery_amd.aero import AeroGeometry, FlowConditions, LbmAeroSolver3D, LbmSolver3DOptions, LesMode, \
    ReferenceValues

pytestmark = pytest.mark.gpu


def schiller_naumann(re):
    return 24.0 / re * (1.0 + 0.15 * re ** 0.687)


def water_at(re):
    """body_case.h water_at: U 0.1 m/s, rho 1000, Re = rho U L / mu with L = 1 m."""
    c = FlowConditions(speed=0.1, density=1000.0)
    c.viscosity = c.density * c.speed * 1.0 / re
    return c


@pytest.mark.parametrize("re,cells,upstream,downstream,width,duration,average,steady", [
    (10.0, 12.0, 6.0, 10.0, 10.0, 30.0, 10.0, True),
    (100.0, 16.0, 4.0, 8.0, 6.0, 30.0, 10.0, True),
    (300.0, 16.0, 4.0, 8.0, 6.0, 30.0, 10.0, False),
    (1000.0, 48.0, 2.0, 4.0, 4.0, 20.0, 10.0, False),
])
def test_sphere_drag_follows_schiller_naumann(re, cells, upstream, downstream, width, duration, average, steady):
    geo = AeroGeometry(g.sphere_mesh(0.5, g.RevolutionResolution(64, 0.05)),
                       ReferenceValues(area=0.25 * math.pi, length=1.0))
    opt = LbmSolver3DOptions(cells_per_length=cells, lattice_velocity=0.05, upstream=upstream,
                             downstream=downstream, width=width, duration=duration, average_window=average,
                             les=LesMode.off, steps_per_call=2000, sample_every=10)
    s = LbmAeroSolver3D(opt)
    s.setup(geo, water_at(re))
    r = s.run()
    c = s.info().config
    want = schiller_naumann(re)
    print(f"measured: sphere Re {re:g} cells/L {cells:g} grid {c.nx}x{c.ny}x{c.nz} steps {r.iterations}: "
          f"Cd {r.cd:.4f}, Schiller-Naumann {want:.4f}, ratio {r.cd / want:.4f}, converged {r.converged}",
          flush=True)
    assert abs(r.cd / want - 1.0) < 0.15
    if steady:
        assert r.converged
