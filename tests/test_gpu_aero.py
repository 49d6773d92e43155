"""The aerodynamic adapter on the device, after the reference's
tests/aero/lbm3d/test_lbm_aero_solver3d.cpp, and the reference's own coarse-sphere Cd
(3.5172643919827955, FP64 CPU) reproduced by the FP64 lattice."""
import math

import numpy as np
import pytest

from orrery_amd import geometry as g
from orrery_amd.aero import AeroError, AeroGeometry, FlowConditions, LbmAeroSolver3D, LesMode, ReferenceValues
from orrery_amd.config import SolidWall
from test_aero_plan import REFERENCE_CD, coarse_options, flow_at, sphere_geometry

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage,tolerance", [("f64", 1e-12), ("f32", 1e-3)])
def test_reference_cd_of_the_coarse_sphere(storage, tolerance):
    s = LbmAeroSolver3D(coarse_options(duration=2.0, average_window=1.0, storage=storage))
    s.setup(sphere_geometry(0.02), flow_at(20.0, 0.02))
    r = s.run()
    print(f"measured: {storage} Cd {r.cd!r} reference {REFERENCE_CD!r} rel {abs(r.cd / REFERENCE_CD - 1):.3e}")
    assert r.iterations == 320
    assert abs(r.cd / REFERENCE_CD - 1.0) < tolerance


def test_sphere_through_the_solver_interface():
    s = LbmAeroSolver3D(coarse_options())
    geo = sphere_geometry(0.02)
    cond = flow_at(20.0, 0.02)
    s.setup(geo, cond)
    nx = s.info().config.nx
    assert nx in (49, 50) and s.info().config.ny == 25
    r = s.run()
    assert r.iterations == 960
    print(f"measured: cd {r.cd} cl {r.cl} cy {r.cy} cm {r.cm}")
    assert math.isfinite(r.cd) and 1.5 < r.cd < 6.0
    assert abs(r.cl) < 0.05 * r.cd and abs(r.cy) < 0.05 * r.cd and abs(r.cm) < 0.05 * r.cd
    q = cond.dynamic_pressure()
    assert r.force[0] == pytest.approx(r.cd * q * geo.reference.area, rel=1e-9)
    assert s.instantaneous()[0] > 0.0
    assert len(s.history()[0]) == 96 and len(s.side_history()) == 96
    surf = r.surface
    assert len(surf.cp) > 0 and surf.points.shape == (len(surf.cp), 3)
    assert surf.areas.sum() == pytest.approx(geo.mesh.surface_area(), rel=1e-12)
    front = int(np.argmax(surf.cp))
    assert 1.0 < surf.cp[front] < 2.0
    assert surf.points[front, 0] < 0.0 and surf.normals[front, 0] < -0.5
    v = s.velocity_field()
    assert v.shape == (25, 25, nx, 3)
    assert s.info().dx == pytest.approx(0.0025)


def test_wall_pressure_equals_the_full_field_readback():
    """The device-sampled wall densities are the values of the full macroscopic field."""
    from orrery_amd import abi
    import ctypes
    s = LbmAeroSolver3D(coarse_options(duration=1.0))
    s.setup(sphere_geometry(0.02), flow_at(20.0, 0.02))
    r = s.run()
    lib = abi.load()
    c = s.info().config
    rho = np.empty(c.cells)
    assert lib.lbm3d_macroscopic(s.lattice_handle(), abi.ptr(rho, ctypes.c_double), None, None, None) == 0
    rho = rho.reshape(c.nz, c.ny, c.nx)
    rho_inf = 0.0
    for z in range(c.nz):
        for y in range(c.ny):
            rho_inf += rho[z, y, 1]
    rho_inf /= c.ny * c.nz
    u = s.info().lattice_velocity
    speed = ctypes.c_double()
    assert lib.lbm3d_max_speed(s.lattice_handle(), ctypes.byref(speed)) == 0
    ux = np.empty(c.cells)
    uy = np.empty(c.cells)
    uz = np.empty(c.cells)
    assert lib.lbm3d_macroscopic(s.lattice_handle(), None, abi.ptr(ux, ctypes.c_double), abi.ptr(uy, ctypes.c_double),
                                 abi.ptr(uz, ctypes.c_double)) == 0
    fluid = s.mask().solid == 0
    assert speed.value == np.sqrt(ux * ux + uy * uy + uz * uz)[fluid].max()
    # Wall samples: the distinct cells of the boundary links, in cell order.
    n = ctypes.c_size_t()
    lib.lbm3d_links(s.lattice_handle(), None, None, None, 0, ctypes.byref(n))
    cells = np.empty(n.value, dtype=np.uint32)
    lib.lbm3d_links(s.lattice_handle(), abi.ptr(cells, ctypes.c_uint32), None, None, n.value, ctypes.byref(n))
    wall = np.unique(cells)
    flat = rho.reshape(-1)
    want = np.array([(1.0 / 3.0) * (flat[k] - rho_inf) / (0.5 * u * u) for k in wall])
    assert np.array_equal(r.surface.cp, want)


def test_angles_rotate_the_wind_frame():
    s = LbmAeroSolver3D(coarse_options(duration=3.0))
    cond = flow_at(20.0, 0.02)
    cond.angle_of_attack = math.radians(30.0)
    cond.sideslip = math.radians(20.0)
    s.setup(sphere_geometry(0.02), cond)
    r = s.run()
    d = cond.drag_direction()
    along = float(r.force @ d)
    total = float(np.linalg.norm(r.force))
    assert along == pytest.approx(total, rel=2e-3)
    assert r.cd > 1.0
    with pytest.raises(AeroError, match="aligned with the wind"):
        s.velocity_field()


def short_sphere_run(steps_per_call, area_scale):
    s = LbmAeroSolver3D(coarse_options(duration=2.5, average_window=1.0, steps_per_call=steps_per_call))
    geo = sphere_geometry(0.02)
    geo.reference.area *= area_scale
    s.setup(geo, flow_at(20.0, 0.02))
    r = s.run()
    t, d, _ = s.history()
    return r, t, d, s.side_history(), s.moving_average()


def test_averages_do_not_depend_on_the_call_size():
    ref = short_sphere_run(400, 1.0)
    assert ref[0].iterations == 400 and len(ref[1]) == 40
    for call in (1, 7, 25):
        run = short_sphere_run(call, 1.0)
        print(f"steps_per_call {call}", flush=True)
        assert run[0].iterations == 400
        assert np.array_equal(run[1], ref[1]) and np.array_equal(run[2], ref[2]) and np.array_equal(run[3], ref[3])
        assert run[0].cd == ref[0].cd and run[4][0] == ref[4][0]


def test_history_and_result_share_the_reference_area():
    unit = short_sphere_run(400, 1.0)
    doubled = short_sphere_run(400, 2.0)
    assert doubled[0].cd == pytest.approx(0.5 * unit[0].cd, rel=1e-12)
    assert doubled[4][0] == pytest.approx(doubled[0].cd, rel=1e-12)
    assert unit[4][0] == pytest.approx(unit[0].cd, rel=1e-12)
    worst = 0.0
    for a, b in zip(unit[2], doubled[2]):
        worst = max(worst, abs(b) if a == 0.0 else abs(b / (0.5 * a) - 1.0))
    assert worst < 1e-12


def test_fast_local_flow_and_divergence_are_reported():
    plate = AeroGeometry(g.box_mesh(0.25, 1.0, 1.0), ReferenceValues(area=1.0, length=1.0))
    squeezed = LbmAeroSolver3D(coarse_options(width=1.3, duration=3.0))
    squeezed.setup(plate, flow_at(20.0, 1.0))
    assert squeezed.result().has_warning("lbm.blockage")
    assert squeezed.run().has_warning("lbm.local_mach")
    unstable = LbmAeroSolver3D(coarse_options(les=LesMode.off, start_ramp=0.0, wall=SolidWall.halfway))
    unstable.setup(sphere_geometry(1.0), flow_at(2e5, 1.0))
    assert unstable.result().has_warning("lbm.reynolds")
    r = unstable.run()
    print(f"measured: unstable cd {r.cd}")
    assert r.has_warning("lbm.diverged") and r.has_warning("lbm.unconverged") and not r.converged


def test_transient_is_not_converged_and_a_steady_wake_is():
    transient = short_sphere_run(400, 1.0)[0]
    assert not transient.converged and transient.has_warning("lbm.unconverged")
    s = LbmAeroSolver3D(coarse_options(duration=40.0, average_window=5.0))
    s.setup(sphere_geometry(0.02), flow_at(10.0, 0.02))
    r = s.run()
    print(f"measured: steady wake cd {r.cd} converged {r.converged}")
    assert r.converged and not r.has_warning("lbm.unconverged")


def test_wall_pressure_against_the_oracle_moments():
    """The device-sampled wall Cp and largest speed of the FP64 lattice against the oracle's
    moments (macroscopic(), lbm3d_cpu.cpp:154-186) of the same state: the device populations
    are handed to the oracle and its moments give the expected samples."""
    import ctypes
    from oracle import Oracle3D
    from orrery_amd import abi
    s = LbmAeroSolver3D(coarse_options(duration=1.0, storage="f64"))
    s.setup(sphere_geometry(0.02), flow_at(20.0, 0.02))
    r = s.run()
    lib = abi.load()
    c = s.info().config
    f = np.empty(19 * c.cells)
    assert lib.lbm3d_populations(s.lattice_handle(), abi.ptr(f, ctypes.c_double)) == 0
    cpu = Oracle3D(c, s.mask())
    cpu.set_populations(f)
    m = cpu.macroscopic()
    rho = m.density.reshape(c.nz, c.ny, c.nx)
    rho_inf = 0.0
    for z in range(c.nz):
        for y in range(c.ny):
            rho_inf += rho[z, y, 1]
    rho_inf /= c.ny * c.nz
    u = s.info().lattice_velocity
    n = ctypes.c_size_t()
    lib.lbm3d_links(s.lattice_handle(), None, None, None, 0, ctypes.byref(n))
    cells = np.empty(n.value, dtype=np.uint32)
    lib.lbm3d_links(s.lattice_handle(), abi.ptr(cells, ctypes.c_uint32), None, None, n.value, ctypes.byref(n))
    wall = np.unique(cells)
    want = (1.0 / 3.0) * (m.density[wall] - rho_inf) / (0.5 * u * u)
    worst = float(np.max(np.abs(r.surface.cp - want)))
    print(f"measured: wall Cp vs oracle moments: max |d| {worst:.3e} over {len(wall)} samples", flush=True)
    assert worst < 1e-12
    speed = ctypes.c_double()
    assert lib.lbm3d_max_speed(s.lattice_handle(), ctypes.byref(speed)) == 0
    fluid = s.mask().solid == 0
    ref = np.sqrt(m.velocity_x ** 2 + m.velocity_y ** 2 + m.velocity_z ** 2)[fluid].max()
    assert speed.value == pytest.approx(ref, rel=1e-14)
