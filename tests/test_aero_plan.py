"""The aerodynamic adapter's setup on the host (LbmAeroSolver3D::setup,
src/aero/lbm3d/lbm_aero_solver3d.cpp:85-339) and its refusals and warnings, after the
reference's tests/aero/lbm3d/test_lbm_aero_solver3d.cpp.

Pins: the reference's own run of the coarse sphere (d = 0.02 m, Re 20, 8 cells per diameter,
2 diameters ahead, 3 behind, 3 wide, 2 convective times averaged over the last one) set up a
49 x 25 x 25 lattice with 270 solid nodes and 1164 boundary links, ran 320 steps and reported
Cd = 3.5172643919827955. The plan is checked against those counts, and the FP64 oracle run on
the plan with the adapter's schedule (gradual start, outlet sponge, convective outlet,
interpolated wall, samples every 10 steps) against that Cd; the mask also against the
reference's voxeliser (oracle/_ref)."""
import math

import numpy as np
import pytest

import refcheck
from oracle import Oracle3D
from orrery_amd import geometry as g
from orrery_amd.aero import (AeroError, AeroGeometry, FlowConditions, LbmAeroSolver3D, LbmSolver3DOptions, LesMode,
                             ReferenceValues)
from orrery_amd.config import SolidWall, WallKind
from orrery_amd.solver import boundary_links

# Survey-recorded, not a pin: the adapter's coarse sphere (49x25x25, 320 steps) run through a
# build of the reference whose oneTBB was replaced by a serial stand-in. The pins of the pieces
# this run combines (inlet, convective outlet, absorbing layer, interpolated wall, force and
# moment) are in tests/test_oracle_pins.py and tests/test_setup_pins.py.
REFERENCE_CD = 3.5172643919827955


def sphere_geometry(d):
    return AeroGeometry(g.sphere_mesh(0.5 * d, g.RevolutionResolution(24, 0.15)),
                        ReferenceValues(area=0.25 * math.pi * d * d, length=d))


def flow_at(reynolds, d):
    c = FlowConditions(speed=0.1, density=1000.0)
    c.viscosity = c.density * c.speed * d / reynolds
    return c


def coarse_options(**kw):
    o = LbmSolver3DOptions(cells_per_length=8.0, lattice_velocity=0.05, upstream=2.0, downstream=3.0, width=3.0,
                           duration=6.0, average_window=2.0, steps_per_call=1000)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def test_coarse_sphere_plan_is_the_references():
    s = LbmAeroSolver3D(coarse_options(duration=2.0))
    p = s.plan(sphere_geometry(0.02), flow_at(20.0, 0.02))
    c = p.config
    assert (c.nx, c.ny, c.nz) == (49, 25, 25)
    assert p.solid_nodes == 270
    assert p.total_steps == 320
    assert p.reynolds == pytest.approx(20.0)
    assert c.tau == pytest.approx(0.56)
    assert c.solid_wall == SolidWall.interpolated and not c.smagorinsky.enabled
    assert (c.sponge.strength, c.sponge.inlet, c.sponge.outlet, c.sponge.sides) == (0.1, 0, 12, 0)
    assert len(boundary_links(c, s.mask()).cell) == 1164
    r = s.result()
    assert r.has_warning("lbm.resolution") and r.has_warning("lbm.blockage")
    assert not r.has_warning("lbm.les") and not r.has_warning("lbm.mach")


@pytest.mark.skipif(not refcheck.available(), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("angles", [(0.0, 0.0), (30.0, 20.0)])
def test_plan_mask_is_the_reference_voxeliser_placed(angles):
    """The body box of the plan, voxelised by the reference's voxelize on the same wind-frame
    mesh, is the plan's mask; the distances are the reference's band divided by the spacing."""
    cond = flow_at(20.0, 0.02)
    cond.angle_of_attack, cond.sideslip = (math.radians(a) for a in angles)
    geo = sphere_geometry(0.02)
    s = LbmAeroSolver3D(coarse_options())
    p = s.plan(geo, cond)
    m = s.mask()
    axes = p.axes
    v = geo.mesh.vertices
    wind = np.empty_like(v)
    for a in range(3):
        wind[:, a] = v[:, 0] * axes[a, 0] + v[:, 1] * axes[a, 1] + v[:, 2] * axes[a, 2]
    dx = p.dx
    origin = [p.origin[a] + (p.first[a] - 0.5) * dx for a in range(3)]
    n = p.count
    solid, _ = refcheck.voxelize(wind, geo.mesh.triangles, origin, dx, *n)
    dist = refcheck.signed_distance_band(wind, geo.mesh.triangles, origin, dx, *n, solid)
    c = p.config
    box = np.s_[p.first[2]:p.first[2] + n[2], p.first[1]:p.first[1] + n[1], p.first[0]:p.first[0] + n[0]]
    want = np.zeros((c.nz, c.ny, c.nx), dtype=np.uint8)
    want[box] = solid.reshape(n[2], n[1], n[0])
    assert np.array_equal(m.solid.reshape(c.nz, c.ny, c.nx), want)
    far = np.full((c.nz, c.ny, c.nx), 1e9)
    far[box] = dist.reshape(n[2], n[1], n[0]) / dx
    assert np.array_equal(m.distance.reshape(c.nz, c.ny, c.nx), far)


def oracle_run(plan, mask, options, conditions, reference):
    """The adapter's schedule on the FP64 oracle: rest start, the inlet ramped per sampling
    interval, forces sampled every sample_every steps, Cd of the window mean."""
    o = Oracle3D(plan.config, mask)
    cells = plan.config.cells
    if plan.ramp_steps > 0.0:
        rest = np.zeros((4, cells))
        rest[0] = 1.0
        o.initialize(rest)
        o.set_inlet_velocity((0.0, 0.0, 0.0))
    every = options.sample_every
    u = plan.config.inlet_ux
    samples = []
    while o.steps() < plan.total_steps:
        now = o.steps()
        if plan.ramp_steps > 0.0:
            start = float(now - now % every)
            frac = 1.0 if start >= plan.ramp_steps else 0.5 * (1.0 - math.cos(math.pi * start / plan.ramp_steps))
            o.set_inlet_velocity((frac * u, 0.0, 0.0))
        o.step(min(plan.total_steps, now - now % every + every) - now)
        if o.steps() % every == 0:
            samples.append(o.force().force[0])
    n = min(plan.window_samples, len(samples))
    mean = 0.0
    for f in samples[len(samples) - n:]:
        mean += f
    mean /= n
    fu = conditions.density * plan.dx ** 4 / plan.dt ** 2
    return mean * fu / (conditions.dynamic_pressure() * reference.area)


def test_oracle_on_the_plan_reproduces_the_reference_cd():
    geo = sphere_geometry(0.02)
    cond = flow_at(20.0, 0.02)
    opt = coarse_options(duration=2.0, average_window=1.0)
    s = LbmAeroSolver3D(opt)
    plan = s.plan(geo, cond)
    cd = oracle_run(plan, s.mask(), opt, cond, geo.reference)
    print(f"measured: oracle Cd {cd!r} reference {REFERENCE_CD!r}")
    assert plan.window_samples == 16
    assert abs(cd / REFERENCE_CD - 1.0) < 1e-14


def test_warnings_outside_the_envelope():
    geo = sphere_geometry(1.0)
    opt = coarse_options(width=6.0)
    les = LbmAeroSolver3D(opt)
    les.plan(geo, flow_at(5000.0, 1.0))
    r = les.result()
    assert r.has_warning("lbm.les") and r.has_warning("lbm.tau") and not r.has_warning("lbm.blockage")
    assert les.info().config.smagorinsky.enabled
    bare = LbmAeroSolver3D(coarse_options(width=6.0, les=LesMode.off))
    bare.plan(geo, flow_at(5000.0, 1.0))
    assert bare.result().has_warning("lbm.les_off") and not bare.info().config.smagorinsky.enabled
    fast = LbmAeroSolver3D(coarse_options(width=6.0, les=LesMode.off))
    fast.plan(geo, FlowConditions(speed=150.0))
    assert fast.result().has_warning("lbm.mach") and fast.result().has_warning("lbm.reynolds")
    plate = AeroGeometry(g.box_mesh(0.1, 1.0, 1.0), ReferenceValues(area=1.0, length=1.0))
    fine = coarse_options(width=3.0, cells_per_length=16.0)
    thin = LbmAeroSolver3D(fine)
    thin.plan(plate, flow_at(20.0, 1.0))
    assert thin.result().has_warning("lbm.resolution")
    sphere = LbmAeroSolver3D(fine)
    sphere.plan(geo, flow_at(20.0, 1.0))
    assert not sphere.result().has_warning("lbm.resolution")
    w = [x for x in les.result().warnings if x.code == "lbm.les"][0]
    assert w.message == "Smagorinsky subgrid closure active at Re 5000 (APPROXIMATION A2: eddy viscosity, not resolved turbulence)"


def test_unusable_problems_are_refused():
    s = LbmAeroSolver3D(coarse_options())
    geo = sphere_geometry(1.0)
    still = flow_at(100.0, 1.0)
    still.speed = 0.0
    with pytest.raises(AeroError, match="speed, density and viscosity"):
        s.plan(geo, still)
    with pytest.raises(AeroError, match="reference length and area"):
        s.plan(AeroGeometry(geo.mesh, ReferenceValues(area=0.0, length=1.0)), flow_at(100.0, 1.0))
    with pytest.raises(AeroError, match="free slip or periodic"):
        LbmAeroSolver3D(coarse_options(sides=WallKind.no_slip)).plan(geo, flow_at(100.0, 1.0))
    with pytest.raises(AeroError, match="lattice Mach number"):
        LbmAeroSolver3D(coarse_options(lattice_velocity=0.06)).plan(geo, flow_at(100.0, 1.0))
    sliver = AeroGeometry(g.box_mesh(0.01, 0.5, 0.5), ReferenceValues(area=1.0, length=1.0))
    with pytest.raises(AeroError, match="2 x 6 x 6 needs at least 3 nodes"):
        LbmAeroSolver3D(coarse_options(cells_per_length=1.0, upstream=0.01, downstream=0.01)).plan(
            sliver, flow_at(20.0, 1.0))
    with pytest.raises(AeroError) as e:
        LbmAeroSolver3D(coarse_options(max_cells=20000)).plan(geo, flow_at(20.0, 1.0))
    assert " x 25 x 25" in str(e.value) and "max_cells" in str(e.value) and "cells per reference length" in str(e.value)
    with pytest.raises(AeroError, match="lattice steps"):
        LbmAeroSolver3D(coarse_options(duration=1e100)).plan(geo, flow_at(20.0, 1.0))
    with pytest.raises(AeroError, match="lattice steps"):
        LbmAeroSolver3D(coarse_options(average_window=1e100)).plan(geo, flow_at(20.0, 1.0))
    with pytest.raises(RuntimeError, match="before setup"):
        LbmAeroSolver3D(coarse_options()).step()


def test_angles_rotate_the_wind_frame():
    cond = flow_at(20.0, 0.02)
    cond.angle_of_attack = math.radians(30.0)
    cond.sideslip = math.radians(20.0)
    s = LbmAeroSolver3D(coarse_options(duration=3.0))
    p = s.plan(sphere_geometry(0.02), cond)
    assert np.allclose(p.axes[0], cond.drag_direction(), atol=0, rtol=0)
    assert np.allclose(p.axes @ p.axes.T, np.eye(3), atol=1e-15)
    assert np.allclose(np.cross(p.axes[0], p.axes[1]), p.axes[2], atol=1e-15)
