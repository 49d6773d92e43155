"""Pins the oracle (the FP64 CPU restatement) with the reference's own closed-form tests
(tests/aero/lbm3d/test_lbm3d_flows.cpp) and with the values recorded from the reference build.

A plausible slip in the restatement fails at least one of these: a wrong sign or weight in the
equilibrium or the forcing breaks the exact Poiseuille/Couette parabolas; a wrong moving-wall
density breaks the co-moving walls; a wrong outlet or inlet rule breaks the uniform stream; a
wrong mirror table breaks mass conservation; a wrong link direction breaks the momentum balance.
"""
import math

import numpy as np
import pytest

from cases import (MAGIC_TAU, couette3d_config, duct_config, duct_velocity, periodic_box, poiseuille3d_config,
                   poiseuille3d_velocity, profile_error, sphere_channel, uniform_state3d, uniform_state_error, wall)
from oracle import Oracle3D, validate3
from orrery_amd import (Lbm3DConfig, Outlet, Smagorinsky, SolidWall, WallKind, XBoundary, box_mask, fluid_mask_3d,
                        sphere_mask)


def run_poiseuille(tau, steps):
    c = poiseuille3d_config(4, 24, 4, tau, 1e-6)
    o = Oracle3D(c, fluid_mask_3d(4, 24, 4))
    o.step(steps)
    return c, o.macroscopic()


def test_poiseuille_value_survey_recorded():
    # Survey-recorded, not a pin: a build of the reference's CPU solver whose oneTBB was replaced
    # by a serial stand-in gave, at 4x24x4, tau = 1/2 + sqrt(3/16), g = 1e-6, 2000 steps,
    # u(y = 12) = 4.9431445994122588e-4. The restatement reproduces it bit for bit; the pins of
    # this path are the closed forms below (parabola, duct series, Couette, conservation).
    _, m = run_poiseuille(MAGIC_TAU, 2000)
    u = m.velocity_x.reshape(4, 24, 4)[:, 12, :]
    assert np.all(u == float("4.9431445994122588e-4"))


def test_poiseuille_matches_the_parabola():
    # test_lbm3d_flows.cpp:57-68
    c, m = run_poiseuille(MAGIC_TAU, 14000)
    assert profile_error(m, lambda x, y, z: poiseuille3d_velocity(c, y)) < 1e-9
    c, m = run_poiseuille(0.8, 14000)
    assert profile_error(m, lambda x, y, z: poiseuille3d_velocity(c, y)) < 2e-3


def test_square_duct_matches_the_series():
    # test_lbm3d_flows.cpp:70-81
    c = duct_config(4, 20, 20, MAGIC_TAU, 1e-6)
    o = Oracle3D(c, fluid_mask_3d(4, 20, 20))
    o.step(8000)
    assert profile_error(o.macroscopic(), lambda x, y, z: duct_velocity(c, y, z)) < 1e-3


def test_couette_between_resting_and_moving_wall():
    # test_lbm3d_flows.cpp:83-93
    c = couette3d_config(4, 4, 16, 0.8, 0.02)
    o = Oracle3D(c, fluid_mask_3d(4, 4, 16))
    o.step(8000)
    assert profile_error(o.macroscopic(), lambda x, y, z: 0.02 * (z + 0.5) / 16.0) < 1e-9


@pytest.mark.parametrize("rho", [0.7, 1.0, 2.0])
def test_co_moving_walls_keep_the_uniform_state(rho):
    # test_lbm3d_flows.cpp:95-112
    c = periodic_box(4, 8, 8, 0.7)
    u = 0.03
    c.y_low = wall(WallKind.no_slip, u)
    c.y_high = wall(WallKind.no_slip, u)
    c.z_low = wall(WallKind.no_slip, u)
    c.z_high = wall(WallKind.no_slip, u)
    o = Oracle3D(c, fluid_mask_3d(4, 8, 8))
    o.initialize(uniform_state3d(c, rho, u))
    o.step(300)
    assert uniform_state_error(o.macroscopic(), rho, (u, 0.0, 0.0), u) < 1e-12


def mass_case():
    c = periodic_box(12, 10, 9, 0.62)
    c.y_low = wall(WallKind.slip)
    c.y_high = wall(WallKind.slip)
    c.z_low = wall(WallKind.no_slip, 0.02)
    c.z_high = wall(WallKind.no_slip, -0.01)
    c.smagorinsky = Smagorinsky(True, 0.15)
    mask = sphere_mask(12, 10, 9, (6.0, 4.6, 4.2), 2.4)
    planes = np.zeros((4, c.cells))
    for z in range(9):
        for y in range(10):
            for x in range(12):
                a = 2.0 * math.pi * x / 12.0
                planes[:, x + 12 * (y + 10 * z)] = (1.0 + 0.01 * math.sin(a), 0.03 * math.cos(a),
                                                    0.01 * math.sin(0.7 * y), 0.01 * math.cos(0.5 * z))
    return c, mask, planes


def test_mass_is_conserved_with_slip_and_moving_walls():
    # test_lbm3d_flows.cpp:146-168
    c, mask, planes = mass_case()
    o = Oracle3D(c, mask)
    o.initialize(planes)
    before = o.mass()
    o.step(1500)
    assert abs(o.mass() - before) / before < 1e-12


def test_momentum_exchange_balances_the_body_force():
    # test_lbm3d_flows.cpp:170-188
    c = periodic_box(4, 16, 4, 0.9)
    c.body_force_x = 2e-6
    mask = box_mask(4, 16, 4, (0.0, 0.0, 0.0), (3.0, 0.0, 3.0))
    o = Oracle3D(c, mask)
    o.step(6000)
    f = o.force()
    expected = c.body_force_x * (4 * 15 * 4)
    assert f.force[0] == pytest.approx(expected, rel=1e-6)
    assert abs(f.force[1]) < 1e-12 and abs(f.force[2]) < 1e-12


@pytest.mark.parametrize("outlet", [Outlet.convective, Outlet.extrapolation])
def test_uniform_stream_through_inlet_and_outlet(outlet):
    # test_lbm3d_flows.cpp:190-209
    c = Lbm3DConfig(nx=16, ny=6, nz=5, tau=0.7, outlet=outlet, y_low=wall(WallKind.slip), y_high=wall(WallKind.slip),
                    z_low=wall(WallKind.periodic), z_high=wall(WallKind.periodic), inlet_ux=0.05)
    o = Oracle3D(c, fluid_mask_3d(16, 6, 5))
    o.step(400)
    assert uniform_state_error(o.macroscopic(), 1.0, (0.05, 0.0, 0.0), 1.0) < 1e-13


def test_inlet_imposes_its_velocity_next_to_a_sphere():
    # test_lbm3d_flows.cpp:211-230
    c, mask = sphere_channel(40, 20, 20, 8.0, 2.0, 0.05, 20.0)
    c.inlet_uy = 0.004
    c.inlet_uz = -0.003
    o = Oracle3D(c, mask)
    o.step(600)
    m = o.macroscopic()
    face = np.array([40 * (y + 20 * z) for z in range(20) for y in range(20)])
    worst = max(np.abs(m.velocity_x[face] - 0.05).max(), np.abs(m.velocity_y[face] - 0.004).max(),
                np.abs(m.velocity_z[face] + 0.003).max())
    assert worst < 1e-13
    assert o.force().force[0] > 0.0


def test_oracle_is_deterministic_across_call_sizes():
    # test_lbm3d_flows.cpp:232-240
    c, mask = sphere_channel(32, 16, 16, 6.0, 2.0, 0.05, 50.0)
    a = Oracle3D(c, mask)
    b = Oracle3D(c, mask)
    a.step(150)
    b.step(100)
    b.step(50)
    assert np.array_equal(a.populations(), b.populations())


def fraction_of(fluid, solid, second_solid):
    # test_lbm3d_flows.cpp:249-269
    c = periodic_box(8, 8, 8, 0.7)
    c.solid_wall = SolidWall.interpolated
    mask = fluid_mask_3d(8, 8, 8)
    mask.distance = np.ones(512)
    mask.solid[mask.index(2, 4, 4)] = 1
    mask.distance[mask.index(2, 4, 4)] = solid
    mask.distance[mask.index(3, 4, 4)] = fluid
    if second_solid:
        mask.solid[mask.index(4, 4, 4)] = 1
        mask.distance[mask.index(4, 4, 4)] = -1.0
    assert validate3(c, mask) is None
    links = Oracle3D(c, mask).links()
    for cell, d, q in zip(links.cell, links.direction, links.q):
        if cell == mask.index(3, 4, 4) and d == 1:
            return q
    raise AssertionError("no link")


def test_wall_fractions_of_the_interpolated_wall():
    # test_lbm3d_flows.cpp:273-295
    inf = math.inf
    big = np.finfo(np.float64).max
    assert fraction_of(0.25, -0.75, False) == 0.25
    assert fraction_of(0.75, -0.25, False) == 0.75
    assert fraction_of(0.0, -1.0, False) == 0.0
    assert fraction_of(1.0, 0.0, False) == 1.0
    assert fraction_of(5e-324, 0.0, False) == 1.0
    assert fraction_of(0.25, -0.75, True) == 0.5
    assert fraction_of(0.0, -1.0, True) == 0.5
    assert fraction_of(inf, -1.0, False) == 0.5
    assert fraction_of(1.0, -inf, False) == 0.5
    assert fraction_of(inf, -inf, False) == 0.5
    assert fraction_of(2.0, 1.0, False) == 0.5
    assert fraction_of(-0.5, -1.0, False) == 0.5
    assert fraction_of(0.0, 0.0, False) == 0.5
    assert fraction_of(big, -big, False) == 0.5
    assert fraction_of(big, -0.5 * big, False) == pytest.approx(2.0 / 3.0, rel=1e-15)


def shear_initial(c):
    planes = np.zeros((4, c.cells))
    for z in range(c.nz):
        for y in range(c.ny):
            for x in range(c.nx):
                planes[:, x + c.nx * (y + c.ny * z)] = (1.0, 0.03 + 0.01 * math.sin(0.8 * y), 0.01 * math.cos(0.6 * z), 0.0)
    return planes


@pytest.mark.parametrize("fluid,solid", [(math.inf, -1.0), (1.0, -math.inf), (2.0, 1.0),
                                         (np.finfo(np.float64).max, -np.finfo(np.float64).max)])
def test_unusable_distances_reflect_like_the_halfway_wall(fluid, solid):
    # test_lbm3d_flows.cpp:297-344
    halfway = periodic_box(8, 8, 8, 0.7)
    interpolated = halfway.copy()
    interpolated.solid_wall = SolidWall.interpolated
    mask = box_mask(8, 8, 8, (3.0, 3.0, 3.0), (4.0, 4.0, 4.0))
    mask.distance = np.where(mask.solid != 0, solid, fluid).astype(np.float64)
    b = Oracle3D(interpolated, mask)
    assert np.all(b.links().q == 0.5)
    a = Oracle3D(halfway, mask)
    planes = shear_initial(halfway)
    a.initialize(planes)
    b.initialize(planes)
    a.step(20)
    b.step(20)
    fb = b.populations()
    assert np.all(np.isfinite(fb))
    assert np.array_equal(fb, a.populations())
    assert a.force() == b.force()


def probe_closure(config, a, b, gamma, after, node):
    """lbm3d_test_support.h:187-239"""
    from oracle import load
    lib = load()
    cells = config.cells
    nx, ny = config.nx, config.ny
    cx = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
    cy = [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1]
    cz = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1]

    def initial(x, y, z):
        u = [0.0, 0.0, 0.0]
        u[a] = gamma * ((x, y, z)[b] - 3.5)
        return u

    def stress(f):
        rho = sum(f)
        j = [sum(c[i] * f[i] for i in range(19)) for c in (cx, cy, cz)]
        u = [v / rho for v in j]
        pi = np.zeros((3, 3))
        for i in range(19):
            neq = f[i] - lib.orc3d_equilibrium(i, rho, u[0], u[1], u[2])
            c = (cx[i], cy[i], cz[i])
            for p in range(3):
                for s in range(3):
                    pi[p, s] += c[p] * c[s] * neq
        return pi, rho

    here = node[0] + nx * (node[1] + ny * node[2])
    before = []
    post = []
    for i in range(19):
        src = (node[0] - cx[i], node[1] - cy[i], node[2] - cz[i])
        u = initial(*src)
        before.append(lib.orc3d_equilibrium(i, 1.0, u[0], u[1], u[2]))
        post.append(after[i * cells + here])
    pi_before, rho = stress(before)
    pi_after, _ = stress(post)
    q2 = float(np.sum(pi_before * pi_before))
    tau = config.tau
    c2 = config.smagorinsky.constant ** 2
    measured = 1.0 / (1.0 - pi_after[a, b] / pi_before[a, b])
    predicted = 0.5 * (tau + math.sqrt(tau * tau + 18.0 * math.sqrt(2.0) * c2 * math.sqrt(q2) / rho))
    continuum = 0.5 * (tau + math.sqrt(tau * tau + 12.0 * c2 * abs(gamma)))
    return measured, predicted, continuum


def shear_planes(c, a, b, gamma):
    planes = np.zeros((4, c.cells))
    for z in range(c.nz):
        for y in range(c.ny):
            for x in range(c.nx):
                u = [0.0, 0.0, 0.0]
                u[a] = gamma * ((x, y, z)[b] - 3.5)
                planes[:, x + c.nx * (y + c.ny * z)] = (1.0, u[0], u[1], u[2])
    return planes


@pytest.mark.parametrize("a,b", [(0, 1), (0, 2), (1, 0), (1, 2), (2, 0), (2, 1)])
def test_smagorinsky_closure_relaxes_a_manufactured_shear(a, b):
    # test_lbm3d_flows.cpp:346-384
    tau, gamma = 0.52, 0.01
    for closure in (True, False):
        c = periodic_box(8, 8, 8, tau)
        c.smagorinsky = Smagorinsky(closure, 0.4)
        o = Oracle3D(c, fluid_mask_3d(8, 8, 8))
        o.initialize(shear_planes(c, a, b, gamma))
        o.step(1)
        after = o.populations()
        for k in (1, 3, 6):
            node = [2, 5, 4]
            node[b] = k
            measured, predicted, continuum = probe_closure(c, a, b, gamma, after, node)
            if not closure:
                assert measured == pytest.approx(tau, rel=1e-10)
                continue
            assert measured == pytest.approx(predicted, rel=1e-10)
            assert abs(measured - continuum) < 0.02 * (continuum - tau)
            assert measured - tau > 0.005


def test_configurations_are_validated():
    # test_lbm3d_flows.cpp:386-430
    mask = fluid_mask_3d(8, 8, 8)
    ok = periodic_box(8, 8, 8, 0.6)
    assert validate3(ok, mask) is None
    small = ok.copy(); small.nz = 2
    assert validate3(small, fluid_mask_3d(8, 8, 2))
    t = ok.copy(); t.tau = 0.5
    assert validate3(t, mask)
    inf = ok.copy(); inf.body_force_z = math.inf
    assert validate3(inf, mask)
    half = ok.copy(); half.z_low = wall(WallKind.slip)
    assert validate3(half, mask)
    fast = ok.copy(); fast.y_low = wall(WallKind.no_slip, 0.06); fast.y_high = wall(WallKind.no_slip, 0.0)
    assert "lattice Mach number" in validate3(fast, mask)
    ch = ok.copy(); ch.x_boundary = XBoundary.inflow_outflow; ch.inlet_ux = 0.05
    assert validate3(ch, mask) is None
    ch.inlet_uz = 0.04
    assert validate3(ch, mask)
    ch.inlet_uz = 0.0
    blocked = box_mask(8, 8, 8, (7.0, 3.0, 3.0), (7.0, 3.0, 3.0))
    assert validate3(ch, blocked) == "lbm3d: the inlet and outlet faces must be fluid"
    ch.inlet_ux = 0.0
    assert validate3(ch, mask) == "lbm3d: the convective outlet needs a positive inlet speed"
    ch.outlet = Outlet.extrapolation
    assert validate3(ch, mask) is None


def test_shear_wave_decays_at_the_lattice_viscosity():
    # Closed form: a shear wave u_x = A sin(k y) in a fully periodic box decays as
    # exp(-nu k^2 t) with nu = (tau - 1/2) / 3, up to the lattice's O(k^2) dispersion (0.5% here).
    c = periodic_box(8, 32, 6, 0.8)
    y = (np.arange(c.cells) // c.nx) % c.ny
    planes = np.zeros((4, c.cells))
    planes[0] = 1.0
    planes[1] = 1e-3 * np.sin(2.0 * math.pi * y / c.ny)
    o = Oracle3D(c, fluid_mask_3d(8, 32, 6))
    o.initialize(planes)
    o.step(400)
    got = np.asarray(o.macroscopic().velocity_x)[y == 8].mean()
    k = 2.0 * math.pi / c.ny
    assert got == pytest.approx(1e-3 * math.exp(-(c.tau - 0.5) / 3.0 * k * k * 400), rel=1e-2)
