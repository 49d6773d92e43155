"""Pins the D2Q9 oracle with the reference's closed-form tests (tests/aero/lbm/test_lbm2d_flows.cpp)."""
import math

import numpy as np
import pytest

from cases import MAGIC_TAU, cylinder_channel_2d, l2_relative
from oracle import Oracle2D
from orrery_amd import (Lbm2DConfig, Smagorinsky, Wall, WallKind, XBoundary, circle_mask, fluid_mask,
                        viscosity_of_tau)


def periodic_box(nx, ny, tau):
    return Lbm2DConfig(nx=nx, ny=ny, tau=tau, x_boundary=XBoundary.periodic, bottom=Wall(WallKind.periodic),
                       top=Wall(WallKind.periodic))


def poiseuille_config(nx, ny, tau, g):
    c = periodic_box(nx, ny, tau)
    c.bottom = Wall(WallKind.no_slip)
    c.top = Wall(WallKind.no_slip)
    c.body_force_x = g
    return c


def run_to_steady(solver, chunk, limit, tol):
    """test_lbm2d_flows.cpp:24-42"""
    prev = solver.macroscopic()
    while solver.steps() < limit:
        solver.step(chunk)
        now = solver.macroscopic()
        change = np.abs(now.velocity_x - prev.velocity_x).max()
        peak = np.abs(now.velocity_x).max()
        prev = now
        if change <= tol * peak:
            break
    return prev


def profile_error(fields, exact):
    diff = 0.0
    norm = 0.0
    for y in range(fields.ny):
        e = exact(y)
        u = fields.velocity_x[fields.nx * y: fields.nx * (y + 1)]
        diff += float(np.sum((u - e) ** 2))
        norm += fields.nx * e * e
    return math.sqrt(diff / norm)


@pytest.mark.parametrize("tau,tol", [(MAGIC_TAU, 1e-9), (0.8, 2e-3), (0.6, 2e-3)])
def test_poiseuille_matches_the_parabola(tau, tol):
    c = poiseuille_config(4, 32, tau, 1e-6)
    o = Oracle2D(c, fluid_mask(4, 32))
    f = run_to_steady(o, 2000, 400000, 1e-12)
    nu = viscosity_of_tau(tau)
    exact = lambda y: c.body_force_x / (2 * nu) * (y + 0.5) * (32 - 0.5 - y)  # noqa: E731
    assert profile_error(f, exact) < tol
    assert np.abs(f.velocity_y).max() < 1e-9 * exact(15.5)


def test_couette_matches_the_linear_profile():
    c = periodic_box(4, 16, 0.9)
    c.bottom = Wall(WallKind.no_slip)
    c.top = Wall(WallKind.no_slip, 0.02)
    f = run_to_steady(Oracle2D(c, fluid_mask(4, 16)), 1000, 200000, 1e-13)
    assert profile_error(f, lambda y: 0.02 * (y + 0.5) / 16.0) < 1e-9
    c.smagorinsky = Smagorinsky(True, 0.17)
    f = run_to_steady(Oracle2D(c, fluid_mask(4, 16)), 1000, 200000, 1e-13)
    assert profile_error(f, lambda y: 0.02 * (y + 0.5) / 16.0) < 1e-7


@pytest.mark.parametrize("rho", [1.0, 2.0, 0.7])
def test_co_moving_walls(rho):
    c = periodic_box(4, 12, 0.8)
    c.bottom = Wall(WallKind.no_slip, 0.02)
    c.top = Wall(WallKind.no_slip, 0.02)
    o = Oracle2D(c, fluid_mask(4, 12))
    planes = np.zeros((3, c.cells))
    planes[0] = rho
    planes[1] = 0.02
    o.initialize(planes)
    o.step(200)
    m = o.macroscopic()
    dev = max(np.abs(m.density - rho).max() / rho, np.abs(m.velocity_x - 0.02).max() / 0.02,
              np.abs(m.velocity_y).max() / 0.02)
    assert dev < 1e-12


def test_momentum_exchange_balances_the_body_force():
    c = periodic_box(6, 24, 0.9)
    c.body_force_x = 2e-6
    mask = fluid_mask(6, 24)
    mask.solid[:6] = 1
    o = Oracle2D(c, mask)
    run_to_steady(o, 2000, 300000, 1e-12)
    f = o.force()
    assert f.x == pytest.approx(2e-6 * 6 * 23, rel=1e-6)
    assert abs(f.y) < 1e-9 * f.x


def taylor_green(n, u0, c):
    k = 2.0 * math.pi / n
    planes = np.zeros((3, c.cells))
    for y in range(c.ny):
        for x in range(c.nx):
            planes[:, x + c.nx * y] = (1.0 - 0.75 * u0 * u0 * (math.cos(2 * k * x) + math.cos(2 * k * y)),
                                       -u0 * math.cos(k * x) * math.sin(k * y), u0 * math.sin(k * x) * math.cos(k * y))
    return planes


@pytest.mark.parametrize("walls,speed,les", [(WallKind.periodic, 0.0, False), (WallKind.slip, 0.0, True),
                                             (WallKind.no_slip, 0.03, False)])
def test_mass_is_conserved(walls, speed, les):
    c = periodic_box(48, 40, 0.56)
    c.bottom = Wall(walls, 0.0)
    c.top = Wall(walls, speed)
    c.smagorinsky = Smagorinsky(les, 0.15)
    o = Oracle2D(c, circle_mask(48, 40, 20.0, 19.0, 6.0))
    o.initialize(taylor_green(40, 0.04, c))
    m0 = o.mass()
    o.step(1500)
    assert abs(o.mass() - m0) / m0 < 1e-12


def kinetic_energy(f):
    return float(np.sum(0.5 * f.density * (f.velocity_x ** 2 + f.velocity_y ** 2)))


def test_taylor_green_decays_at_the_lattice_viscosity():
    n, u0, tau = 64, 0.02, 0.65
    c = periodic_box(n, n, tau)
    o = Oracle2D(c, fluid_mask(n, n))
    o.initialize(taylor_green(n, u0, c))
    start = o.macroscopic()
    o.step(2000)
    end = o.macroscopic()
    k = 2 * math.pi / n
    rate = -math.log(kinetic_energy(end) / kinetic_energy(start)) / 2000
    assert rate == pytest.approx(4.0 * viscosity_of_tau(tau) * k * k, rel=0.01)
    les = periodic_box(n, n, tau)
    les.smagorinsky = Smagorinsky(True, 0.3)
    o2 = Oracle2D(les, fluid_mask(n, n))
    o2.initialize(taylor_green(n, u0, les))
    o2.step(2000)
    assert kinetic_energy(o2.macroscopic()) < kinetic_energy(end)


def test_rest_stays_at_rest_without_force():
    c = periodic_box(32, 24, 0.7)
    c.bottom = Wall(WallKind.no_slip)
    c.top = Wall(WallKind.slip)
    o = Oracle2D(c, circle_mask(32, 24, 15.0, 12.0, 5.0))
    o.step(200)
    f = o.force()
    assert abs(f.x) < 1e-15 and abs(f.y) < 1e-15
    m = o.macroscopic()
    assert np.sqrt(m.velocity_x ** 2 + m.velocity_y ** 2).max() < 1e-16


def test_cylinder_is_deterministic_and_drags_along_the_stream():
    c, m = cylinder_channel_2d(120, 60, 30.0, 30.5, 10.0, 0.05, 100.0)
    a = Oracle2D(c, m)
    b = Oracle2D(c, m)
    a.step(300)
    b.step(300)
    assert np.array_equal(a.populations(), b.populations())
    assert a.force() == b.force()
    assert a.force().x > 0.0
