"""Pins of the oracle pieces that no flow closed form reached alone:

- the interpolated (Bouzidi) reflection and the momentum-exchange force with its moment arm
  x - q c_j, against the reference's own `reflected` (src/aero/lbm3d/lbm3d_internal.h:185-197)
  and `add_link_force` / `link_wall_point` (src/aero/lbm3d/lbm3d_common.cpp:213-241), compiled
  from the reference sources into oracle/_ref, bit for bit on random states;
- the absorbing-layer relaxation f* <- f* - sigma (f* - f_eq(1, u_inlet)) (lbm3d_cpu.cpp:94-99,
  called with the current inlet velocity at :317) by its closed form on a uniform state;
- the convective outlet (prev + U up) / (1 + U) (lbm3d_cpu.cpp:242-250) by the geometric decay
  of its history when the outlet node's own populations are held fixed.
"""
import math

import numpy as np
import pytest

import refcheck
from cases import sphere_channel
from oracle import Oracle3D, load as load_oracle, ptr
from orrery_amd import Lbm3DConfig, Outlet, SolidWall, Sponge, Wall, WallKind, XBoundary, sphere_mask

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def feq(rho, u):
    """Second-order D3Q19 equilibrium (d3q19.h), per direction."""
    cu = CX * u[0] + CY * u[1] + CZ * u[2]
    uu = u[0] ** 2 + u[1] ** 2 + u[2] ** 2
    return W * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu)


def lattices_with_walls():
    """A sphere channel with an interpolated wall (q spans (0, 1), both Bouzidi branches and
    the exact 1/2 case from the distance field), an off-centre body (so the moments are not
    zero by symmetry), and the same lattice with the halfway wall."""
    c, m = sphere_channel(30, 18, 16, 8.0, 1.4, 0.05, 60.0)
    m = sphere_mask(30, 18, 16, (11.3, 8.1, 6.7), 4.2)
    c.solid_wall = SolidWall.interpolated
    halfway = c.copy()
    halfway.solid_wall = SolidWall.halfway
    return [(c, m), (halfway, m)]


@pytest.mark.skipif(not refcheck.available(), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("which", [0, 1])
def test_force_and_moment_equal_the_references_link_sum(which):
    """The oracle's force() (lbm3d_cpu.cpp:192-203 restated) equals the reference's own link
    loop over `reflected` and `add_link_force` bit for bit, on random population states."""
    import ctypes
    ref = refcheck.load()
    ref.ref3d_force_of.restype = None
    ref.ref3d_force_of.argtypes = [ctypes.POINTER(type(lattices_with_walls()[0][0].to_c())),
                                   ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double)]
    config, mask = lattices_with_walls()[which]
    cpu = Oracle3D(config, mask)
    links = cpu.links()
    assert len(links.q) > 200
    if which == 0:
        # the distance field gives both branches of the interpolation
        assert (links.q < 0.5).sum() > 20 and (links.q > 0.5).sum() > 20
    else:
        assert np.all(links.q == 0.5)
    rng = np.random.default_rng(40 + which)
    c = config.to_c()
    solid = np.ascontiguousarray(mask.solid, dtype=np.uint8)
    dist = None if mask.distance is None else np.ascontiguousarray(mask.distance)
    for trial in range(3):
        f = rng.uniform(0.0, 0.2, 19 * config.cells)
        cpu.set_populations(f)
        got = cpu.force()
        F = np.zeros(3)
        M = np.zeros(3)
        ref.ref3d_force_of(ctypes.byref(c), ptr(solid, ctypes.c_uint8), ptr(dist), ptr(f), ptr(F), ptr(M))
        assert tuple(F) == got.force, trial
        assert tuple(M) == got.moment, trial
        assert np.abs(M).max() > 1e-3  # the arm matters


@pytest.mark.skipif(not refcheck.available(), reason="oracle/_ref not built (reference tree absent)")
def test_reflected_branches_equal_the_reference():
    """The reference's `reflected` on hand-set states reproduces the three forms of the
    interpolated bounce-back (Bouzidi et al. 2001): q = 1/2 the plain reflection, q < 1/2
    2q f_o(x) + (1 - 2q) f_o(x + c_j), q > 1/2 (f_o(x) + (2q - 1) f_j(x)) / (2q)."""
    import ctypes
    ref = refcheck.load()
    ref.ref3d_reflected.restype = ctypes.c_double
    ref.ref3d_reflected.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_size_t, ctypes.c_size_t,
                                    ctypes.c_size_t, ctypes.c_size_t, ctypes.c_double]
    cells = 5
    rng = np.random.default_rng(3)
    f = rng.uniform(0.01, 0.1, 19 * cells)
    opp = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17]
    here, farther = 1, 3
    for i in range(1, 19):
        o = opp[i]
        out = f[o * cells + here]
        for q in (0.0, 0.125, 0.3, 0.5, 0.7, 1.0):
            got = ref.ref3d_reflected(ptr(f), cells, here, farther, i, q)
            if q == 0.5:
                want = out
            elif q < 0.5:
                want = 2 * q * out + (1 - 2 * q) * f[o * cells + farther]
            else:
                want = (out + (2 * q - 1) * f[i * cells + here]) / (2 * q)
            assert got == pytest.approx(want, rel=1e-15, abs=0), (i, q)


def sponge_channel(strength, inlet, outlet):
    """Inflow/outflow channel with periodic sides (no side layers, no walls)."""
    return Lbm3DConfig(nx=24, ny=3, nz=3, tau=0.8, x_boundary=XBoundary.inflow_outflow,
                       outlet=Outlet.extrapolation, y_low=Wall(WallKind.periodic), y_high=Wall(WallKind.periodic),
                       z_low=Wall(WallKind.periodic), z_high=Wall(WallKind.periodic), inlet_ux=0.05,
                       sponge=Sponge(strength, inlet, outlet, 0))


def test_sponge_relaxes_towards_the_current_inlet_equilibrium():
    """One step from a uniform equilibrium (rho 1.02, u != inlet): the interior nodes pull the
    same state from every side, collide it to itself, then relax by sigma(x) towards
    f_eq(1, u_inlet) with the inlet velocity set on the solver (not the configured one). Closed
    form f = f_eq(rho, u) - sigma(x) (f_eq(rho, u) - f_eq(1, u_inlet))."""
    from orrery_amd import fluid_mask_3d
    c = sponge_channel(0.35, 6, 9)
    mask = fluid_mask_3d(c.nx, c.ny, c.nz)
    cpu = Oracle3D(c, mask)
    rho, u = 1.02, np.array([0.031, 0.012, -0.007])
    planes = np.zeros((4, c.cells))
    planes[0] = rho
    planes[1:] = u[:, None]
    cpu.initialize(planes)
    inlet = np.array([0.04, 0.003, -0.002])
    cpu.set_inlet_velocity(inlet)
    cpu.step(1)
    f = cpu.populations().reshape(19, c.cells)
    lib = load_oracle()
    cc = c.to_c()
    start = feq(rho, u)
    target = feq(1.0, inlet)
    checked = 0
    for x in range(1, c.nx - 1):
        sigma = lib.orc3d_sponge(cc, x, 1, 1)
        if refcheck.available():
            assert sigma == refcheck.load().ref3d_sponge(cc, x, 1, 1)
        # closed form of sigma(x) = s ((T - d) / T)^2 at distance d from the nearer face
        want_sigma = 0.35 * max(((6 - x) / 6) ** 2 if x < 6 else 0.0,
                                ((9 - (c.nx - 1 - x)) / 9) ** 2 if c.nx - 1 - x < 9 else 0.0)
        assert sigma == pytest.approx(want_sigma, rel=1e-15, abs=1e-300)
        cell = x + c.nx * (1 + c.ny * 1)
        want = start - sigma * (start - target)
        assert np.max(np.abs(f[:, cell] - want)) < 2e-16, x
        checked += sigma > 0
    assert checked >= 10
    # the configured inlet speed as the target would be a different state
    wrong = start - 0.35 * (start - feq(1.0, np.array([0.05, 0.0, 0.0])))
    assert np.max(np.abs(f[:, 1 + c.nx * 4] - wrong)) > 1e-5


def test_convective_outlet_history_relaxes_geometrically():
    """With a full-strength layer at the outlet (sigma = 1 there) the outlet node leaves every
    step at f_eq(1, u_inlet), so the value its zero-gradient rule copies is fixed, and its
    convective history h <- (h + U up) / (1 + U), U the configured inlet speed, decays towards
    it by 1 / (1 + U) per step from the initial outlet state."""
    from orrery_amd import fluid_mask_3d
    c = sponge_channel(1.0, 0, 3)
    c.outlet = Outlet.convective
    c.inlet_ux = 0.055
    mask = fluid_mask_3d(c.nx, c.ny, c.nz)
    cpu = Oracle3D(c, mask)
    rho, u = 0.97, np.array([0.02, -0.01, 0.015])
    planes = np.zeros((4, c.cells))
    planes[0] = rho
    planes[1:] = u[:, None]
    cpu.initialize(planes)
    inlet = np.array([0.03, 0.0, 0.0])  # current inlet differs from the configured U
    cpu.set_inlet_velocity(inlet)
    slots = [2, 8, 10, 12, 14]
    a = feq(rho, u)[slots]
    t = feq(1.0, inlet)[slots]
    h0 = cpu.outlet_history()
    assert np.allclose(h0, a[:, None, None], rtol=0, atol=1e-17)
    U = c.inlet_ux
    for n in range(1, 9):
        cpu.step(1)
        h = cpu.outlet_history()
        # step 1 copies the initial outlet state (h stays a); from step 2 on it copies t
        want = t + (a - t) / (1.0 + U) ** (n - 1)
        assert np.max(np.abs(h - want[:, None, None])) < 1e-16 * (n + 1), n
        wrong = t + (a - t) / (1.0 + inlet[0]) ** (n - 1)  # the current inlet speed as U
        if n >= 3:
            assert np.max(np.abs(h - wrong[:, None, None])) > 1e-6
    # the outlet node itself sits at the layer's target after every step
    f = cpu.populations().reshape(19, c.cells)
    cell = c.nx - 1 + c.nx * (1 + c.ny)
    assert np.max(np.abs(f[:, cell] - feq(1.0, inlet))) < 1e-16
