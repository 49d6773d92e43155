"""The 2D aerodynamic adapter's setup on the host (LbmAeroSolver2D::setup,
src/aero/lbm/lbm_aero_solver.cpp:39-198), its refusals and warnings, after the reference's
tests/aero/lbm/test_lbm_aero_solver.cpp; and the two geometry pieces it rests on, Naca4::section
and polygon_mask, after tests/aero/lbm/test_solid_mask.cpp and tests/geom/test_naca.cpp.

Pins: the sections and masks equal the reference's own functions compiled from its sources
(oracle/_ref) bit for bit; the mask areas match pi r^2 and the NACA section area (the reference's
closed form, Naca4::section_area) to the reference's tolerances; the placement is pinned by
where the body lands (leftmost node `upstream` lengths from the inlet, centred rows, a nose-up
pitch putting the trailing edge below the leading edge); the step counts and spacing by the
reference test's expectations (2000 steps, 0.002 m)."""
import math

import numpy as np
import pytest

import refcheck
from orrery_amd import geometry as g
from orrery_amd import polygon_mask
from orrery_amd.aero import AeroError, AeroGeometry, FlowConditions, LbmAeroSolver2D, LbmSolverOptions, LesMode, \
    ReferenceValues
from orrery_amd.config import Outlet, WallKind, XBoundary


def cylinder_geometry(d, n=96):
    """Circle of diameter d (m) centred at the origin, counter-clockwise."""
    a = 2.0 * math.pi * np.arange(n) / n
    section = np.stack([0.5 * d * np.cos(a), 0.5 * d * np.sin(a)], axis=1)
    return AeroGeometry(section=section, reference=ReferenceValues(area=d * 1.0, length=d))


def flow_at(reynolds, d):
    """Water-like fluid at Re = rho V d / mu."""
    c = FlowConditions(speed=0.1, density=1000.0)
    c.viscosity = c.density * c.speed * d / reynolds
    return c


def cylinder_options(**kw):
    o = LbmSolverOptions(cells_per_length=10.0, lattice_velocity=0.05, height=12.0, downstream=10.0, duration=10.0,
                         average_window=4.0, steps_per_call=1000)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def naca0012_geometry(stations=60):
    return AeroGeometry(section=g.naca_section(0.0, 0.0, 0.12, stations),
                        reference=ReferenceValues(area=1.0, length=1.0, moment_centre=(0.25, 0.0, 0.0)))


# ---- Naca4::section and polygon_mask ------------------------------------------------------------

@pytest.mark.skipif(not refcheck.available(), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("digits,n", [("0012", 60), ("2412", 48), ("4415", 100), ("0006", 3)])
def test_naca_section_is_the_references(digits, n):
    m, p, t = int(digits[0]) / 100.0, int(digits[1]) / 10.0, int(digits[2:]) / 100.0
    assert np.array_equal(g.naca_section(m, p, t, n), refcheck.naca_section(m, p, t, n))
    assert refcheck.naca_section_area(0.0, 0.0, 0.12) == pytest.approx(0.0817, rel=1e-3)  # used below


def test_naca_section_shape():
    """Counter-clockwise from the trailing edge, leading edge at index n, a symmetric section
    mirrored about the chord, its shoelace area the section's closed-form area (0.0817 t/0.12)."""
    s = g.naca_section(0.0, 0.0, 0.12, 200)
    assert s.shape == (400, 2)
    assert tuple(s[0]) == (1.0, 0.0) and tuple(s[200]) == (0.0, 0.0)
    assert np.allclose(s[1:200, 0], s[399:200:-1, 0]) and np.allclose(s[1:200, 1], -s[399:200:-1, 1])
    assert (s[1:200, 1] > 0).all()  # upper surface first: counter-clockwise
    x, z = s[:, 0], s[:, 1]
    area = 0.5 * np.sum(x * np.roll(z, -1) - np.roll(x, -1) * z)
    assert area == pytest.approx(0.0817, rel=2e-3)  # integral of 2 y_t over the chord
    with pytest.raises(g.GeometryError, match="n must be in"):
        g.naca_section(0.0, 0.0, 0.12, 2)


def place(section, le_x, le_y, chord, angle):
    """place_section (solid_mask.cpp:75-90): nose-up rotation about the leading edge."""
    c, s = math.cos(angle), math.sin(angle)
    x = section[:, 0] * chord
    z = section[:, 1] * chord
    return np.stack([le_x + x * c + z * s, le_y - x * s + z * c], axis=1)


def test_naca_mask_matches_the_section_area():
    """The reference's mask test: NACA 0012, chord 120 nodes, leading edge at (20, 40)."""
    section = g.naca_section(0.0, 0.0, 0.12, 60)
    level = polygon_mask(180, 80, place(section, 20.0, 40.0, 120.0, 0.0))
    area = 0.0817 * 120.0 * 120.0
    assert level.count() == pytest.approx(area, rel=0.03)
    at = lambda m, x, y: bool(m.solid[x + m.nx * y])  # noqa: E731
    assert at(level, 50, 40) and not at(level, 50, 50)
    pitched_pts = place(section, 20.0, 40.0, 120.0, math.radians(10.0))
    assert pitched_pts[0, 1] < 40.0 - 120.0 * math.sin(math.radians(10.0)) + 1.0  # trailing edge to -y
    pitched = polygon_mask(180, 80, pitched_pts)
    assert pitched.count() == pytest.approx(area, rel=0.03)
    assert at(pitched, 79, 30) and not at(level, 79, 30)


def test_polygon_mask_closed_forms():
    """A square with half-integer corners covers exactly its integer nodes; a 400-gon of radius
    12 about pi r^2 nodes; fewer than 3 points cover nothing."""
    sq = polygon_mask(20, 20, [(2.5, 3.5), (9.5, 3.5), (9.5, 11.5), (2.5, 11.5)])
    want = np.zeros((20, 20), dtype=np.uint8)
    want[4:12, 3:10] = 1
    assert np.array_equal(sq.solid.reshape(20, 20), want)
    a = 2.0 * math.pi * np.arange(400) / 400
    circle = polygon_mask(100, 80, np.stack([40.3 + 12 * np.cos(a), 39.7 + 12 * np.sin(a)], axis=1))
    assert circle.count() == pytest.approx(math.pi * 144.0, rel=0.02)
    assert polygon_mask(10, 10, [(0.0, 0.0), (5.0, 5.0)]).count() == 0


@pytest.mark.skipif(not refcheck.available(), reason="oracle/_ref not built (reference tree absent)")
def test_polygon_mask_is_the_references():
    rng = np.random.default_rng(7)
    section = g.naca_section(0.02, 0.4, 0.12, 48)
    cases = [place(section, 20.0, 40.0, 120.0, math.radians(a)) for a in (0.0, 4.0, -12.5)]
    for _ in range(5):  # random star polygons, with vertices on node rows and columns
        k = int(rng.integers(3, 40))
        ang = np.sort(rng.uniform(0, 2 * math.pi, k))
        r = rng.uniform(3, 25, k)
        pts = np.stack([50 + r * np.cos(ang), 30 + r * np.sin(ang)], axis=1)
        pts[::3] = np.round(pts[::3])
        cases.append(pts)
    for pts in cases:
        got = polygon_mask(180, 80, pts)
        assert np.array_equal(got.solid, refcheck.polygon_mask(180, 80, pts))


# ---- the plan ---------------------------------------------------------------------------------

def test_cylinder_plan_follows_the_reference_test():
    """test_lbm_aero_solver.cpp:45-96: Re 100, 10 cells per diameter, 2000 steps, 0.002 m."""
    geo = cylinder_geometry(0.02)
    s = LbmAeroSolver2D(cylinder_options())
    p = s.plan(geo, flow_at(100.0, 0.02))
    c = p.config
    assert p.reynolds == pytest.approx(100.0)
    assert p.dx == pytest.approx(0.002)
    assert p.total_steps == 2000
    assert p.window_samples == 80  # 4 L / U at 200 steps per L / U, every 10 steps
    assert c.tau == pytest.approx(0.5 + 3.0 * 0.05 * 10.0 / 100.0)
    assert (c.x_boundary, c.outlet) == (XBoundary.inflow_outflow, Outlet.convective)
    assert (c.bottom.kind, c.top.kind) == (WallKind.slip, WallKind.slip)
    assert c.inlet_ux == 0.05 and not c.smagorinsky.enabled
    r = s.result()
    assert r.has_warning("lbm.resolution") and not r.has_warning("lbm.les") and not r.has_warning("lbm.mach")
    # The domain: upstream + body + downstream lengths, with at most one node to spare.
    span = (8.0 + 10.0) * 0.02 + 0.02
    assert (c.nx - 1) * p.dx >= span - 1e-12 and (c.nx - 2) * p.dx < span
    assert (c.ny - 1) * p.dx >= 12.0 * 0.02 - 1e-12 and (c.ny - 2) * p.dx < 12.0 * 0.02
    # The body: 8 lengths from the inlet, centred between the walls, about pi r^2 nodes.
    m = s.mask().solid.reshape(c.ny, c.nx)
    cols = np.nonzero(m.any(axis=0))[0]
    rows = np.nonzero(m.any(axis=1))[0]
    assert cols[0] == pytest.approx(8.0 * 0.02 / p.dx, abs=1.0)
    assert rows[0] + rows[-1] == c.ny - 1
    assert np.array_equal(m, m[::-1])  # the circle's mask is symmetric about the centre row
    assert p.solid_nodes == pytest.approx(math.pi * 25.0, rel=0.1)
    assert p.origin[0] == pytest.approx(-0.01 - 8.0 * 0.02)


@pytest.mark.parametrize("aoa", [0.0, 8.0, -15.0])
def test_plan_mask_is_the_rotated_section(aoa):
    """The mask is polygon_mask of the section in the wind frame placed by the plan's origin
    (the reference's polygon_mask where oracle/_ref is built); a nose-up angle puts the trailing
    edge below the leading edge, and the moment centre (quarter chord) lands on the chord line."""
    geo = naca0012_geometry()
    cond = flow_at(200.0, 1.0)
    cond.angle_of_attack = math.radians(aoa)
    s = LbmAeroSolver2D(LbmSolverOptions(cells_per_length=16.0, upstream=4.0, downstream=6.0, height=8.0,
                                         duration=12.0, average_window=4.0, steps_per_call=4000))
    p = s.plan(geo, cond)
    c = p.config
    ca, sa = math.cos(cond.angle_of_attack), math.sin(cond.angle_of_attack)
    x, z = geo.section[:, 0], geo.section[:, 1]
    wind = np.stack([(x * ca + z * sa - p.origin[0]) / p.dx, (-x * sa + z * ca - p.origin[1]) / p.dx], axis=1)
    mask = s.mask()
    assert np.array_equal(mask.solid, polygon_mask(c.nx, c.ny, wind).solid)
    if refcheck.available():
        assert np.array_equal(mask.solid, refcheck.polygon_mask(c.nx, c.ny, wind))
    te, le = wind[0], wind[60]
    assert te[1] - le[1] == pytest.approx(-math.sin(cond.angle_of_attack) * 16.0, abs=1e-9)
    assert te[0] - le[0] == pytest.approx(math.cos(cond.angle_of_attack) * 16.0, abs=1e-9)
    centre = le + 0.25 * (te - le)
    assert p.centre == pytest.approx(tuple(centre), abs=1e-9)
    assert mask.count() == pytest.approx(0.0817 * 256.0, abs=6.0)  # a 2-node-thick staircase


def test_warnings_outside_the_envelope():
    """test_lbm_aero_solver.cpp:130-162."""
    geo = cylinder_geometry(1.0)
    opt = LbmSolverOptions(cells_per_length=12.0, height=8.0, downstream=4.0, upstream=2.0)
    les = LbmAeroSolver2D(opt)
    p = les.plan(geo, flow_at(5e4, 1.0))
    r = les.result()
    assert r.has_warning("lbm.les") and r.has_warning("lbm.tau") and p.config.smagorinsky.enabled
    opt.les = LesMode.off
    bare = LbmAeroSolver2D(opt)
    p = bare.plan(geo, flow_at(5e4, 1.0))
    assert bare.result().has_warning("lbm.les_off") and not p.config.smagorinsky.enabled
    fast = FlowConditions(speed=150.0, sideslip=0.1)
    compressible = LbmAeroSolver2D(opt)
    compressible.plan(geo, fast)
    r = compressible.result()
    assert r.has_warning("lbm.mach") and r.has_warning("lbm.sideslip")
    assert [w.message for w in r.warnings if w.code == "lbm.sideslip"] == \
        ["the 2D solver ignores the sideslip angle 0.1 rad"]
    # A section filling 1/8 of the channel: blockage above 0.1.
    tall = LbmAeroSolver2D(LbmSolverOptions(cells_per_length=12.0, height=6.0, downstream=4.0, upstream=2.0))
    pt = tall.plan(cylinder_geometry(1.0), flow_at(100.0, 1.0))
    assert pt.blockage == pytest.approx(1.0 / ((pt.config.ny - 1) * pt.dx))
    assert tall.result().has_warning("lbm.blockage")


@pytest.mark.parametrize("what,message", [
    ("empty", "lbm: the 2D solver needs a section of at least 3 points"),
    ("nan", "lbm: section points must be finite"),
    ("still", "lbm: speed, density and viscosity must be positive and finite, angles finite"),
    ("length", "lbm: reference length and area must be positive and finite"),
    ("mach", "lattice units: lattice Mach number 0.138564 (u_lat 0.08) is not below 0.1; lower the lattice "
             "velocity (weak compressibility, APPROXIMATION A3)"),
    ("duration", "lbm: duration and averaging window must stay below 9.0072e+15 lattice steps"),
    ("window", "lbm: duration and averaging window must stay below 9.0072e+15 lattice steps"),
    ("calls", "lbm: domain, duration and step options must be positive"),
    ("coarse", "lbm: the section covers no lattice node (resolution too low)"),
])
def test_unusable_problems_are_refused(what, message):
    """test_lbm_aero_solver.cpp:164-201, with the reference's messages."""
    geo = cylinder_geometry(1.0)
    cond = flow_at(100.0, 1.0)
    opt = LbmSolverOptions(cells_per_length=10.0, height=8.0, downstream=4.0, upstream=2.0)
    if what == "empty":
        geo = AeroGeometry(section=np.zeros((0, 2)), reference=geo.reference)
    elif what == "nan":
        geo.section[3, 1] = float("nan")
    elif what == "still":
        cond.speed = 0.0
    elif what == "length":
        geo.reference.length = 0.0
    elif what == "mach":
        opt.lattice_velocity = 0.08
    elif what == "duration":
        opt.duration = 1e100
    elif what == "window":
        opt.average_window = 1e100
    elif what == "calls":
        opt.steps_per_call = 0
    elif what == "coarse":
        opt.cells_per_length = 0.01
    s = LbmAeroSolver2D(opt)
    with pytest.raises(AeroError) as err:
        s.plan(geo, cond)
    assert str(err.value) == message
    with pytest.raises(AeroError):
        s.info()  # no lattice after a refusal
    with pytest.raises(RuntimeError, match="before setup"):
        s.step()
