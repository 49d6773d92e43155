"""Integer and setup work, bit-exact: masks, node classes, boundary links (with Bouzidi q), the
boundary-rule resolution, the inlet reconstructions, the sponge and the validation messages.

Three implementations are compared: the reference's own setup code (oracle/_ref, compiled from the
reference sources), the oracle's restatement, and the device library's host setup. Where
oracle/_ref is not built the reference side comes from the committed fixtures
(tests/golden/setup_fixtures.json, written by tests/golden/make_setup_fixtures.py from
oracle/_ref).
"""
import ctypes
import hashlib
import json
import math
import os

import numpy as np
import pytest

import refcheck
from cases import periodic_box, sphere_channel, wall
from oracle import Oracle3D, load as load_oracle
from orrery_amd import (Lbm2DConfig, Lbm3DConfig, Outlet, SolidWall, Sponge, WallKind, XBoundary, boundary_links,
                        box_mask, circle_mask, classify_nodes, fluid_mask_3d, sphere_mask, validate)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "setup_fixtures.json")
needs_ref = pytest.mark.skipif(not refcheck.available(), reason="oracle/_ref not built (reference tree absent)")


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


# ---- the workload lattices ------------------------------------------------------------------
def c1():
    """D2Q9 400x160 channel, cylinder D 32 at (80, 80), u 0.05, Re 100."""
    from cases import cylinder_channel_2d
    return cylinder_channel_2d()


def c2():
    """D3Q19 256x128x128, sphere D 32 at (64, 64, 64), u 0.05, Re 200."""
    from orrery_amd import Smagorinsky, tau_of_viscosity
    c = Lbm3DConfig(nx=256, ny=128, nz=128, tau=tau_of_viscosity(0.05 * 32 / 200),
                    x_boundary=XBoundary.inflow_outflow, y_low=wall(WallKind.slip), y_high=wall(WallKind.slip),
                    z_low=wall(WallKind.slip), z_high=wall(WallKind.slip), inlet_ux=0.05,
                    smagorinsky=Smagorinsky(True, 0.17))
    return c, sphere_mask(256, 128, 128, (64.0, 64.0, 64.0), 16.0)


def test_c1_c2_counts_match_the_reference_record():
    # Recorded from a build of the reference's own solver: C1 797 solid / 312 links / tau 0.548;
    # C2 17,077 solid / 18,282 links / tau 0.524.
    c, m = c1()
    assert m.count() == 797
    assert len(boundary_links(c, m).cell) == 312
    assert c.tau == pytest.approx(0.548, abs=1e-15)
    c, m = c2()
    assert m.count() == 17077
    assert len(boundary_links(c, m).cell) == 18282
    assert c.tau == pytest.approx(0.524, abs=1e-15)


def test_masks_match_the_fixtures():
    g = golden()["masks"]
    c, m = c1()
    assert digest(m.solid) == g["c1_circle"]
    c, m = c2()
    assert digest(m.solid, m.distance) == g["c2_sphere"]
    b = box_mask(12, 10, 9, (2.0, 3.5, 1.0), (7.0, 6.0, 4.0))
    assert digest(b.solid, b.distance) == g["box_12_10_9"]


@needs_ref
def test_masks_match_the_reference_bitwise():
    for (nx, ny, nz, centre, r) in [(256, 128, 128, (64.0, 64.0, 64.0), 16.0), (40, 20, 20, (16.0, 9.75, 9.75), 4.0),
                                    (12, 10, 9, (6.0, 4.6, 4.2), 2.4)]:
        solid, dist = refcheck.sphere_mask(nx, ny, nz, centre, r)
        m = sphere_mask(nx, ny, nz, centre, r)
        assert np.array_equal(m.solid, solid)
        assert np.array_equal(m.distance, dist)
    solid, dist = refcheck.box_mask(12, 10, 9, (2.0, 3.5, 1.0), (7.0, 6.0, 4.0))
    m = box_mask(12, 10, 9, (2.0, 3.5, 1.0), (7.0, 6.0, 4.0))
    assert np.array_equal(m.solid, solid) and np.array_equal(m.distance, dist)
    assert np.array_equal(circle_mask(400, 160, 80.0, 80.0, 16.0).solid, refcheck.circle_mask(400, 160, 80.0, 80.0, 16.0))


def setup_cases():
    """Lattices that exercise every wall kind, both x modes, slip corners and the interpolated wall."""
    cases = {}
    c, m = sphere_channel(24, 12, 10, 5.0, 2.0, 0.05, 20.0)
    cases["channel_slip"] = (c, m)
    c2_, m2 = sphere_channel(24, 12, 10, 5.0, 2.0, 0.05, 20.0)
    c2_.y_low = wall(WallKind.no_slip, 0.01)
    c2_.y_high = wall(WallKind.slip)
    c2_.z_low = wall(WallKind.periodic)
    c2_.z_high = wall(WallKind.periodic)
    c2_.solid_wall = SolidWall.interpolated
    c2_.outlet = Outlet.extrapolation
    cases["mixed_interpolated"] = (c2_, m2)
    p = periodic_box(10, 9, 8, 0.7)
    p.z_low = wall(WallKind.no_slip, 0.02)
    p.z_high = wall(WallKind.no_slip, -0.01)
    cases["periodic_walls"] = (p, box_mask(10, 9, 8, (3.0, 2.0, 2.0), (5.0, 4.0, 6.0)))
    return cases


@pytest.mark.parametrize("name", ["channel_slip", "mixed_interpolated", "periodic_walls"])
def test_links_and_classes_of_library_and_oracle_match_fixtures(name):
    c, m = setup_cases()[name]
    g = golden()["cases"][name]
    links = boundary_links(c, m)
    assert digest(links.cell, links.direction, links.q) == g["links"]
    o = Oracle3D(c, m).links()
    assert digest(o.cell, o.direction, o.q) == g["links"]
    assert digest(classify_nodes(c, m)) == g["classes"]
    assert digest(Oracle3D(c, m).classes()) == g["classes"]


def test_c2_links_match_fixture():
    c, m = c2()
    links = boundary_links(c, m)
    assert digest(links.cell, links.direction, links.q) == golden()["c2_links"]


@needs_ref
@pytest.mark.parametrize("name", ["channel_slip", "mixed_interpolated", "periodic_walls"])
def test_resolve_source_matches_the_reference_everywhere(name):
    c, m = setup_cases()[name]
    ref = refcheck.load()
    orc = load_oracle()
    cc = c.to_c()
    solid = np.ascontiguousarray(m.solid)
    sp = refcheck.ptr(solid, ctypes.c_uint8)
    cell_a, cell_b = ctypes.c_uint64(), ctypes.c_uint64()
    dir_a, dir_b = ctypes.c_int(), ctypes.c_int()
    w_a, w_b = ctypes.c_double(), ctypes.c_double()
    for z in range(c.nz):
        for y in range(c.ny):
            for x in range(c.nx):
                if solid[x + c.nx * (y + c.ny * z)]:
                    continue
                for i in range(19):
                    ka = ref.ref3d_resolve(ctypes.byref(cc), sp, x, y, z, i, ctypes.byref(cell_a), ctypes.byref(dir_a),
                                           ctypes.byref(w_a))
                    kb = orc.orc3d_resolve(ctypes.byref(cc), sp, x, y, z, i, ctypes.byref(cell_b), ctypes.byref(dir_b),
                                           ctypes.byref(w_b))
                    assert (ka, cell_a.value, dir_a.value, w_a.value) == (kb, cell_b.value, dir_b.value, w_b.value)


@needs_ref
def test_sponge_matches_the_reference():
    c, m = sphere_channel(30, 12, 10, 5.0, 2.0, 0.05, 20.0)
    c.sponge = Sponge(0.1, 5, 7, 3)
    ref = refcheck.load()
    orc = load_oracle()
    cc = c.to_c()
    for z in range(c.nz):
        for y in range(c.ny):
            for x in range(c.nx):
                assert ref.ref3d_sponge(ctypes.byref(cc), x, y, z) == orc.orc3d_sponge(ctypes.byref(cc), x, y, z)


def test_inlet_reconstructions_match_fixtures():
    g = golden()["inlet"]
    orc = load_oracle()
    rng = np.random.default_rng(5)
    for k, case in enumerate(g["d3q19"]):
        f = np.array(case["in"], dtype=np.float64)
        orc.orc3d_inlet_west(refcheck.ptr(f), *case["u"])
        assert f.tolist() == case["out"], k
    for k, case in enumerate(g["d2q9"]):
        f = np.array(case["in"], dtype=np.float64)
        orc.orc2d_zou_he_west(refcheck.ptr(f), *case["u"])
        assert f.tolist() == case["out"], k


@needs_ref
def test_inlet_reconstructions_match_the_reference_bitwise():
    ref = refcheck.load()
    orc = load_oracle()
    rng = np.random.default_rng(11)
    for _ in range(200):
        f = 1.0 / 19 + 1e-3 * rng.standard_normal(19)
        u = (0.05 * rng.random(), 0.01 * rng.standard_normal(), 0.01 * rng.standard_normal())
        a = f.copy()
        b = f.copy()
        ref.ref3d_inlet_west(refcheck.ptr(a), *u)
        orc.orc3d_inlet_west(refcheck.ptr(b), *u)
        assert np.array_equal(a, b)
        g = 1.0 / 9 + 1e-3 * rng.standard_normal(9)
        a = g.copy()
        b = g.copy()
        ref.ref2d_zou_he_west(refcheck.ptr(a), u[0], u[1])
        orc.orc2d_zou_he_west(refcheck.ptr(b), u[0], u[1])
        assert np.array_equal(a, b)
        rho = 1.0 + 0.01 * rng.standard_normal()
        for i in range(19):
            assert ref.ref3d_equilibrium(i, rho, *u) == orc.orc3d_equilibrium(i, rho, *u)


def bad_configs():
    mask = fluid_mask_3d(8, 8, 8)
    ok = periodic_box(8, 8, 8, 0.6)
    out = []
    t = ok.copy(); t.tau = 0.5; out.append((t, mask))
    t = ok.copy(); t.body_force_z = math.inf; out.append((t, mask))
    t = ok.copy(); t.z_low = wall(WallKind.slip); out.append((t, mask))
    t = ok.copy(); t.y_low = wall(WallKind.no_slip, 0.06); t.y_high = wall(WallKind.no_slip); out.append((t, mask))
    ch = ok.copy(); ch.x_boundary = XBoundary.inflow_outflow; ch.inlet_ux = 0.05; ch.inlet_uz = 0.04
    out.append((ch, mask))
    ch2 = ch.copy(); ch2.inlet_uz = 0.0
    out.append((ch2, box_mask(8, 8, 8, (7.0, 3.0, 3.0), (7.0, 3.0, 3.0))))
    ch3 = ch2.copy(); ch3.inlet_ux = 0.0; out.append((ch3, mask))
    sp = ch2.copy(); sp.sponge = Sponge(1.5, 2, 2, 2); out.append((sp, mask))
    sp = ch2.copy(); sp.sponge = Sponge(0.1, 20, 2, 2); out.append((sp, mask))
    it = ch2.copy(); it.solid_wall = SolidWall.interpolated; out.append((it, mask))
    nan = sphere_mask(8, 8, 8, (4.0, 4.0, 4.0), 1.5)
    nan.distance = nan.distance.copy(); nan.distance[0] = math.nan
    it2 = ok.copy(); it2.solid_wall = SolidWall.interpolated; out.append((it2, nan))
    return out


def test_validation_messages_match_fixtures():
    expected = golden()["validation"]
    got = [validate(c, m) for c, m in bad_configs()]
    assert got == expected
    assert all(g is not None for g in got)
    from oracle import validate3
    assert [validate3(c, m) for c, m in bad_configs()] == expected


@needs_ref
def test_validation_messages_match_the_reference():
    assert [validate(c, m) for c, m in bad_configs()] == [refcheck.validate3(c, m) for c, m in bad_configs()]


@needs_ref
def test_links_match_the_reference_live():
    for name, (c, m) in setup_cases().items():
        rc, rd, rq = refcheck.boundary_links3(c, m)
        links = boundary_links(c, m)
        assert np.array_equal(links.cell, rc) and np.array_equal(links.direction, rd)
        assert np.array_equal(links.q, rq), name
        assert np.array_equal(classify_nodes(c, m), refcheck.classify3(c, m))
    c, m = c1()
    rc, rd = refcheck.boundary_links2(c, m)
    links = boundary_links(c, m)
    assert np.array_equal(links.cell, rc) and np.array_equal(links.direction, rd)
