"""Experiment runner, host side: the bodies of primitive records and their reference quantities
against closed forms and the silhouette bound, orientation conventions of
src/app/aero/aero_geometry.cpp, and refusals that never touch the device."""
import math

import numpy as np
import pytest

from orrery_amd import experiment as ex
from orrery_amd import runner


def silhouette_bound(mesh, d):
    """Half the sum of |n . d| A: the exact silhouette area of a closed convex mesh
    (reference_quantities.cpp:48-60)."""
    v = mesh.vertices
    t = mesh.triangles.astype(np.int64)
    c = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    return 0.25 * np.abs(c @ np.asarray(d, dtype=float)).sum()


def ref(shape, **dims):
    g = ex.GeometryRef(shape=shape)
    for k, v in dims.items():
        setattr(g.dimensions, k, v)
    return g


def test_cube_quantities_are_exact():
    p = runner.prepare_geometry(ref(ex.AeroShape.cube, side_m=2.0))
    q = p.quantities
    assert q.frontal_area == pytest.approx(4.0, rel=1e-12) and q.planform_area == pytest.approx(4.0, rel=1e-12)
    assert q.wetted_area == pytest.approx(24.0, rel=1e-12) and q.volume == pytest.approx(8.0, rel=1e-12)
    assert q.reference_length == pytest.approx(2.0, rel=1e-12)
    assert np.allclose(q.centre_of_mass, 0.0, atol=1e-12) and np.allclose(p.moment_centre, 0.0, atol=1e-12)


@pytest.mark.parametrize("shape,dims", [(ex.AeroShape.sphere, {"diameter_m": 1.3}),
                                        (ex.AeroShape.capsule, {"diameter_m": 3.9}),
                                        (ex.AeroShape.cone, {"diameter_m": 1.0, "length_m": 2.0}),
                                        (ex.AeroShape.cylinder, {"diameter_m": 0.5, "length_m": 2.0})])
def test_silhouettes_against_the_convex_bound(shape, dims):
    """Convex bodies: the raster silhouette equals the bound to within a raster cell along the
    perimeter (512 cells along the larger side)."""
    p = runner.prepare_geometry(ref(shape, **dims))
    for area, d in ((p.quantities.frontal_area, (1, 0, 0)), (p.quantities.planform_area, (0, 0, 1))):
        assert area == pytest.approx(silhouette_bound(p.mesh, d), rel=1e-2)


def _reference_quantities_cases():
    from golden.make_reference_quantities import CASES, RESOLUTIONS
    return [(key, shape, dims, res) for key, (shape, dims) in CASES.items() for res in RESOLUTIONS]


@pytest.mark.parametrize("key,shape,dims,res", _reference_quantities_cases())
def test_quantities_match_the_reference(key, shape, dims, res):
    """runner.reference_quantities against the reference's reference_quantities
    (src/geom/reference_quantities.cpp, built into oracle/_ref; the committed fixture
    tests/golden/reference_quantities.json where it is not built): the rasterised silhouettes bit
    for bit, the sums to rounding."""
    import json
    import os
    import refcheck
    mesh = runner.prepare_geometry(ref(shape, **dims)).mesh
    if refcheck.available():
        want = refcheck.reference_quantities(mesh.vertices, mesh.triangles, res)
    else:
        with open(os.path.join(os.path.dirname(__file__), "golden", "reference_quantities.json")) as f:
            want = json.load(f)[f"{key}@{res}"]
    q = runner.reference_quantities(mesh, res)
    assert q.frontal_area == want["frontal_area"] and q.planform_area == want["planform_area"]
    assert q.reference_length == want["reference_length"]
    assert q.wetted_area == pytest.approx(want["wetted_area"], rel=1e-13)
    assert q.volume == pytest.approx(want["volume"], rel=1e-13)
    size = want["reference_length"]
    for got, name in zip(q.centre_of_mass, ("centre_x", "centre_y", "centre_z")):
        assert got == pytest.approx(want[name], abs=1e-13 * size)


def test_sphere_quantities_against_the_closed_form():
    r = 0.65
    q = runner.prepare_geometry(ref(ex.AeroShape.sphere, diameter_m=2 * r)).quantities
    assert q.frontal_area == pytest.approx(math.pi * r * r, rel=1e-2)
    assert q.wetted_area == pytest.approx(4 * math.pi * r * r, rel=1e-2)
    assert q.volume == pytest.approx(4 / 3 * math.pi * r ** 3, rel=2e-2)
    assert q.reference_length == pytest.approx(2 * r, rel=1e-3)


def test_orientation_conventions():
    """Cone and capsule noses point upstream (-x), the cylinder's axis lies along y."""
    cone = runner.prepare_geometry(ref(ex.AeroShape.cone, diameter_m=1.0, length_m=2.0)).mesh
    lo, hi = cone.bounds()
    assert hi[0] - lo[0] == pytest.approx(2.0, rel=1e-12)
    apex = cone.vertices[np.argmin(cone.vertices[:, 0])]
    assert abs(apex[1]) < 1e-12 and abs(apex[2]) < 1e-12  # the apex on the axis, upstream
    cyl = runner.prepare_geometry(ref(ex.AeroShape.cylinder, diameter_m=0.5, length_m=2.0)).mesh
    lo, hi = cyl.bounds()
    assert hi[1] - lo[1] == pytest.approx(2.0, rel=1e-12) and hi[0] - lo[0] < 0.51
    cap = runner.prepare_geometry(ref(ex.AeroShape.capsule, diameter_m=3.912)).mesh
    v = cap.vertices
    # the heat shield (widest section) sits at the upstream end
    widest = v[np.argmax(np.hypot(v[:, 1], v[:, 2])), 0]
    assert widest < v[:, 0].min() + 0.3 * (v[:, 0].max() - v[:, 0].min())


def test_wing_moment_centre_is_the_quarter_chord():
    p = runner.prepare_geometry(ex.GeometryRef(shape=ex.AeroShape.naca_wing, naca="2412",
                                               dimensions=ex.ShapeDimensions(chord_m=0.8, span_m=3.0)))
    lo, hi = p.mesh.bounds()
    assert p.moment_centre[0] == pytest.approx(lo[0] + 0.2, abs=1e-12)
    assert p.quantities.reference_length == pytest.approx(0.8, rel=1e-9)


def test_reference_values_follow_the_settings():
    p = runner.prepare_geometry(ref(ex.AeroShape.cube, side_m=2.0))
    s = ex.AeroSolverSettings(reference_area=ex.ReferenceAreaKind.wetted)
    rv = runner.reference_values(p, s)
    assert rv.area == pytest.approx(24.0) and rv.length == pytest.approx(2.0)
    s.reference_length_m = 0.5
    assert runner.reference_values(p, s).length == 0.5


@pytest.mark.parametrize("solver,shape,message", [
    (ex.SolverKind.panel, ex.AeroShape.sphere, "solver 'panel' is not part of this library"),
    (None, ex.AeroShape.sphere, "solver 'auto'"),
    (ex.SolverKind.lbm, ex.AeroShape.naca_section, "lbm: the section covers no lattice node (resolution too low)"),
    (ex.SolverKind.lbm, ex.AeroShape.ellipsoid, "ellipsoid primitive is not built"),
])
def test_unrunnable_records_fail_with_a_message(solver, shape, message):
    """Refusals before the device lattice exists: the solver choice, the primitive, and the 2D
    adapter's setup of a section too coarse to cover a node (0.01 cells per chord)."""
    e = ex.AeroExperiment(name="x")
    e.geometry.shape = shape
    e.solver.solver = solver
    e.solver.lbm_cells_per_length = 0.01
    with pytest.raises(runner.RunnerError) as err:
        runner.run_experiment(e)
    assert message in str(err.value)
    r = runner.ExperimentRunner()
    try:
        r.send(runner.StartRun(e, to_convergence=True))
        p = r.wait({runner.RunState.failed}, timeout=30)
        assert message in p.status.message and p.status.run == 1
    finally:
        r.close()
