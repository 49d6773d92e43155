"""Body geometry: primitive tessellations, winding-number voxelisation and the signed-distance band
bit-identical to the reference (src/geom/primitives.cpp, naca.cpp, voxel_grid.cpp,
winding_number.cpp), checked against oracle/_ref where it is built and against the committed
digests in tests/golden/geometry_fixtures.json everywhere; and independently of the reference
against what the mathematics fixes (tests/geom/test_voxel.cpp:26-128 and test_primitives.cpp of
the reference): direct-sum winding numbers, exact prism volumes, closed oriented surfaces,
sphere volume and distance."""
import json
import math
import os

import numpy as np
import pytest

import refcheck
from orrery_amd import geometry as g
from orrery_amd.masks import fluid_mask_3d
from test_setup_pins import digest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "geometry_fixtures.json")
needs_ref = pytest.mark.skipif(not refcheck.available(), reason="oracle/_ref not built (reference tree absent)")


def transform(v, r, t=(0.0, 0.0, 0.0)):
    """v <- R v + t, one product at a time (the library's order, no fused multiply-add)."""
    r = np.asarray(r)
    out = np.empty_like(v)
    for a in range(3):
        out[:, a] = r[a, 0] * v[:, 0] + r[a, 1] * v[:, 1] + r[a, 2] * v[:, 2] + t[a]
    return out


def mesh_cases():
    """name -> (ours, reference-binding call)."""
    cap = (3.912, 4.694, 0.196, 0.5759586531581288, 0.7)
    return {
        "sphere_r10": (lambda: g.sphere_mesh(10.0), lambda: refcheck.revolution_mesh(0, 10.0)),
        "sphere_r16_64": (lambda: g.sphere_mesh(16.0, g.RevolutionResolution(64, 0.05)),
                          lambda: refcheck.revolution_mesh(0, 16.0, around=64, arc_step=0.05)),
        "cylinder": (lambda: g.cylinder_mesh(3.0, 7.0, g.RevolutionResolution(48, 0.1)),
                     lambda: refcheck.revolution_mesh(1, 3.0, 7.0, around=48, arc_step=0.1)),
        "cone": (lambda: g.cone_mesh(2.0, 5.0), lambda: refcheck.revolution_mesh(2, 2.0, 5.0)),
        "capsule_apollo": (lambda: g.capsule_mesh(), lambda: refcheck.revolution_mesh(3, *cap)),
        "capsule_coarse": (lambda: g.capsule_mesh(resolution=g.RevolutionResolution(32, 0.2)),
                           lambda: refcheck.revolution_mesh(3, *cap, around=32, arc_step=0.2)),
        "wedge": (lambda: g.wedge_mesh(2.0, 0.2, 1.5), lambda: refcheck.wedge_mesh(2.0, 0.2, 1.5)),
        "naca0012": (lambda: g.naca_wing(0.0, 0.0, 0.12, 1.0, 4.0, 40, 8),
                     lambda: refcheck.naca_wing(0.0, 0.0, 0.12, 1.0, 4.0, 40, 8)),
        "naca2412": (lambda: g.naca_wing(0.02, 0.4, 0.12, 1.0, 2.0, 33, 3),
                     lambda: refcheck.naca_wing(0.02, 0.4, 0.12, 1.0, 2.0, 33, 3)),
    }


def voxel_cases():
    """name -> (mesh vertices and triangles as placed, grid spec). The bodies are the served
    shapes (sphere, cone, cylinder, capsule, wedge) and the wing, pitched as a request would."""
    out = {}

    def around(name, mesh, angle, cells, pad=2):
        v = transform(mesh.vertices, g.rotation_y(math.radians(angle)))
        lo, hi = v.min(axis=0), v.max(axis=0)
        h = float((hi - lo).max()) / cells
        n = [int(math.ceil((hi[a] - lo[a]) / h)) + 2 * pad for a in range(3)]
        # A quarter cell off the lattice planes, as the adapter places bodies.
        origin = tuple(float(lo[a] - (pad + 0.25) * h) for a in range(3))
        out[name] = (v, mesh.triangles, g.GridSpec(origin, h, *n))

    around("capsule_pitch20", g.capsule_mesh(), 20.0, 40)
    around("sphere_r10", g.sphere_mesh(10.0), 0.0, 24)
    around("cone_pitch10", g.cone_mesh(2.0, 5.0), 10.0, 30)
    around("cylinder_pitch30", g.cylinder_mesh(3.0, 7.0, g.RevolutionResolution(48, 0.1)), 30.0, 30)
    around("wedge_pitch5", g.wedge_mesh(2.0, 0.2, 1.5), 5.0, 36)
    around("naca0012_aoa8", g.naca_wing(0.0, 0.0, 0.12, 1.0, 4.0, 40, 8), -8.0, 48)
    return out


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def voxel_digests(v, t, spec):
    m = g.TriangleMesh(v, t)
    vg = g.voxelize(m, spec)
    sdf = g.signed_distance_band(m, vg, 3.0)
    return vg, sdf, {"solid": int(vg.solid.sum()), "mask": digest(vg.solid), "winding": digest(vg.winding),
                     "distance": digest(sdf.distance), "in_band": int(sdf.in_band.sum())}


# ---- against the reference -----------------------------------------------------------------

@pytest.mark.parametrize("name", list(mesh_cases()))
def test_mesh_matches_the_golden_digest(name):
    m = mesh_cases()[name][0]()
    want = golden()["meshes"][name]
    assert [len(m.vertices), len(m.triangles)] == want["size"]
    assert digest(m.vertices, m.triangles) == want["digest"]


@needs_ref
@pytest.mark.parametrize("name", list(mesh_cases()))
def test_mesh_is_bitwise_the_reference(name):
    ours, ref = mesh_cases()[name]
    m = ours()
    v, t = ref()
    assert np.array_equal(m.vertices, v) and np.array_equal(m.triangles, t)


@pytest.mark.parametrize("name", list(voxel_cases()))
def test_voxels_and_distance_match_the_golden_digests(name):
    v, t, spec = voxel_cases()[name]
    _, _, d = voxel_digests(v, t, spec)
    want = golden()["voxels"][name]
    assert [spec.nx, spec.ny, spec.nz] == want["grid"]
    assert d == want["digests"]


@needs_ref
@pytest.mark.parametrize("name", ["capsule_pitch20", "naca0012_aoa8", "cone_pitch10"])
def test_voxels_and_distance_are_bitwise_the_reference(name):
    v, t, spec = voxel_cases()[name]
    vg, sdf, _ = voxel_digests(v, t, spec)
    solid, w = refcheck.voxelize(v, t, spec.origin, spec.spacing, spec.nx, spec.ny, spec.nz)
    assert np.array_equal(vg.solid, solid) and np.array_equal(vg.winding, w)
    d = refcheck.signed_distance_band(v, t, spec.origin, spec.spacing, spec.nx, spec.ny, spec.nz, solid)
    assert np.array_equal(sdf.distance, d)


def test_transform_is_the_plain_product():
    m = g.capsule_mesh()
    r = g.rotation_y(math.radians(20.0)) @ g.rotation_z(math.radians(7.0))
    got = m.transformed(r, pivot=(0.0, 0.0, 0.0), translation=(1.5, -2.0, 0.25))
    assert np.array_equal(got.vertices, transform(m.vertices, r, (1.5, -2.0, 0.25)))
    assert np.array_equal(got.triangles, m.triangles)


# ---- against the mathematics ---------------------------------------------------------------

def direct_winding(v, t, p):
    """Sum of signed solid angles / 4 pi (Van Oosterom and Strackee 1983), no tree."""
    a = v[t[:, 0]][None] - p[:, None]
    b = v[t[:, 1]][None] - p[:, None]
    c = v[t[:, 2]][None] - p[:, None]
    la, lb, lc = (np.linalg.norm(x, axis=2) for x in (a, b, c))
    det = np.einsum("pti,pti->pt", a, np.cross(b, c))
    den = (la * lb * lc + np.einsum("pti,pti->pt", a, b) * lc + np.einsum("pti,pti->pt", b, c) * la
           + np.einsum("pti,pti->pt", c, a) * lb)
    return (2.0 * np.arctan2(det, den)).sum(axis=1) / (4.0 * math.pi)


def test_winding_number_tree_against_the_direct_sum():
    m = g.sphere_mesh(1.0, g.RevolutionResolution(24, 0.2))
    rng = np.random.default_rng(7)
    p = rng.uniform(-1.6, 1.6, size=(400, 3))
    direct = direct_winding(m.vertices, m.triangles.astype(np.int64), p)
    tree = m.winding_numbers(p)
    r = np.linalg.norm(p, axis=1)
    assert np.abs(direct[r < 0.8] - 1.0).max() < 1e-12
    assert np.abs(direct[r > 1.2]).max() < 1e-12
    assert np.abs(tree - direct).max() < 0.05
    assert ((tree >= 0.5) != (direct >= 0.5)).sum() == 0


@pytest.mark.parametrize("name", list(mesh_cases()))
def test_meshes_are_closed_and_outward(name):
    m = mesh_cases()[name][0]()
    t = m.triangles.astype(np.int64)
    edges = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    directed = {tuple(e) for e in edges.tolist()}
    assert len(directed) == len(edges)  # no directed edge twice
    assert all((b, a) in directed for a, b in directed)  # every edge met once each way
    assert m.volume() > 0.0


def test_prism_volumes_are_exact():
    n, r, h = 48, 3.0, 7.0
    assert g.cylinder_mesh(r, h, g.RevolutionResolution(n, 0.1)).volume() == pytest.approx(
        0.5 * n * r * r * math.sin(2.0 * math.pi / n) * h, rel=1e-13)
    length, half, span = 2.0, 0.2, 1.5
    assert g.wedge_mesh(length, half, span).volume() == pytest.approx(length * length * math.tan(half) * span,
                                                                      rel=1e-13)
    assert g.box_mesh(1.0, 2.0, 3.0).volume() == pytest.approx(6.0, rel=1e-14)
    # Symmetric section with the closed trailing edge (last coefficient -0.1036): the area is
    # 10 t (2/3 0.2969 - 0.1260/2 - 0.3516/3 + 0.2843/4 - 0.1036/5) c^2, to within the tessellation.
    area = 10.0 * 0.12 * (0.2969 * 2.0 / 3.0 - 0.1260 / 2.0 - 0.3516 / 3.0 + 0.2843 / 4.0 - 0.1036 / 5.0)
    assert g.naca_wing(0.0, 0.0, 0.12, 1.0, 4.0, 200, 4).volume() == pytest.approx(area * 4.0, rel=1e-4)


def test_sphere_voxels_and_distance_against_the_closed_form():
    radius = 10.0
    m = g.sphere_mesh(radius)
    h = 0.8
    n = 30
    spec = g.GridSpec((-0.5 * n * h + 0.13, -0.5 * n * h - 0.07, -0.5 * n * h + 0.21), h, n, n, n)
    vg = g.voxelize(m, spec)
    idx = np.arange(spec.cells)
    i, j, k = idx % n, (idx // n) % n, idx // (n * n)
    c = np.stack([spec.origin[0] + (i + 0.5) * h, spec.origin[1] + (j + 0.5) * h, spec.origin[2] + (k + 0.5) * h], 1)
    rr = np.linalg.norm(c, axis=1)
    assert abs(vg.solid.sum() * h ** 3 / (4.0 / 3.0 * math.pi * radius ** 3) - 1.0) < 0.02
    assert vg.solid[rr < radius - 0.05].all() and not vg.solid[rr > radius + 0.05].any()
    sdf = g.signed_distance_band(m, vg, 3.0)
    band = sdf.in_band != 0
    assert np.array_equal(band, np.abs(sdf.distance) < 3.0 * h)
    # The tessellation is convex, inside the sphere and around the ball of radius h (its nearest
    # facet plane), so the distance to it is within r - h of the distance to the sphere.
    v = m.vertices
    t = m.triangles.astype(np.int64)
    n_ = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    plane = np.abs(np.einsum("ij,ij->i", n_ / np.linalg.norm(n_, axis=1)[:, None], v[t[:, 0]]))
    gap = radius - plane.min()
    assert 0.0 < gap < 0.02
    err = sdf.distance[band] - (rr[band] - radius)
    assert np.abs(err).max() <= gap + 1e-12
    assert (np.sign(err[sdf.distance[band] > 0]) >= 0).all()  # outside: never nearer than the sphere
    assert ((sdf.distance < 0) == (vg.solid != 0)).all()


def test_placement_into_the_lattice():
    mask = fluid_mask_3d(12, 10, 8)
    spec = g.GridSpec((0.0, 0.0, 0.0), 0.5, 4, 3, 2)
    solid = np.zeros(24, dtype=np.uint8)
    solid[[0, 5, 23]] = 1
    vg = g.VoxelGrid(spec, solid, np.zeros(24))
    g.place_voxels(mask, vg, (3, 2, 1))
    want = {(3, 2, 1), (4, 3, 1), (6, 4, 2)}
    got = {(x, y, z) for z in range(8) for y in range(10) for x in range(12) if mask.at(x, y, z)}
    assert got == want
    sdf = g.SignedDistanceField(spec, 1.5, np.arange(24, dtype=np.float64))
    g.place_distance(mask, sdf, (3, 2, 1))
    assert mask.distance[mask.index(3, 2, 1)] == 0.0 and mask.distance[mask.index(6, 4, 2)] == 46.0
    assert mask.distance[mask.index(0, 0, 0)] == g.FAR_DISTANCE
    with pytest.raises(g.GeometryError, match="does not fit"):
        g.place_voxels(mask, vg, (9, 2, 1))


def test_refusals():
    with pytest.raises(g.GeometryError, match="nose radius"):
        g.capsule_mesh(g.CapsuleShape(nose_radius=1.0))
    with pytest.raises(g.GeometryError):
        g.wedge_mesh(1.0, 2.0, 1.0)
    with pytest.raises(g.GeometryError):
        g.voxelize(g.box_mesh(1, 1, 1), g.GridSpec((0, 0, 0), -1.0, 2, 2, 2))
