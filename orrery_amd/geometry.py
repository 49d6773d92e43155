"""Body geometry for the lattice: triangle meshes of the reference's primitives, winding-number
voxelisation and the signed-distance band, and their placement into a lattice mask.

The arithmetic runs in the library (lbm_mesh_* / lbm_voxelize / lbm_signed_distance_band in
include/orrery_lbm.h) and reproduces the reference's masks and distances bit for bit:
primitives (src/geom/primitives.cpp:292-426,557-604), naca_wing (src/geom/naca.cpp:184-257),
voxelize / signed_distance_band (src/geom/voxel_grid.cpp:81-138) over MeshTree
(src/geom/winding_number.cpp:155-307). The placement helpers follow place_voxels /
place_distance (src/aero/lbm3d/solid_mask3d.cpp:107-163).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .masks import FAR_DISTANCE, SolidMask3D


class GeometryError(ValueError):
    pass


def _lib():
    return abi.load()


def _check(lib, status: int) -> None:
    if status != abi.LBM_OK:
        raise GeometryError(lib.lbm_geom_last_error().decode())


@dataclass
class TriangleMesh:
    """Vertices [nv, 3] (float64) and triangles [nt, 3] (uint32), as geom::TriangleMesh."""

    vertices: np.ndarray
    triangles: np.ndarray

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, dtype=np.float64).reshape(-1, 3)
        self.triangles = np.ascontiguousarray(self.triangles, dtype=np.uint32).reshape(-1, 3)

    def _handle(self):
        lib = _lib()
        h = ctypes.c_void_p()
        _check(lib, lib.lbm_mesh_create(abi.ptr(self.vertices, ctypes.c_double), len(self.vertices),
                                        abi.ptr(self.triangles, ctypes.c_uint32), len(self.triangles),
                                        ctypes.byref(h)))
        return _Handle(lib, h)

    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        return self.vertices.min(axis=0), self.vertices.max(axis=0)

    def transformed(self, rotation=None, pivot=(0.0, 0.0, 0.0), translation=(0.0, 0.0, 0.0)) -> "TriangleMesh":
        """v <- R (v - pivot) + translation."""
        r = np.eye(3) if rotation is None else np.asarray(rotation, dtype=np.float64)
        with self._handle() as h:
            lib = _lib()
            _check(lib, lib.lbm_mesh_transform(h, abi.ptr(np.ascontiguousarray(r.reshape(9)), ctypes.c_double),
                                               abi.ptr(np.asarray(pivot, dtype=np.float64), ctypes.c_double),
                                               abi.ptr(np.asarray(translation, dtype=np.float64), ctypes.c_double)))
            return _mesh_of(lib, h)

    def winding_numbers(self, points) -> np.ndarray:
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.empty(len(p))
        with self._handle() as h:
            lib = _lib()
            _check(lib, lib.lbm_winding_numbers(h, abi.ptr(p, ctypes.c_double), len(p), abi.ptr(out, ctypes.c_double)))
        return out

    def surface_area(self) -> float:
        v = self.vertices
        t = self.triangles.astype(np.int64)
        c = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
        return float(0.5 * np.sqrt((c * c).sum(axis=1)).sum())

    def volume(self) -> float:
        v = self.vertices
        t = self.triangles.astype(np.int64)
        return float(np.einsum("ij,ij->i", v[t[:, 0]], np.cross(v[t[:, 1]], v[t[:, 2]])).sum() / 6.0)


class _Handle:
    def __init__(self, lib, h):
        self.lib = lib
        self.h = h

    def __enter__(self):
        return self.h

    def __exit__(self, *exc):
        self.lib.lbm_mesh_destroy(self.h)


def _mesh_of(lib, h) -> TriangleMesh:
    nv = ctypes.c_size_t()
    nt = ctypes.c_size_t()
    _check(lib, lib.lbm_mesh_size(h, ctypes.byref(nv), ctypes.byref(nt)))
    v = np.empty((nv.value, 3))
    t = np.empty((nt.value, 3), dtype=np.uint32)
    _check(lib, lib.lbm_mesh_data(h, abi.ptr(v, ctypes.c_double), abi.ptr(t, ctypes.c_uint32)))
    return TriangleMesh(v, t)


def _made(call) -> TriangleMesh:
    lib = _lib()
    h = ctypes.c_void_p()
    _check(lib, call(lib, ctypes.byref(h)))
    with _Handle(lib, h):
        return _mesh_of(lib, h)


@dataclass
class RevolutionResolution:
    """Segments around the axis, and the largest angle (radians) of one piece of a profile arc
    (primitives.h:43-47)."""

    around: int = 64
    arc_step: float = 0.05


@dataclass
class CapsuleShape:
    """Capsule of revolution (primitives.h:99-114); the defaults are the Apollo-like shape."""

    max_diameter: float = 3.912
    nose_radius: float = 4.694
    shoulder_radius: float = 0.196
    afterbody_half_angle: float = 0.5759586531581288
    top_diameter: float = 0.7

    @staticmethod
    def apollo() -> "CapsuleShape":
        return CapsuleShape()


def _revolution(kind, a, b=0.0, c=0.0, d=0.0, e=0.0, resolution: RevolutionResolution | None = None):
    r = resolution or RevolutionResolution()
    return _made(lambda lib, out: lib.lbm_mesh_revolution(kind, a, b, c, d, e, r.around, r.arc_step, out))


def sphere_mesh(radius: float, resolution: RevolutionResolution | None = None) -> TriangleMesh:
    """sphere_profile(radius).mesh(resolution), centred on the origin."""
    return _revolution(0, radius, resolution=resolution)


def cylinder_mesh(radius: float, height: float, resolution: RevolutionResolution | None = None) -> TriangleMesh:
    """cylinder_profile(radius, height).mesh(resolution): axis z, base at z = 0."""
    return _revolution(1, radius, height, resolution=resolution)


def cone_mesh(radius: float, height: float, resolution: RevolutionResolution | None = None) -> TriangleMesh:
    """cone_profile(radius, height).mesh(resolution): base at z = 0, apex at z = height."""
    return _revolution(2, radius, height, resolution=resolution)


def capsule_mesh(shape: CapsuleShape | None = None, resolution: RevolutionResolution | None = None) -> TriangleMesh:
    """capsule_profile(shape).mesh(resolution): heat shield at the bottom, axis z."""
    s = shape or CapsuleShape()
    return _revolution(3, s.max_diameter, s.nose_radius, s.shoulder_radius, s.afterbody_half_angle, s.top_diameter,
                       resolution)


def wedge_mesh(length: float, half_angle: float, span: float) -> TriangleMesh:
    """Wedge::mesh: apex edge along y at x = 0, base at x = length."""
    return _made(lambda lib, out: lib.lbm_mesh_wedge(length, half_angle, span, out))


def box_mesh(lx: float, ly: float, lz: float) -> TriangleMesh:
    """box_mesh of the given sizes, centred on the origin."""
    return _made(lambda lib, out: lib.lbm_mesh_box(lx, ly, lz, out))


def naca_wing(max_camber: float, camber_position: float, thickness: float, chord: float, span: float, n: int,
              span_divisions: int) -> TriangleMesh:
    """naca_wing(Naca4{...}, chord, span, n, span_divisions).mesh: chord along x from the
    leading edge at the origin, span along y centred on 0."""
    return _made(lambda lib, out: lib.lbm_mesh_naca_wing(max_camber, camber_position, thickness, chord, span, n,
                                                         span_divisions, out))


def naca_section(max_camber: float, camber_position: float, thickness: float, n: int) -> np.ndarray:
    """Naca4::section(n) (src/geom/naca.cpp:155-172): [2 n, 2] points (x, z) per unit chord,
    counter-clockwise from the trailing edge over the upper surface."""
    lib = _lib()
    out = np.empty((2 * max(int(n), 0), 2))
    _check(lib, lib.lbm_naca_section(max_camber, camber_position, thickness, int(n),
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if out.size else None))
    return out


@dataclass
class GridSpec:
    """Cell (i, j, k) at index i + nx (j + ny k), centre origin + (i + 1/2, j + 1/2, k + 1/2) spacing."""

    origin: tuple
    spacing: float
    nx: int
    ny: int
    nz: int

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def shape(self) -> tuple[int, int, int]:
        return (self.nz, self.ny, self.nx)


@dataclass
class VoxelGrid:
    spec: GridSpec
    solid: np.ndarray  # uint8 [cells]
    winding: np.ndarray  # float64 [cells]
    rechecked: int | None = None  # device voxelisation: cells decided by the host evaluation


@dataclass
class SignedDistanceField:
    spec: GridSpec
    band: float
    distance: np.ndarray  # float64 [cells]
    in_band: np.ndarray = field(default=None)  # uint8 [cells]


def voxelize(mesh: TriangleMesh, spec: GridSpec, threshold: float = 0.5, device: bool = False) -> VoxelGrid:
    """Solid where the generalised winding number at the cell centre is at least `threshold`.
    With `device`, the cells are evaluated on the GPU (the same mask bit for bit; near-threshold
    cells are decided by the host evaluation, counted in `rechecked`)."""
    lib = _lib()
    solid = np.empty(spec.cells, dtype=np.uint8)
    winding = np.empty(spec.cells)
    o = np.asarray(spec.origin, dtype=np.float64)
    rechecked = None
    with mesh._handle() as h:
        if device:
            n = ctypes.c_size_t()
            _check(lib, lib.lbm_voxelize_device(h, abi.ptr(o, ctypes.c_double), spec.spacing, spec.nx, spec.ny, spec.nz,
                                                threshold, abi.ptr(solid, ctypes.c_uint8),
                                                abi.ptr(winding, ctypes.c_double), ctypes.byref(n)))
            rechecked = n.value
        else:
            _check(lib, lib.lbm_voxelize(h, abi.ptr(o, ctypes.c_double), spec.spacing, spec.nx, spec.ny, spec.nz,
                                         threshold, abi.ptr(solid, ctypes.c_uint8), abi.ptr(winding, ctypes.c_double)))
    return VoxelGrid(spec, solid, winding, rechecked)


def signed_distance_band(mesh: TriangleMesh, voxels: VoxelGrid, band_cells: float = 3.0,
                         device: bool = False) -> SignedDistanceField:
    """Distance to the surface, exact within band_cells * spacing, negative in solid voxels
    (with `device`, computed on the GPU, bit for bit the same)."""
    lib = _lib()
    spec = voxels.spec
    if voxels.solid.size != spec.cells:
        raise GeometryError("signed_distance_band: voxel grid does not match its spec")
    d = np.empty(spec.cells)
    band = np.empty(spec.cells, dtype=np.uint8)
    o = np.asarray(spec.origin, dtype=np.float64)
    solid = np.ascontiguousarray(voxels.solid, dtype=np.uint8)
    fn = lib.lbm_signed_distance_band_device if device else lib.lbm_signed_distance_band
    with mesh._handle() as h:
        _check(lib, fn(h, abi.ptr(o, ctypes.c_double), spec.spacing, spec.nx, spec.ny, spec.nz,
                       abi.ptr(solid, ctypes.c_uint8), band_cells, abi.ptr(d, ctypes.c_double),
                       abi.ptr(band, ctypes.c_uint8)))
    return SignedDistanceField(spec, band_cells * spec.spacing, d, band)


def _check_placement(mask: SolidMask3D, spec: GridSpec, values: int, offset) -> None:
    """check_placement (solid_mask3d.cpp:107-126)."""
    if values != spec.cells:
        raise GeometryError("lbm3d: the voxel vector does not match its grid")
    ox, oy, oz = offset
    if (ox > mask.nx or spec.nx > mask.nx - ox or oy > mask.ny or spec.ny > mask.ny - oy or oz > mask.nz
            or spec.nz > mask.nz - oz):
        raise GeometryError("lbm3d: the voxel grid does not fit in the lattice")


def place_voxels(mask: SolidMask3D, voxels: VoxelGrid, offset) -> None:
    """Sets the mask's solid flag at offset + (i, j, k) for every solid voxel (solid_mask3d.cpp:130-143)."""
    spec = voxels.spec
    _check_placement(mask, spec, voxels.solid.size, offset)
    ox, oy, oz = offset
    m = mask.solid.reshape(mask.nz, mask.ny, mask.nx)
    box = m[oz:oz + spec.nz, oy:oy + spec.ny, ox:ox + spec.nx]
    box |= voxels.solid.reshape(spec.shape()) != 0


def place_distance(mask: SolidMask3D, sdf: SignedDistanceField, offset) -> None:
    """Copies distance / spacing into the mask's distance field, FAR_DISTANCE elsewhere
    (solid_mask3d.cpp:145-163)."""
    spec = sdf.spec
    _check_placement(mask, spec, sdf.distance.size, offset)
    if not (spec.spacing > 0.0 and math.isfinite(spec.spacing)):
        raise GeometryError("lbm3d: the distance grid needs a positive spacing")
    if mask.distance is None or mask.distance.size != mask.solid.size:
        mask.distance = np.full(mask.solid.size, FAR_DISTANCE)
    ox, oy, oz = offset
    d = mask.distance.reshape(mask.nz, mask.ny, mask.nx)
    d[oz:oz + spec.nz, oy:oy + spec.ny, ox:ox + spec.nx] = sdf.distance.reshape(spec.shape()) / spec.spacing


def rotation_y(angle: float) -> np.ndarray:
    """Rotation about +y by `angle` radians (pitch)."""
    c, s = math.cos(angle), math.sin(angle)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def rotation_z(angle: float) -> np.ndarray:
    """Rotation about +z by `angle` radians (yaw)."""
    c, s = math.cos(angle), math.sin(angle)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


__all__ = ["TriangleMesh", "RevolutionResolution", "CapsuleShape", "sphere_mesh", "cylinder_mesh", "cone_mesh",
           "capsule_mesh", "wedge_mesh", "box_mesh", "naca_wing", "naca_section", "GridSpec", "VoxelGrid", "SignedDistanceField",
           "voxelize", "signed_distance_band", "place_voxels", "place_distance", "rotation_y", "rotation_z",
           "FAR_DISTANCE", "GeometryError"]
