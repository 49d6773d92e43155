"""Solid masks of the lattices (SolidMask3D / SolidMask2D of the reference).

Node (x, y, z) has index x + nx (y + ny z). A node is solid where the signed distance function
is <= 0; the 3D masks keep the distance (lattice units, negative inside) for the interpolated
wall. The arithmetic follows src/aero/lbm3d/solid_mask3d.cpp:22-102 and
src/aero/lbm/solid_mask.cpp:10-43 operation for operation, so the masks are bit-identical.
"""
from __future__ import annotations

import dataclasses

import numpy as np

MAX_GRID_CELLS = 1 << 27  # geom::max_grid_cells (src/geom/include/orrery/geom/voxel_grid.h:29)
FAR_DISTANCE = 1e9        # far_distance (solid_mask3d.h:82)


@dataclasses.dataclass
class SolidMask3D:
    nx: int
    ny: int
    nz: int
    solid: np.ndarray                    # uint8, 1 = solid
    distance: np.ndarray | None = None   # float64 signed distance, or None

    def index(self, x: int, y: int, z: int) -> int:
        return x + self.nx * (y + self.ny * z)

    def at(self, x: int, y: int, z: int) -> bool:
        return bool(self.solid[self.index(x, y, z)])

    def count(self) -> int:
        return int(np.count_nonzero(self.solid))


@dataclasses.dataclass
class SolidMask2D:
    nx: int
    ny: int
    solid: np.ndarray

    def index(self, x: int, y: int) -> int:
        return x + self.nx * y

    def count(self) -> int:
        return int(np.count_nonzero(self.solid))


def lattice_cells_3d(nx: int, ny: int, nz: int) -> int | None:
    """nx ny nz, None for a zero dimension or more than max_grid_cells (solid_mask3d.cpp:22-36)."""
    if nx <= 0 or ny <= 0 or nz <= 0:
        return None
    cells = nx * ny * nz
    return cells if cells <= MAX_GRID_CELLS else None


def fluid_mask_3d(nx: int, ny: int, nz: int) -> SolidMask3D:
    if lattice_cells_3d(nx, ny, nz) is None:
        raise ValueError("lbm3d: mask dimensions must be positive with at most geom::max_grid_cells nodes")
    return SolidMask3D(nx, ny, nz, np.zeros(nx * ny * nz, dtype=np.uint8))


def _grid3(nx, ny, nz):
    z, y, x = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(ny, dtype=np.float64),
                          np.arange(nx, dtype=np.float64), indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


def mask_from_sdf_3d(nx: int, ny: int, nz: int, sdf) -> SolidMask3D:
    """`sdf(x, y, z)` on float64 arrays of node coordinates (mask_from_sdf_3d, :57-75)."""
    mask = fluid_mask_3d(nx, ny, nz)
    x, y, z = _grid3(nx, ny, nz)
    d = np.asarray(sdf(x, y, z), dtype=np.float64)
    mask.distance = d
    mask.solid = (d <= 0.0).astype(np.uint8)
    return mask


def sphere_mask(nx: int, ny: int, nz: int, centre, radius: float) -> SolidMask3D:
    """Sphere with its exact signed distance (sphere_mask, :77-85)."""
    c0, c1, c2 = (float(v) for v in centre)
    r = float(radius)

    def sdf(x, y, z):
        dx = x - c0
        dy = y - c1
        dz = z - c2
        return np.sqrt(dx * dx + dy * dy + dz * dz) - r

    return mask_from_sdf_3d(nx, ny, nz, sdf)


def box_mask(nx: int, ny: int, nz: int, lo, hi) -> SolidMask3D:
    """Nodes with lo <= position <= hi, exact box distance (box_mask, :87-102)."""
    lo = [float(v) for v in lo]
    hi = [float(v) for v in hi]

    def sdf(x, y, z):
        p = (x, y, z)
        outside = np.zeros_like(x)
        inside = np.full_like(x, -np.inf)
        for a in range(3):
            excess = np.abs(p[a] - 0.5 * (lo[a] + hi[a])) - 0.5 * (hi[a] - lo[a])
            pos = np.maximum(excess, 0.0)
            outside = outside + pos * pos
            inside = np.maximum(inside, excess)
        return np.sqrt(outside) + np.minimum(inside, 0.0)

    return mask_from_sdf_3d(nx, ny, nz, sdf)


def fluid_mask(nx: int, ny: int) -> SolidMask2D:
    if nx <= 0 or ny <= 0 or nx * ny > MAX_GRID_CELLS:
        raise ValueError("solid mask: dimensions must be at least 1 and the node count at most "
                         "geom::max_grid_cells")
    return SolidMask2D(nx, ny, np.zeros(nx * ny, dtype=np.uint8))


def mask_from_sdf(nx: int, ny: int, sdf) -> SolidMask2D:
    mask = fluid_mask(nx, ny)
    y, x = np.meshgrid(np.arange(ny, dtype=np.float64), np.arange(nx, dtype=np.float64), indexing="ij")
    d = np.asarray(sdf(x.ravel(), y.ravel()), dtype=np.float64)
    mask.solid = (d <= 0.0).astype(np.uint8)
    return mask


def circle_mask(nx: int, ny: int, cx: float, cy: float, radius: float) -> SolidMask2D:
    """Circle of radius r centred at (cx, cy) (circle_mask, src/aero/lbm/solid_mask.cpp:39-43)."""
    return mask_from_sdf(nx, ny, lambda x, y: np.hypot(x - float(cx), y - float(cy)) - float(radius))


def polygon_mask(nx: int, ny: int, polygon) -> SolidMask2D:
    """Nodes inside a closed polygon [n, 2] of (x, z) points in lattice units, even-odd rule
    (polygon_mask, src/aero/lbm/solid_mask.cpp:45-73), computed by the library."""
    import ctypes

    from . import abi
    lib = abi.load()
    pts = np.ascontiguousarray(np.asarray(polygon, dtype=np.float64).reshape(-1, 2))
    mask = fluid_mask(nx, ny)
    if lib.lbm_polygon_mask(nx, ny, abi.ptr(pts, ctypes.c_double), len(pts),
                            abi.ptr(mask.solid, ctypes.c_uint8)) != abi.LBM_OK:
        raise ValueError(lib.lbm_geom_last_error().decode())
    return mask
