"""Python host API over the C ABI: the Lbm3DSolver / Lbm2DSolver contract of the reference.

Mirrors src/aero/lbm3d/include/orrery/aero/lbm3d/lbm3d.h:196-247 and
src/aero/lbm/include/orrery/aero/lbm/lbm2d.h:173-211: make_hip_solver() stands where
make_gpu_solver() stands in the reference (lbm3d_gpu.cpp:487-513). Errors the reference returns
as core::Result failures raise LbmError with the same message.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import abi
from .config import Lbm2DConfig, Lbm3DConfig
from .masks import SolidMask2D, SolidMask3D


@dataclasses.dataclass(frozen=True)
class NodeCounts:
    """Node classes of a D3Q19 lattice (lbm3d_counts)."""
    bulk_nodes: int
    boundary_nodes: int
    solid_nodes: int
    links: int


class LbmError(RuntimeError):
    """A refused configuration or a failed device operation (code and message of the ABI)."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def _check(lib, rc: int) -> None:
    if rc != abi.LBM_OK:
        raise LbmError(rc, lib.lbm_last_error().decode())


@dataclasses.dataclass
class MacroFields3D:
    nx: int
    ny: int
    nz: int
    density: np.ndarray
    velocity_x: np.ndarray
    velocity_y: np.ndarray
    velocity_z: np.ndarray
    solid: np.ndarray


@dataclasses.dataclass
class MacroFields2D:
    nx: int
    ny: int
    density: np.ndarray
    velocity_x: np.ndarray
    velocity_y: np.ndarray
    solid: np.ndarray


@dataclasses.dataclass
class Force3D:
    force: tuple
    moment: tuple


@dataclasses.dataclass
class Force2D:
    x: float
    y: float
    moment: float


@dataclasses.dataclass
class BoundaryLinks:
    cell: np.ndarray
    direction: np.ndarray
    q: np.ndarray


def device_memory() -> tuple[int, int]:
    """(available, max_alloc) bytes of the current device."""
    lib = abi.load()
    a = ctypes.c_size_t()
    m = ctypes.c_size_t()
    _check(lib, lib.lbm_device_memory(ctypes.byref(a), ctypes.byref(m)))
    return a.value, m.value


def set_memory_limit(nbytes: int) -> None:
    """Caps the reported free device memory (0 clears); the admission checks honour it."""
    abi.load().lbm_set_memory_limit(int(nbytes))


class Lbm3DSolver:
    """A D3Q19 lattice resident on the GPU (one HIP stream)."""

    def __init__(self, config: Lbm3DConfig, mask: SolidMask3D, storage: str = "f32"):
        lib = abi.load()
        self._lib = lib
        self._config = config.copy()
        self._mask = mask
        self._storage = storage
        self._c = config.to_c()
        solid = np.ascontiguousarray(mask.solid, dtype=np.uint8)
        distance = None if mask.distance is None else np.ascontiguousarray(mask.distance, dtype=np.float64)
        handle = ctypes.c_void_p()
        if storage not in abi.STORAGE:
            raise LbmError(abi.LBM_EINVAL, f"unknown storage {storage!r}")
        rc = lib.lbm3d_create(ctypes.byref(self._c), abi.ptr(solid, ctypes.c_uint8),
                              abi.ptr(distance, ctypes.c_double), abi.STORAGE[storage], ctypes.byref(handle))
        _check(lib, rc)
        self._h = handle

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.lbm3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- Lbm3DSolver contract ----
    def config(self) -> Lbm3DConfig:
        return self._config

    def mask(self) -> SolidMask3D:
        return self._mask

    @property
    def storage(self) -> str:
        return self._storage

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def steps(self) -> int:
        return int(self._lib.lbm3d_steps(self._h))

    def initialize(self, initial=None) -> None:
        """Equilibrium of `initial`: None (uniform inlet state), an array of shape (4, cells)
        with (rho, ux, uy, uz), or a callable (x, y, z) -> (rho, ux, uy, uz) (InitialState3D)."""
        if initial is None:
            _check(self._lib, self._lib.lbm3d_initialize(self._h, None))
            return
        c = self._config
        if callable(initial):
            planes = np.zeros((4, c.cells))
            for z in range(c.nz):
                for y in range(c.ny):
                    for x in range(c.nx):
                        planes[:, x + c.nx * (y + c.ny * z)] = initial(x, y, z)
        else:
            planes = np.ascontiguousarray(initial, dtype=np.float64).reshape(4, c.cells)
        _check(self._lib, self._lib.lbm3d_initialize(self._h, abi.ptr(planes, ctypes.c_double)))

    def initialize_perturbed(self, seed: int, amplitude: float) -> None:
        _check(self._lib, self._lib.lbm3d_initialize_perturbed(self._h, seed, amplitude))

    def step(self, count: int) -> None:
        _check(self._lib, self._lib.lbm3d_step(self._h, int(count)))

    def synchronize(self) -> None:
        _check(self._lib, self._lib.lbm3d_synchronize(self._h))

    def set_inlet_velocity(self, velocity) -> None:
        v = np.asarray(velocity, dtype=np.float64)
        _check(self._lib, self._lib.lbm3d_set_inlet(self._h, abi.ptr(v, ctypes.c_double)))

    def inlet_velocity(self) -> tuple:
        v = np.zeros(3)
        _check(self._lib, self._lib.lbm3d_get_inlet(self._h, abi.ptr(v, ctypes.c_double)))
        return tuple(v)

    def macroscopic(self) -> MacroFields3D:
        c = self._config
        out = np.empty((4, c.cells))
        p = [abi.ptr(out[k], ctypes.c_double) for k in range(4)]
        _check(self._lib, self._lib.lbm3d_macroscopic(self._h, *p))
        return MacroFields3D(c.nx, c.ny, c.nz, out[0], out[1], out[2], out[3],
                             np.asarray(self._mask.solid, dtype=np.uint8))

    def populations(self) -> np.ndarray:
        f = np.empty(19 * self._config.cells)
        _check(self._lib, self._lib.lbm3d_populations(self._h, abi.ptr(f, ctypes.c_double)))
        return f

    def force(self) -> Force3D:
        f = np.zeros(3)
        m = np.zeros(3)
        _check(self._lib, self._lib.lbm3d_force(self._h, abi.ptr(f, ctypes.c_double), abi.ptr(m, ctypes.c_double)))
        return Force3D(tuple(f), tuple(m))

    def mass(self) -> float:
        v = ctypes.c_double()
        _check(self._lib, self._lib.lbm3d_mass(self._h, ctypes.byref(v)))
        return v.value

    def links(self) -> BoundaryLinks:
        n = ctypes.c_size_t()
        _check(self._lib, self._lib.lbm3d_links(self._h, None, None, None, 0, ctypes.byref(n)))
        cell = np.empty(n.value, dtype=np.uint32)
        d = np.empty(n.value, dtype=np.uint32)
        q = np.empty(n.value)
        _check(self._lib, self._lib.lbm3d_links(self._h, abi.ptr(cell, ctypes.c_uint32), abi.ptr(d, ctypes.c_uint32),
                                                abi.ptr(q, ctypes.c_double), n.value, ctypes.byref(n)))
        return BoundaryLinks(cell, d, q)

    # ---- device extras ----
    def device_bytes(self) -> int:
        v = ctypes.c_size_t()
        _check(self._lib, self._lib.lbm3d_device_bytes(self._h, ctypes.byref(v)))
        return v.value

    def set_launch(self, block: int = 0, cells_per_thread: int = 0) -> None:
        _check(self._lib, self._lib.lbm3d_set_launch(self._h, block, cells_per_thread))

    def set_occupancy(self, groups_a: int, groups_b: int) -> None:
        """Cap on the bulk kernel's workgroups per compute unit for the A step and the B step (0:
        none); results do not change."""
        _check(self._lib, self._lib.lbm3d_set_occupancy(self._h, int(groups_a), int(groups_b)))

    def occupancy(self) -> tuple[int, int]:
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _check(self._lib, self._lib.lbm3d_get_occupancy(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def counts(self) -> "NodeCounts":
        """lbm3d_counts: bulk nodes, boundary nodes, solid nodes and solid links of the lattice."""
        v = [ctypes.c_size_t() for _ in range(4)]
        _check(self._lib, self._lib.lbm3d_counts(self._h, *[ctypes.byref(x) for x in v]))
        return NodeCounts(*(x.value for x in v))

    def stream_counts(self) -> tuple:
        """lbm3d_stream_counts: (bulk, side, listed) fluid nodes: the bulk kernel streams the bulk
        nodes, and in A steps the side nodes (boundary nodes on slip or periodic y and z faces),
        which a side kernel updates in B steps; the boundary kernel the listed nodes."""
        v = [ctypes.c_size_t() for _ in range(3)]
        _check(self._lib, self._lib.lbm3d_stream_counts(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_schedule(self, schedule: str) -> None:
        """"serial" (bulk then boundary kernel, the default), "fused" (one kernel per step) or
        "streams" (the two kernels concurrently on two streams); all give identical results."""
        _check(self._lib, self._lib.lbm3d_set_schedule(self._h, abi.SCHEDULE[schedule]))

    def set_graphs(self, enabled: bool) -> None:
        """Whether step() replays captured HIP graphs (the default) or launches every step."""
        _check(self._lib, self._lib.lbm3d_set_graphs(self._h, 1 if enabled else 0))

    def step_bytes(self) -> float:
        v = ctypes.c_double()
        _check(self._lib, self._lib.lbm3d_step_bytes(self._h, ctypes.byref(v)))
        return v.value


class Lbm2DSolver:
    """A D2Q9 lattice resident on the GPU."""

    def __init__(self, config: Lbm2DConfig, mask: SolidMask2D, storage: str = "f32"):
        lib = abi.load()
        self._lib = lib
        self._config = config
        self._mask = mask
        self._storage = storage
        self._c = config.to_c()
        solid = np.ascontiguousarray(mask.solid, dtype=np.uint8)
        handle = ctypes.c_void_p()
        if storage not in abi.STORAGE:
            raise LbmError(abi.LBM_EINVAL, f"unknown storage {storage!r}")
        _check(lib, lib.lbm2d_create(ctypes.byref(self._c), abi.ptr(solid, ctypes.c_uint8), abi.STORAGE[storage],
                                     ctypes.byref(handle)))
        self._h = handle

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.lbm2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self) -> Lbm2DConfig:
        return self._config

    def mask(self) -> SolidMask2D:
        return self._mask

    def steps(self) -> int:
        return int(self._lib.lbm2d_steps(self._h))

    def set_graphs(self, enabled: bool) -> None:
        """Whether step() replays captured HIP graphs (the default) or launches every step."""
        _check(self._lib, self._lib.lbm2d_set_graphs(self._h, 1 if enabled else 0))

    def set_schedule(self, schedule: str) -> None:
        """"fused" (bulk and boundary in one launch, the default) or "serial" (two kernels)."""
        _check(self._lib, self._lib.lbm2d_set_schedule(self._h, abi.SCHEDULE[schedule]))

    def initialize(self, initial=None) -> None:
        if initial is None:
            _check(self._lib, self._lib.lbm2d_initialize(self._h, None))
            return
        c = self._config
        if callable(initial):
            planes = np.zeros((3, c.cells))
            for y in range(c.ny):
                for x in range(c.nx):
                    planes[:, x + c.nx * y] = initial(x, y)
        else:
            planes = np.ascontiguousarray(initial, dtype=np.float64).reshape(3, c.cells)
        _check(self._lib, self._lib.lbm2d_initialize(self._h, abi.ptr(planes, ctypes.c_double)))

    def initialize_perturbed(self, seed: int, amplitude: float) -> None:
        _check(self._lib, self._lib.lbm2d_initialize_perturbed(self._h, seed, amplitude))

    def step(self, count: int) -> None:
        _check(self._lib, self._lib.lbm2d_step(self._h, int(count)))

    def synchronize(self) -> None:
        _check(self._lib, self._lib.lbm2d_synchronize(self._h))

    def macroscopic(self) -> MacroFields2D:
        c = self._config
        out = np.empty((3, c.cells))
        p = [abi.ptr(out[k], ctypes.c_double) for k in range(3)]
        _check(self._lib, self._lib.lbm2d_macroscopic(self._h, *p))
        return MacroFields2D(c.nx, c.ny, out[0], out[1], out[2], np.asarray(self._mask.solid, dtype=np.uint8))

    def sample(self, cells) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(rho, ux, uy) at the listed cell indices, computed on the device: the values
        macroscopic() reports there."""
        idx = np.ascontiguousarray(cells, dtype=np.uint32)
        out = np.empty((3, len(idx)))
        p = [abi.ptr(out[k], ctypes.c_double) for k in range(3)]
        _check(self._lib, self._lib.lbm2d_sample(self._h, abi.ptr(idx, ctypes.c_uint32), len(idx), *p))
        return out[0], out[1], out[2]

    def max_speed(self) -> float:
        """Largest hypot(ux, uy) over the fluid nodes, reduced on the device."""
        v = ctypes.c_double()
        _check(self._lib, self._lib.lbm2d_max_speed(self._h, ctypes.byref(v)))
        return v.value

    def populations(self) -> np.ndarray:
        f = np.empty(9 * self._config.cells)
        _check(self._lib, self._lib.lbm2d_populations(self._h, abi.ptr(f, ctypes.c_double)))
        return f

    def force(self) -> Force2D:
        v = np.zeros(3)
        _check(self._lib, self._lib.lbm2d_force(self._h, abi.ptr(v[0:1], ctypes.c_double),
                                                abi.ptr(v[1:2], ctypes.c_double), abi.ptr(v[2:3], ctypes.c_double)))
        return Force2D(float(v[0]), float(v[1]), float(v[2]))

    def mass(self) -> float:
        v = ctypes.c_double()
        _check(self._lib, self._lib.lbm2d_mass(self._h, ctypes.byref(v)))
        return v.value

    def links(self) -> BoundaryLinks:
        n = ctypes.c_size_t()
        _check(self._lib, self._lib.lbm2d_links(self._h, None, None, 0, ctypes.byref(n)))
        cell = np.empty(n.value, dtype=np.uint32)
        d = np.empty(n.value, dtype=np.uint32)
        _check(self._lib, self._lib.lbm2d_links(self._h, abi.ptr(cell, ctypes.c_uint32), abi.ptr(d, ctypes.c_uint32),
                                                n.value, ctypes.byref(n)))
        return BoundaryLinks(cell, d, np.full(n.value, 0.5))


def make_hip_solver(config, mask, storage: str = "f32"):
    """The device solver for a 3D or 2D configuration (make_gpu_solver's counterpart)."""
    if isinstance(config, Lbm3DConfig):
        return Lbm3DSolver(config, mask, storage)
    return Lbm2DSolver(config, mask, storage)


# ---- host-only setup (no device) -----------------------------------------------------------
def validate(config, mask) -> str | None:
    """The reference's validate(): None when the configuration can run, else the reason."""
    lib = abi.load()
    c = config.to_c()
    solid = np.ascontiguousarray(mask.solid, dtype=np.uint8)
    if isinstance(config, Lbm3DConfig):
        distance = None if mask.distance is None else np.ascontiguousarray(mask.distance, dtype=np.float64)
        rc = lib.lbm3d_validate(ctypes.byref(c), abi.ptr(solid, ctypes.c_uint8), abi.ptr(distance, ctypes.c_double))
    else:
        rc = lib.lbm2d_validate(ctypes.byref(c), abi.ptr(solid, ctypes.c_uint8))
    return None if rc == abi.LBM_OK else lib.lbm_last_error().decode()


def classify_nodes(config: Lbm3DConfig, mask: SolidMask3D) -> np.ndarray:
    """Node classes 0 solid, 1 bulk, 2 boundary (classify_nodes, lbm3d_internal.h:219-248)."""
    lib = abi.load()
    c = config.to_c()
    out = np.empty(config.cells, dtype=np.uint8)
    solid = np.ascontiguousarray(mask.solid, dtype=np.uint8)
    _check(lib, lib.lbm3d_classify(ctypes.byref(c), abi.ptr(solid, ctypes.c_uint8), abi.ptr(out, ctypes.c_uint8)))
    return out


def boundary_links(config, mask) -> BoundaryLinks:
    """boundary_links() of the reference, computed by the library's host setup."""
    lib = abi.load()
    c = config.to_c()
    solid = np.ascontiguousarray(mask.solid, dtype=np.uint8)
    n = ctypes.c_size_t()
    if isinstance(config, Lbm3DConfig):
        distance = None if mask.distance is None else np.ascontiguousarray(mask.distance, dtype=np.float64)
        args = (ctypes.byref(c), abi.ptr(solid, ctypes.c_uint8), abi.ptr(distance, ctypes.c_double))
        _check(lib, lib.lbm3d_boundary_links(*args, None, None, None, 0, ctypes.byref(n)))
        cell = np.empty(n.value, dtype=np.uint32)
        d = np.empty(n.value, dtype=np.uint32)
        q = np.empty(n.value)
        _check(lib, lib.lbm3d_boundary_links(*args, abi.ptr(cell, ctypes.c_uint32), abi.ptr(d, ctypes.c_uint32),
                                             abi.ptr(q, ctypes.c_double), n.value, ctypes.byref(n)))
        return BoundaryLinks(cell, d, q)
    args = (ctypes.byref(c), abi.ptr(solid, ctypes.c_uint8))
    _check(lib, lib.lbm2d_boundary_links(*args, None, None, 0, ctypes.byref(n)))
    cell = np.empty(n.value, dtype=np.uint32)
    d = np.empty(n.value, dtype=np.uint32)
    _check(lib, lib.lbm2d_boundary_links(*args, abi.ptr(cell, ctypes.c_uint32), abi.ptr(d, ctypes.c_uint32),
                                         n.value, ctypes.byref(n)))
    return BoundaryLinks(cell, d, np.full(n.value, 0.5))
