"""ctypes binding of the oracle (oracle/liborrery_oracle.so), the FP64 CPU restatement of the
reference LBM solver. Test infrastructure: the same interface as orrery_amd's device solvers so
parity tests read like the reference's GPU-vs-CPU tests (tests/aero/lbm3d/test_lbm3d_gpu.cpp)."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from orrery_amd import abi
from orrery_amd.config import Lbm2DConfig, Lbm3DConfig
from orrery_amd.solver import BoundaryLinks, Force2D, Force3D, MacroFields2D, MacroFields3D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBRARY = os.environ.get("ORRERY_ORACLE_LIBRARY") or os.path.join(ROOT, "oracle", "liborrery_oracle.so")

P = ctypes.POINTER
_f64p = P(ctypes.c_double)
_u8p = P(ctypes.c_uint8)
_u32p = P(ctypes.c_uint32)
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    lib = ctypes.CDLL(LIBRARY)
    sig = {
        "orc_set_threads": (ctypes.c_int, [ctypes.c_int]),
        "orc_perturbation": (ctypes.c_double, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32]),
        "orc3d_validate": (ctypes.c_int, [P(abi.CConfig3D), _u8p, _f64p, ctypes.c_char_p, ctypes.c_size_t]),
        "orc3d_create": (ctypes.c_void_p, [P(abi.CConfig3D), _u8p, _f64p, ctypes.c_char_p, ctypes.c_size_t]),
        "orc3d_destroy": (None, [ctypes.c_void_p]),
        "orc3d_initialize": (None, [ctypes.c_void_p, _f64p]),
        "orc3d_initialize_perturbed": (None, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double]),
        "orc3d_step": (None, [ctypes.c_void_p, ctypes.c_uint64]),
        "orc3d_steps": (ctypes.c_uint64, [ctypes.c_void_p]),
        "orc3d_set_inlet": (None, [ctypes.c_void_p, _f64p]),
        "orc3d_force": (None, [ctypes.c_void_p, _f64p, _f64p]),
        "orc3d_mass": (ctypes.c_double, [ctypes.c_void_p]),
        "orc3d_macroscopic": (None, [ctypes.c_void_p, _f64p, _f64p, _f64p, _f64p]),
        "orc3d_populations": (None, [ctypes.c_void_p, _f64p]),
        "orc3d_set_populations": (None, [ctypes.c_void_p, _f64p]),
        "orc3d_outlet_history": (ctypes.c_size_t, [ctypes.c_void_p, _f64p]),
        "orc3d_links": (ctypes.c_size_t, [ctypes.c_void_p, _u32p, _u32p, _f64p, ctypes.c_size_t]),
        "orc3d_classes": (None, [ctypes.c_void_p, _u8p]),
        "orc3d_resolve": (ctypes.c_int, [P(abi.CConfig3D), _u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                         ctypes.c_int, P(ctypes.c_uint64), P(ctypes.c_int), _f64p]),
        "orc3d_link_fraction": (ctypes.c_double, [P(abi.CConfig3D), _u8p, _f64p, ctypes.c_uint64, ctypes.c_uint64,
                                                  ctypes.c_uint64, ctypes.c_int]),
        "orc3d_sponge": (ctypes.c_double, [P(abi.CConfig3D), ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]),
        "orc3d_inlet_west": (None, [_f64p, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "orc3d_collide": (None, [_f64p, P(abi.CConfig3D), ctypes.c_double, _f64p]),
        "orc3d_equilibrium": (ctypes.c_double, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double]),
        "orc2d_validate": (ctypes.c_int, [P(abi.CConfig2D), _u8p, ctypes.c_char_p, ctypes.c_size_t]),
        "orc2d_create": (ctypes.c_void_p, [P(abi.CConfig2D), _u8p, ctypes.c_char_p, ctypes.c_size_t]),
        "orc2d_destroy": (None, [ctypes.c_void_p]),
        "orc2d_initialize": (None, [ctypes.c_void_p, _f64p]),
        "orc2d_initialize_perturbed": (None, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double]),
        "orc2d_step": (None, [ctypes.c_void_p, ctypes.c_uint64]),
        "orc2d_steps": (ctypes.c_uint64, [ctypes.c_void_p]),
        "orc2d_force": (None, [ctypes.c_void_p, _f64p, _f64p, _f64p]),
        "orc2d_mass": (ctypes.c_double, [ctypes.c_void_p]),
        "orc2d_macroscopic": (None, [ctypes.c_void_p, _f64p, _f64p, _f64p]),
        "orc2d_populations": (None, [ctypes.c_void_p, _f64p]),
        "orc2d_links": (ctypes.c_size_t, [ctypes.c_void_p, _u32p, _u32p, ctypes.c_size_t]),
        "orc2d_resolve": (ctypes.c_int, [P(abi.CConfig2D), _u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                         P(ctypes.c_uint64), P(ctypes.c_int), _f64p]),
        "orc2d_zou_he_west": (None, [_f64p, ctypes.c_double, ctypes.c_double]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a, t=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(P(t))


class OracleError(ValueError):
    pass


def validate3(config: Lbm3DConfig, mask) -> str | None:
    lib = load()
    c = config.to_c()
    buf = ctypes.create_string_buffer(512)
    solid = np.ascontiguousarray(mask.solid, dtype=np.uint8)
    dist = None if mask.distance is None else np.ascontiguousarray(mask.distance)
    rc = lib.orc3d_validate(ctypes.byref(c), ptr(solid, ctypes.c_uint8), ptr(dist), buf, 512)
    return buf.value.decode() if rc else None


class Oracle3D:
    """CPU FP64 D3Q19 solver (make_cpu_solver of the reference)."""

    def __init__(self, config: Lbm3DConfig, mask):
        self._lib = load()
        self._config = config.copy()
        self._mask = mask
        self._c = config.to_c()
        solid = np.ascontiguousarray(mask.solid, dtype=np.uint8)
        dist = None if mask.distance is None else np.ascontiguousarray(mask.distance)
        buf = ctypes.create_string_buffer(512)
        h = self._lib.orc3d_create(ctypes.byref(self._c), ptr(solid, ctypes.c_uint8), ptr(dist), buf, 512)
        if not h:
            raise OracleError(buf.value.decode())
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.orc3d_destroy(self._h)
            self._h = None

    def config(self):
        return self._config

    def steps(self):
        return int(self._lib.orc3d_steps(self._h))

    def initialize(self, initial=None):
        if initial is None:
            self._lib.orc3d_initialize(self._h, None)
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
        self._lib.orc3d_initialize(self._h, ptr(planes))

    def initialize_perturbed(self, seed, amplitude):
        self._lib.orc3d_initialize_perturbed(self._h, seed, amplitude)

    def step(self, count):
        self._lib.orc3d_step(self._h, int(count))

    def set_inlet_velocity(self, v):
        v = np.asarray(v, dtype=np.float64)
        self._lib.orc3d_set_inlet(self._h, ptr(v))

    def macroscopic(self) -> MacroFields3D:
        c = self._config
        out = np.empty((4, c.cells))
        self._lib.orc3d_macroscopic(self._h, *[ptr(out[k]) for k in range(4)])
        return MacroFields3D(c.nx, c.ny, c.nz, out[0], out[1], out[2], out[3],
                             np.asarray(self._mask.solid, dtype=np.uint8))

    def populations(self):
        f = np.empty(19 * self._config.cells)
        self._lib.orc3d_populations(self._h, ptr(f))
        return f

    def set_populations(self, f):
        """Replaces the state (logical layout f[i * cells + cell]); a test hook."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.size == 19 * self._config.cells
        self._lib.orc3d_set_populations(self._h, ptr(f))

    def outlet_history(self):
        """The convective-outlet history, [5 slots][nz][ny] (slots: directions 2, 8, 10, 12, 14)."""
        n = self._lib.orc3d_outlet_history(self._h, None)
        out = np.empty(n)
        self._lib.orc3d_outlet_history(self._h, ptr(out))
        return out.reshape(5, self._config.nz, self._config.ny)

    def force(self) -> Force3D:
        f = np.zeros(3)
        m = np.zeros(3)
        self._lib.orc3d_force(self._h, ptr(f), ptr(m))
        return Force3D(tuple(f), tuple(m))

    def mass(self):
        return float(self._lib.orc3d_mass(self._h))

    def links(self) -> BoundaryLinks:
        n = self._lib.orc3d_links(self._h, None, None, None, 0)
        cell = np.empty(n, dtype=np.uint32)
        d = np.empty(n, dtype=np.uint32)
        q = np.empty(n)
        self._lib.orc3d_links(self._h, ptr(cell, ctypes.c_uint32), ptr(d, ctypes.c_uint32), ptr(q), n)
        return BoundaryLinks(cell, d, q)

    def classes(self):
        out = np.empty(self._config.cells, dtype=np.uint8)
        self._lib.orc3d_classes(self._h, ptr(out, ctypes.c_uint8))
        return out


class Oracle2D:
    """CPU FP64 D2Q9 solver."""

    def __init__(self, config: Lbm2DConfig, mask):
        self._lib = load()
        self._config = config
        self._mask = mask
        self._c = config.to_c()
        solid = np.ascontiguousarray(mask.solid, dtype=np.uint8)
        buf = ctypes.create_string_buffer(512)
        h = self._lib.orc2d_create(ctypes.byref(self._c), ptr(solid, ctypes.c_uint8), buf, 512)
        if not h:
            raise OracleError(buf.value.decode())
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.orc2d_destroy(self._h)
            self._h = None

    def steps(self):
        return int(self._lib.orc2d_steps(self._h))

    def initialize(self, initial=None):
        if initial is None:
            self._lib.orc2d_initialize(self._h, None)
            return
        c = self._config
        if callable(initial):
            planes = np.zeros((3, c.cells))
            for y in range(c.ny):
                for x in range(c.nx):
                    planes[:, x + c.nx * y] = initial(x, y)
        else:
            planes = np.ascontiguousarray(initial, dtype=np.float64).reshape(3, c.cells)
        self._lib.orc2d_initialize(self._h, ptr(planes))

    def initialize_perturbed(self, seed, amplitude):
        self._lib.orc2d_initialize_perturbed(self._h, seed, amplitude)

    def step(self, count):
        self._lib.orc2d_step(self._h, int(count))

    def macroscopic(self) -> MacroFields2D:
        c = self._config
        out = np.empty((3, c.cells))
        self._lib.orc2d_macroscopic(self._h, *[ptr(out[k]) for k in range(3)])
        return MacroFields2D(c.nx, c.ny, out[0], out[1], out[2], np.asarray(self._mask.solid, dtype=np.uint8))

    def populations(self):
        f = np.empty(9 * self._config.cells)
        self._lib.orc2d_populations(self._h, ptr(f))
        return f

    def force(self) -> Force2D:
        v = np.zeros(3)
        self._lib.orc2d_force(self._h, ptr(v[0:1]), ptr(v[1:2]), ptr(v[2:3]))
        return Force2D(float(v[0]), float(v[1]), float(v[2]))

    def mass(self):
        return float(self._lib.orc2d_mass(self._h))

    def links(self) -> BoundaryLinks:
        n = self._lib.orc2d_links(self._h, None, None, 0)
        cell = np.empty(n, dtype=np.uint32)
        d = np.empty(n, dtype=np.uint32)
        self._lib.orc2d_links(self._h, ptr(cell, ctypes.c_uint32), ptr(d, ctypes.c_uint32), n)
        return BoundaryLinks(cell, d, np.full(n, 0.5))
