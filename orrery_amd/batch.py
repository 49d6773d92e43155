"""Batched lattices: many resident D3Q19 lattices advanced together by the library's batch engine
(lbm_batch_* in include/orrery_lbm.h), with per-lattice force histories."""
from __future__ import annotations

import ctypes

import numpy as np

from . import abi
from .config import Lbm3DConfig
from .masks import SolidMask3D
from .solver import LbmError, _check


class BatchLattice:
    """A lattice owned by a batch: the Lbm3DSolver queries on its borrowed handle."""

    def __init__(self, batch: "LatticeBatch", lattice_id: int, config: Lbm3DConfig, mask: SolidMask3D, storage: str):
        self.batch = batch
        self.id = lattice_id
        self.config = config
        self.mask = mask
        self.storage = storage

    def _h(self):
        h = ctypes.c_void_p()
        _check(self.batch._lib, self.batch._lib.lbm_batch_lattice(self.batch._b, self.id, ctypes.byref(h)))
        return h

    def steps(self) -> int:
        return int(self.batch._lib.lbm3d_steps(self._h()))

    def populations(self) -> np.ndarray:
        f = np.empty(19 * self.config.cells)
        _check(self.batch._lib, self.batch._lib.lbm3d_populations(self._h(), abi.ptr(f, ctypes.c_double)))
        return f

    def force(self):
        f = np.zeros(3)
        m = np.zeros(3)
        _check(self.batch._lib, self.batch._lib.lbm3d_force(self._h(), abi.ptr(f, ctypes.c_double),
                                                            abi.ptr(m, ctypes.c_double)))
        return tuple(f), tuple(m)

    def history(self) -> tuple[np.ndarray, np.ndarray]:
        """(time [n], force [n, 6]) of the samples taken so far."""
        lib = self.batch._lib
        n = ctypes.c_size_t()
        _check(lib, lib.lbm_batch_history(self.batch._b, self.id, None, None, 0, ctypes.byref(n)))
        t = np.empty(n.value)
        f = np.empty((n.value, 6))
        _check(lib, lib.lbm_batch_history(self.batch._b, self.id, abi.ptr(t, ctypes.c_double),
                                          abi.ptr(f, ctypes.c_double), n.value, ctypes.byref(n)))
        return t, f

    def set_inlet_velocity(self, v) -> None:
        v = np.asarray(v, dtype=np.float64)
        _check(self.batch._lib, self.batch._lib.lbm_batch_set_inlet(self.batch._b, self.id, abi.ptr(v, ctypes.c_double)))


class LatticeBatch:
    """Resident lattices admitted against an HBM budget (bytes; 0 = none), sampled every
    `sample_every` steps."""

    def __init__(self, budget_bytes: int = 0, sample_every: int = 10):
        lib = abi.load()
        self._lib = lib
        b = ctypes.c_void_p()
        _check(lib, lib.lbm_batch_create(int(budget_bytes), int(sample_every), ctypes.byref(b)))
        self._b = b
        self.sample_every = sample_every
        self.lattices: dict[int, BatchLattice] = {}

    def close(self):
        if getattr(self, "_b", None):
            self._lib.lbm_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def required_bytes(config: Lbm3DConfig, mask: SolidMask3D, storage: str = "f32") -> int:
        lib = abi.load()
        c = config.to_c()
        solid = np.ascontiguousarray(mask.solid, dtype=np.uint8)
        d = None if mask.distance is None else np.ascontiguousarray(mask.distance)
        n = ctypes.c_size_t()
        _check(lib, lib.lbm_batch_required_bytes(ctypes.byref(c), abi.ptr(solid, ctypes.c_uint8),
                                                 abi.ptr(d, ctypes.c_double), abi.STORAGE[storage], ctypes.byref(n)))
        return n.value

    def add(self, config: Lbm3DConfig, mask: SolidMask3D, storage: str = "f32", seed: int = 0,
            amplitude: float = 0.0) -> BatchLattice:
        c = config.to_c()
        solid = np.ascontiguousarray(mask.solid, dtype=np.uint8)
        d = None if mask.distance is None else np.ascontiguousarray(mask.distance)
        i = ctypes.c_int32()
        _check(self._lib, self._lib.lbm_batch_add(self._b, ctypes.byref(c), abi.ptr(solid, ctypes.c_uint8),
                                                  abi.ptr(d, ctypes.c_double), abi.STORAGE[storage], seed, amplitude,
                                                  ctypes.byref(i)))
        lat = BatchLattice(self, i.value, config.copy(), mask, storage)
        self.lattices[i.value] = lat
        return lat

    def remove(self, lattice: BatchLattice) -> None:
        _check(self._lib, self._lib.lbm_batch_remove(self._b, lattice.id))
        self.lattices.pop(lattice.id, None)

    def run(self, steps: int) -> None:
        _check(self._lib, self._lib.lbm_batch_run(self._b, int(steps)))

    def run_timed(self, steps: int) -> dict:
        """run(steps) with each kernel between HIP events on the batch's stream (no graph
        replay; the lattices advance exactly as run() advances them). Returns the mean step
        time and, per storage format present, the bulk and boundary kernel times per step and
        their node counts, plus the force sample time."""
        t = abi.CBatchTiming()
        _check(self._lib, self._lib.lbm_batch_run_timed(self._b, int(steps), ctypes.byref(t)))
        groups = {}
        for name, code in abi.STORAGE.items():
            if t.bulk_nodes[code] > 0 or t.bulk_ms[code] > 0:
                groups[name] = {"bulk_ms": t.bulk_ms[code], "boundary_ms": t.boundary_ms[code],
                                "bulk_nodes": int(t.bulk_nodes[code]), "boundary_nodes": int(t.boundary_nodes[code])}
        return {"step_ms": t.step_ms, "force_ms": t.force_ms, "groups": groups}

    def set_launch(self, storage: str, block: int, cells_per_thread: int) -> None:
        _check(self._lib, self._lib.lbm_batch_set_launch(self._b, abi.STORAGE[storage], block, cells_per_thread))

    def set_occupancy(self, storage: str, groups_per_cu: int) -> None:
        """Caps the bulk launch of one storage group at `groups_per_cu` resident workgroups per
        compute unit (0: no cap); results do not change."""
        _check(self._lib, self._lib.lbm_batch_set_occupancy(self._b, abi.STORAGE[storage], int(groups_per_cu)))

    def bytes(self) -> tuple[int, int]:
        u = ctypes.c_size_t()
        t = ctypes.c_size_t()
        _check(self._lib, self._lib.lbm_batch_bytes(self._b, ctypes.byref(u), ctypes.byref(t)))
        return u.value, t.value

    def __len__(self):
        return len(self.lattices)


__all__ = ["LatticeBatch", "BatchLattice", "LbmError"]
