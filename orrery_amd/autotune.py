"""Autotuner of the step launch: which schedule, bulk launch shape (threads per block, cells
per thread) and bulk occupancy caps (workgroups per compute unit, A and B step) run a lattice
fastest on this device.

The library does the timing (lbm3d_autotune: every shape on a device copy of the lattice, which
is restored afterwards, so tuning changes no result). This module adds a JSON cache keyed by the
device and the lattice's storage and dimensions, so a shape is timed once per device and shape
class; lbm_batch_autotune tunes each storage group of a batch on a scratch copy of its largest
lattice.

Cache file: $ORRERY_LBM_TUNE_CACHE, else ~/.cache/orrery_amd/launch.json.
"""
from __future__ import annotations

import ctypes
import dataclasses
import json
import os

from . import abi
from .solver import LbmError, Lbm3DSolver, _check

_SCHEDULE_NAMES = {v: k for k, v in abi.SCHEDULE.items()}
_STORAGE_NAMES = {v: k for k, v in abi.STORAGE.items()}


@dataclasses.dataclass(frozen=True)
class LaunchChoice:
    schedule: str
    block: int
    cells_per_thread: int
    ms_per_step: float
    groups_a: int = 0  # occupancy caps (workgroups per compute unit), 0: none
    groups_b: int = 0

    @staticmethod
    def _from_c(c: abi.CLaunchChoice) -> "LaunchChoice":
        return LaunchChoice(_SCHEDULE_NAMES.get(int(c.schedule), str(c.schedule)), int(c.block),
                            int(c.cells_per_thread), float(c.ms_per_step), int(c.groups_a), int(c.groups_b))


def device_name() -> str:
    buf = ctypes.create_string_buffer(256)
    lib = abi.load()
    _check(lib, lib.lbm_device_name(buf, len(buf)))
    return buf.value.decode()


def cache_path() -> str:
    return os.environ.get("ORRERY_LBM_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "orrery_amd",
                                                                   "launch.json")


def _load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _store_cache(path: str, data: dict) -> None:
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def cache_key(solver: Lbm3DSolver, device: str | None = None) -> str:
    c = solver.config()
    return f"{device or device_name()}|{solver.storage}|{c.nx}x{c.ny}x{c.nz}"


def current(solver: Lbm3DSolver) -> tuple[str, int, int]:
    """(schedule, block, cells per thread) the solver steps with."""
    lib = abi.load()
    s, b, v = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    _check(lib, lib.lbm3d_get_launch(solver.handle, ctypes.byref(s), ctypes.byref(b), ctypes.byref(v)))
    return _SCHEDULE_NAMES.get(s.value, str(s.value)), b.value, v.value


def apply(solver: Lbm3DSolver, choice: LaunchChoice) -> None:
    solver.set_schedule(choice.schedule)
    solver.set_launch(choice.block, choice.cells_per_thread)
    solver.set_occupancy(choice.groups_a, choice.groups_b)


def tune(solver: Lbm3DSolver, steps: int = 10, serial_only: bool = False) -> tuple[LaunchChoice, list[LaunchChoice]]:
    """Times every candidate on the solver's lattice (restored afterwards) and applies the
    fastest; returns it and every candidate timed."""
    lib = abi.load()
    tried = (abi.CLaunchChoice * 64)()
    n = ctypes.c_size_t()
    best = abi.CLaunchChoice()
    _check(lib, lib.lbm3d_autotune(solver.handle, int(steps), 1 if serial_only else 0, tried, len(tried),
                                   ctypes.byref(n), ctypes.byref(best)))
    return LaunchChoice._from_c(best), [LaunchChoice._from_c(tried[k]) for k in range(min(n.value, len(tried)))]


def autotune(solver: Lbm3DSolver, steps: int = 10, cache: str | None = None, refresh: bool = False) -> LaunchChoice:
    """The cached choice for this device and lattice shape, or a fresh tuning stored in the
    cache; applied to the solver either way."""
    path = cache or cache_path()
    device = device_name()
    key = cache_key(solver, device)
    data = _load_cache(path)
    if not refresh and key in data:
        e = data[key]
        choice = LaunchChoice(e["schedule"], int(e["block"]), int(e["cells_per_thread"]), float(e["ms_per_step"]),
                              int(e.get("groups_a", 0)), int(e.get("groups_b", 0)))
        apply(solver, choice)
        return choice
    choice, _ = tune(solver, steps)
    data = _load_cache(path)  # another process may have written meanwhile
    data[key] = dataclasses.asdict(choice)
    _store_cache(path, data)
    return choice


def autotune_batch(batch, steps: int = 10) -> dict[str, LaunchChoice]:
    """lbm_batch_autotune: the bulk launch shape of each storage group of the batch, tuned on a
    scratch copy of its largest lattice and applied to the group."""
    lib = abi.load()
    best = (abi.CLaunchChoice * 4)()
    _check(lib, lib.lbm_batch_autotune(batch._b, int(steps), best))
    return {_STORAGE_NAMES[k]: LaunchChoice._from_c(best[k]) for k in range(4) if best[k].block > 0}


__all__ = ["LaunchChoice", "autotune", "autotune_batch", "tune", "current", "apply", "device_name", "cache_path",
           "cache_key", "LbmError"]
