"""Builds tests/adapter_check.cpp, the driver of the C++ drop-in (include/orrery_lbm_solver.hpp,
include/orrery_lbm_aero.hpp) through the reference's own solver interface. It compiles only
where the reference's headers are present; the built driver carries no reference file with it
and finds the library through its run path (orrery_amd/ beside tests/)."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REFERENCE = "/root/reference"
INCLUDES = ["src/aero/lbm3d/include", "src/aero/lbm/include", "src/aero/compute/include", "src/core/include",
            "src/geom/include", "src/aero/include"]
DRIVER = os.path.join(HERE, "_build", "adapter_check")


def possible() -> bool:
    return os.path.isdir(REFERENCE) and shutil.which("g++") is not None


def build(out: str = DRIVER) -> str:
    """Compiles the driver into `out`, linked to orrery_amd/liborrery_lbm.so by a run path
    relative to `out` when `out` is the in-tree default, else by the library's own folder."""
    from orrery_amd.abi import LIBRARY
    os.makedirs(os.path.dirname(out), exist_ok=True)
    lib_dir = os.path.dirname(LIBRARY)
    rpath = "$ORIGIN/../../orrery_amd" if os.path.abspath(out) == DRIVER else lib_dir
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include")]
    for inc in INCLUDES:
        cmd += ["-I", os.path.join(REFERENCE, inc)]
    cmd += [os.path.join(HERE, "adapter_check.cpp"), "-o", out, LIBRARY, f"-Wl,-rpath,{rpath}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out
