"""Build recipe of liborrery_lbm.so (HIP, gfx950), compiled in-tree next to this file.

    python -m orrery_amd.build

Every HIP source is compiled for gfx950 without floating-point contraction (the binary64
kernels reproduce the CPU solver's rounding) and without packed-FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["lbm_setup.hip", "lbm3d_kernels.hip", "lbm3d_abi.hip", "lbm2d_kernels.hip", "lbm2d_abi.hip",
           "lbm_batch.hip", "geom_device.hip"]
# Host-only sources, compiled by the host C++ compiler with the floating-point policy of the
# reference's geometry and history code (-O2, no contraction, no fast-math) so that their FP64
# decisions (voxel masks) match it bit for bit.
HOST_SOURCES = ["geom.cpp", "lbm_history.cpp", "aero3d.cpp", "aero2d.cpp"]
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-pthread"]
OUTPUT = os.path.join(HERE, "liborrery_lbm.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def sources() -> list[str]:
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = False, extra: list[str] | None = None, jobs: int | None = None) -> str:
    """Compiles the library; returns its path. Raises CalledProcessError on failure.

    Every source is compiled to its own object (each HIP translation unit carries and registers
    its own gfx950 code object), `jobs` at a time, then the objects are linked."""
    out_dir = os.path.join(CSRC, "_build")
    os.makedirs(out_dir, exist_ok=True)
    commands = []
    objects = []
    for src in HOST_SOURCES:
        obj = os.path.join(out_dir, src.replace(".cpp", ".o"))
        commands.append([CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objects.append(obj)
    for src in sources():
        obj = os.path.join(out_dir, os.path.basename(src).replace(".hip", ".hip.o"))
        commands.append([HIPCC, *FLAGS, *(extra or []), "-c", src, "-o", obj])
        objects.append(obj)
    workers = jobs or max(1, min(len(commands), os.cpu_count() or 1, 8))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for f in [pool.submit(_run, c, verbose) for c in commands]:
            f.result()
    tmp = OUTPUT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objects, "-lpthread"], verbose)
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
