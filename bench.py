"""Throughput of the D3Q19 device solver: GLUPS and the HBM roofline of its step kernel.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: one D3Q19 BGK + Smagorinsky lattice of 512 x 256 x 256 nodes (33.5 M) in a channel with
a velocity inlet, a convective outlet and slip sides, around a voxelised body, FP32 populations
stored in place (A-A), started from a seeded +-1e-4 perturbation of the uniform inlet state. A step
is one lattice update of every node (bulk kernel + boundary kernel). `value` = nx ny nz K / time
in GLUPS, timed with HIP events on the solver's stream after W warm-up steps. The roofline object
prices the bulk kernel (the dominant one) at 152 algorithmic bytes per bulk node against MI355X's
8 TB/s; cpu_baseline is the oracle (the FP64 CPU restatement of the reference solver) on a
bounded number of steps of the same lattice on all host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # bytes/s, MI355X published


def workload(nx=512, ny=256, nz=256, diameter=64.0, reynolds=1000.0, u=0.05):
    from orrery_amd import Lbm3DConfig, Smagorinsky, Wall, WallKind, XBoundary, sphere_mask, tau_of_viscosity
    config = Lbm3DConfig(nx=nx, ny=ny, nz=nz, tau=tau_of_viscosity(u * diameter / reynolds),
                         x_boundary=XBoundary.inflow_outflow, y_low=Wall(WallKind.slip), y_high=Wall(WallKind.slip),
                         z_low=Wall(WallKind.slip), z_high=Wall(WallKind.slip), inlet_ux=u,
                         smagorinsky=Smagorinsky(True, 0.17))
    centre = (2.0 * diameter, 0.5 * (ny - 1) + 0.25, 0.5 * (nz - 1) + 0.25)
    return config, sphere_mask(nx, ny, nz, centre, 0.5 * diameter)


def time_storage(config, mask, storage, steps, warmup, per_kernel_steps):
    from orrery_amd import Lbm3DSolver, abi
    import ctypes
    lib = abi.load()
    solver = Lbm3DSolver(config, mask, storage)
    solver.initialize_perturbed(2024, 1e-4)
    solver.step(warmup)
    solver.synchronize()
    ms = ctypes.c_double()
    bulk = ctypes.c_double()
    bnd = ctypes.c_double()
    rc = lib.lbm3d_time_steps(solver.handle, steps, 0, ctypes.byref(ms), None, None)
    if rc:
        raise RuntimeError(lib.lbm_last_error().decode())
    total_ms = ms.value
    rc = lib.lbm3d_time_steps(solver.handle, per_kernel_steps, 1, ctypes.byref(ms), ctypes.byref(bulk),
                              ctypes.byref(bnd))
    if rc:
        raise RuntimeError(lib.lbm_last_error().decode())
    counts = [ctypes.c_size_t() for _ in range(4)]
    lib.lbm3d_counts(solver.handle, *[ctypes.byref(c) for c in counts])
    fx = solver.force().force[0]
    solver.close()
    return {
        "ms_per_step": total_ms / steps,
        "bulk_ms": bulk.value / per_kernel_steps,
        "boundary_ms": bnd.value / per_kernel_steps,
        "bulk_nodes": counts[0].value,
        "boundary_nodes": counts[1].value,
        "solid_nodes": counts[2].value,
        "links": counts[3].value,
        "force_x": fx,
    }


def cpu_baseline(config, mask, steps):
    """The oracle on the same lattice: one warm-up step, then `steps` timed steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import Oracle3D, load
    threads = os.cpu_count() or 1
    load().orc_set_threads(threads)
    oracle = Oracle3D(config, mask)
    oracle.initialize_perturbed(2024, 1e-4)
    oracle.step(1)
    t0 = time.perf_counter()
    oracle.step(steps)
    dt = time.perf_counter() - t0
    del oracle
    cells = config.cells
    return {"value": cells * steps / dt / 1e9, "unit": "GLUPS", "cores": threads, "kind": "port",
            "sample": f"{steps} FP64 steps of the same {config.nx}x{config.ny}x{config.nz} lattice after 1 warm-up "
                      f"step ({dt:.1f} s), OpenMP over rows, host CPU {platform.processor() or platform.machine()}"}


def copy_calibration(mib):
    """A plain device-to-device copy of `mib` MiB (torch's elementwise copy kernel), three times."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"bench: copied {mib} MiB three times (calibration kernel)", file=sys.stderr, flush=True)
    del src, dst
    torch.cuda.empty_cache()


TRAFFIC = os.path.join(ROOT, "profiles", "d3q19_bulk_f32_512x256x256.traffic.json")


def measured_traffic():
    """HBM bytes per launch of the f32 bulk kernel on this workload, from the two counter runs
    (FETCH_SIZE, WRITE_SIZE) summarised in profiles/, scaled by the copy-kernel calibration."""
    if not os.path.exists(TRAFFIC):
        return None
    with open(TRAFFIC) as f:
        return json.load(f)


def log(message):
    print(f"bench: {message}", file=sys.stderr, flush=True)


def polar_bench(args):
    """Batched Cd solves per hour: the wing-section sweep in one batch against the same cases one
    lattice at a time (host-paced force reads), same kernels, same convergence test."""
    from orrery_amd import polar
    cases = polar.wing_cases(args.lattices)
    log(f"polar: voxelising {len(cases)} wing sections")
    lattices = [polar.wing_lattice(c) for c in cases]
    log("polar: batched run")
    batched, tb = polar.run_batched(cases, lattices, log=log)
    pick = list(range(0, len(cases), max(1, args.sequential_every)))
    log(f"polar: one-at-a-time run of {len(pick)} cases")
    sequential, ts = polar.run_sequential([cases[k] for k in pick], [lattices[k] for k in pick], log=log)
    # The one-at-a-time rate on the cases it ran, scaled by their share of the batched work
    # (cell-steps), so a subset of cases gives a like-for-like estimate.
    work = {s.case.index: s.case.cells * s.steps for s in batched}
    share = sum(work[k] for k in pick) / sum(work.values())
    seq_total = ts / share
    by_case = {s.case.index: s for s in batched}
    same = all((s.steps, s.cd, s.cl) == (by_case[s.case.index].steps, by_case[s.case.index].cd,
                                         by_case[s.case.index].cl) for s in sequential)
    rate = len(batched) / tb * 3600.0
    line = {
        "metric": "batched Cd solves per hour (wing-section sweep, convergence |dCd| < 1e-4 over 500 steps)",
        "value": round(rate, 1), "unit": "solves/h", "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(cases)} NACA 0012 sections, dims in [96^3, 320x160x160], chord 48-96 cells, "
                               "AoA -4..18 deg, Re 500-2000, LES, interpolated wall, periodic span",
                   "cells": sum(c.cells for c in cases), "parallelism": "single GPU"},
        "batched_s": round(tb, 2), "sequential_s_estimated": round(seq_total, 2),
        "sequential_cases_timed": len(pick), "speedup": round(seq_total / tb, 3),
        "same_results_as_sequential": same,
        "converged": sum(s.converged for s in batched),
        "lift_slope_per_rad": round(polar.lift_slope(batched), 4),
        "glups_batched": round(sum(work.values()) / tb / 1e9, 2),
    }
    print(json.dumps(line), flush=True)
    return 0


def serve_bench(args):
    from orrery_amd import serve
    result = serve.serve_bench(args.requests, log=log)
    print(json.dumps(result), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--storages", default="f32,f16s,bf16s")
    ap.add_argument("--workload", choices=["channel", "polar", "serve"], default="channel",
                    help="channel: the 512x256x256 lattice (the headline line); polar: the batched wing-section "
                         "sweep, batched against one at a time; serve: the continuous-batching server")
    ap.add_argument("--lattices", type=int, default=48, help="polar: number of wing cases")
    ap.add_argument("--sequential-every", type=int, default=1,
                    help="polar: time the one-at-a-time baseline on every k-th case (k > 1 bounds the run)")
    ap.add_argument("--requests", type=int, default=32, help="serve: number of requests")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device with a plain copy kernel (3 launches), the "
                         "known-size reference for the HBM byte counters of a profiler run")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(independent lattices shard across GPUs by HBM budget, one lattice stays on one GPU)",
              file=sys.stderr)
        return 2

    if args.workload == "polar":
        return polar_bench(args)
    if args.workload == "serve":
        return serve_bench(args)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
    config, mask = workload()
    per_kernel_steps = min(200, max(20, args.steps // 5))
    results = {}
    for storage in args.storages.split(","):
        print(f"bench: timing {storage} ({args.warmup} warm-up, {args.steps} timed steps)", file=sys.stderr,
              flush=True)
        results[storage] = time_storage(config, mask, storage, args.steps, args.warmup, per_kernel_steps)
    main_r = results["f32"]
    cells = config.cells
    glups = cells / (main_r["ms_per_step"] * 1e-3) / 1e9
    bytes_per_node = 2 * 19 * 4
    achieved = bytes_per_node * main_r["bulk_nodes"] / (main_r["bulk_ms"] * 1e-3)
    cpu = None
    if args.cpu_steps > 0:
        print("bench: timing the CPU baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(config, mask, args.cpu_steps)
    traffic = measured_traffic()
    line = {
        "metric": "D3Q19 GLUPS (FP32 A-A, BGK+Smagorinsky) and % HBM roofline",
        "value": round(glups, 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "D3Q19 512x256x256 channel, sphere D64 (Re 1000, u 0.05, Smagorinsky C 0.17), "
                        "velocity inlet, convective outlet, slip sides, FP32 A-A in place",
            "cells": cells,
            "bulk_nodes": main_r["bulk_nodes"],
            "boundary_nodes": main_r["boundary_nodes"],
            "solid_nodes": main_r["solid_nodes"],
            "links": main_r["links"],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved / 1e9, 1),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4),
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "traffic_over_algorithmic": round(traffic["bytes_per_launch"] / (bytes_per_node * main_r["bulk_nodes"]), 4)
            if traffic else None,
            "kernel": "d3q19_bulk<StoreF32>",
            "bulk_ms": round(main_r["bulk_ms"], 4),
            "boundary_ms": round(main_r["boundary_ms"], 4),
            "step_roofline_frac": round(bytes_per_node * cells / (main_r["ms_per_step"] * 1e-3) / HBM_PEAK, 4),
        },
        "cpu_baseline": cpu,
        "storage_variants": {
            k: {"glups": round(cells / (r["ms_per_step"] * 1e-3) / 1e9, 3),
                "bytes_per_node": 2 * 19 * {"f32": 4, "f16s": 2, "bf16s": 2, "f64": 8}[k],
                "hbm_frac_bulk": round(2 * 19 * {"f32": 4, "f16s": 2, "bf16s": 2, "f64": 8}[k] * r["bulk_nodes"]
                                       / (r["bulk_ms"] * 1e-3) / HBM_PEAK, 4),
                "ms_per_step": round(r["ms_per_step"], 4)}
            for k, r in results.items()
        },
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
