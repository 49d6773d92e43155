"""Throughput of the D3Q19 device solver: GLUPS and the HBM roofline of its step kernel.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: one D3Q19 BGK + Smagorinsky lattice of 512 x 256 x 256 nodes (33.5 M) in a channel with
a velocity inlet, a convective outlet and slip sides, around a voxelised Apollo-type capsule
(diameter 64 cells, pitched 20 degrees, Re 1000), FP32 populations stored in place (A-A), started
from a seeded +-1e-4 perturbation of the uniform inlet state. A step is one lattice update of
every node (bulk kernel + boundary kernel). `value` = nx ny nz K / time in GLUPS, timed with HIP
events on the solver's stream after W warm-up steps. The roofline object prices the bulk kernel
(the dominant one) at 152 algorithmic bytes per node it updates (bulk nodes, and in A steps the
side nodes of the slip faces, which a side kernel updates in B steps) against MI355X's 8 TB/s;
cpu_baseline is
the oracle (the FP64 CPU restatement of the reference solver) on a bounded number of steps of the
same lattice on all host cores. Each storage format (f32, f16s, bf16s) is also run on to step
2000 and its drag force and velocity field compared with the f64 storage's (the FP64 oracle
arithmetic on the device) at that step.

The same line carries the metric's second half, measured in the same run:
- `batched`: the 48 NACA 0012 wing sections (96^3 .. 320 x 160 x 160, FP32) advanced in one
  batch to convergence, against the same solves one lattice at a time (every 8th timed, scaled
  by cell-steps): converged solves per hour both ways, the speed-up, and the Cd of the 8
  smallest converged sections against the FP64 device lattice over the same samples;
- `served`: 16 requests with seeded Poisson arrivals at 4/s (sphere, cone, cylinder, capsule,
  wedge; 64^3 .. 384 x 192 x 192; f32 or f16s) through the continuous-batching server:
  requests/s, converged solves per hour, p50/p99 latency from arrival, the one-at-a-time
  estimate, and the Cd of the 8 smallest converged requests against the FP64 device lattice.

Other workloads (--workload): polar (the batched section alone), serve (the served section
alone), d2q9 (the 400 x 160 cylinder channel), sphere (the 256 x 128 x 128 sphere channel),
reference (the reference's own GPU benchmarks).
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


def workload(nx=512, ny=256, nz=256, diameter=64.0, pitch_deg=20.0, reynolds=1000.0, u=0.05):
    """The 512 x 256 x 256 channel around a voxelised Apollo-type capsule (heat-shield diameter 64
    cells, heat shield facing the flow, pitched 20 degrees about y) at Re 1000 on the diameter."""
    import math
    from orrery_amd import Lbm3DConfig, Smagorinsky, Wall, WallKind, XBoundary, geometry, tau_of_viscosity
    from orrery_amd.polar import place_body
    from orrery_amd.serve import body_mesh
    config = Lbm3DConfig(nx=nx, ny=ny, nz=nz, tau=tau_of_viscosity(u * diameter / reynolds),
                         x_boundary=XBoundary.inflow_outflow, y_low=Wall(WallKind.slip), y_high=Wall(WallKind.slip),
                         z_low=Wall(WallKind.slip), z_high=Wall(WallKind.slip), inlet_ux=u,
                         smagorinsky=Smagorinsky(True, 0.17))
    centre = (2.0 * diameter, 0.5 * (ny - 1) + 0.25, 0.5 * (nz - 1) + 0.25)
    mesh = body_mesh("capsule", diameter).transformed(geometry.rotation_y(math.radians(pitch_deg)), (0.0, 0.0, 0.0),
                                                       centre)
    return config, place_body(mesh, nx, ny, nz, interpolated=False, device=True)


def time_storage(config, mask, storage, steps, warmup, per_kernel_steps, compare_steps=0, ref=None, tune_steps=0,
                 groups=(0, 0)):
    """Times `steps` steps after `warmup` (one event pair) and `per_kernel_steps` more per kernel,
    then steps on to `compare_steps` and reads the force and velocity there: returned as the
    reference fields for the f64 storage, compared with `ref` (the f64 run's) otherwise. With
    `tune_steps`, the autotuner picks the schedule and launch shape after the warm-up (the
    lattice is restored after tuning)."""
    import numpy as np
    from orrery_amd import Lbm3DSolver, abi, autotune
    import ctypes
    lib = abi.load()
    solver = Lbm3DSolver(config, mask, storage)
    solver.initialize_perturbed(2024, 1e-4)
    solver.step(warmup)
    solver.synchronize()
    launch = None
    if tune_steps > 0:
        best, tried = autotune.tune(solver, tune_steps)
        launch = {"schedule": best.schedule, "block": best.block, "cells_per_thread": best.cells_per_thread,
                  "groups_per_cu": [best.groups_a, best.groups_b],
                  "tuned_ms_per_step": round(best.ms_per_step, 4),
                  "default_ms_per_step": round(next(t.ms_per_step for t in reversed(tried) if t.schedule == "serial" and t.block == 256
                                                    and t.cells_per_thread == (2 if storage in ("f16s", "bf16s") else 1)
                                                    and t.groups_a == 0 and t.groups_b == 0), 4),
                  "candidates": len(tried)}
    else:
        solver.set_occupancy(*groups)
        sched, block, cpt = autotune.current(solver)
        launch = {"schedule": sched, "block": block, "cells_per_thread": cpt, "groups_per_cu": list(solver.occupancy())}
    ms = ctypes.c_double()
    rc = lib.lbm3d_time_steps(solver.handle, steps, 0, ctypes.byref(ms), None, None)
    if rc:
        raise RuntimeError(lib.lbm_last_error().decode())
    total_ms = ms.value
    kernels = (ctypes.c_double * 6)()
    if lib.lbm3d_time_kernels(solver.handle, per_kernel_steps, kernels):
        raise RuntimeError(lib.lbm_last_error().decode())
    counts = [ctypes.c_size_t() for _ in range(4)]
    lib.lbm3d_counts(solver.handle, *[ctypes.byref(c) for c in counts])
    split = stream_counts(lib, solver.handle)
    out = {
        "launch": launch,
        "ms_per_step": total_ms / steps,
        # mean over A and B steps of each kernel; the side kernel runs in B steps only
        "bulk_ms": (kernels[0] + kernels[1]) / 2,
        "bulk_ms_a_b": [round(kernels[0], 4), round(kernels[1], 4)],
        "side_ms_b": kernels[3],
        "boundary_ms": (kernels[4] + kernels[5]) / 2,
        "bulk_nodes": counts[0].value,
        "boundary_nodes": counts[1].value,
        "solid_nodes": counts[2].value,
        "links": counts[3].value,
        # the bulk kernel updates bulk and side nodes in A steps and bulk nodes in B steps; the
        # side kernel side nodes in B steps; the boundary kernel the listed boundary nodes
        "bulk_kernel_nodes": split["bulk"] + split["side"] / 2,
        "side_nodes": split["side"],
        "listed_nodes": split["listed"],
    }
    if compare_steps > 0:
        solver.step(max(0, compare_steps - solver.steps()))
        fx = solver.force().force[0]
        m = solver.macroscopic()
        fluid = np.asarray(m.solid) == 0
        velocity = np.stack([m.velocity_x[fluid], m.velocity_y[fluid], m.velocity_z[fluid]])
        del m
        out["compared_at_step"] = solver.steps()
        out["force_x"] = fx
        if ref is None:
            out["fields"] = velocity
        else:
            out["force_x_rel_error_vs_f64"] = abs(fx / ref["force_x"] - 1.0)
            out["velocity_l2_vs_f64"] = float(np.sqrt(((velocity - ref["fields"]) ** 2).sum() /
                                                      (ref["fields"] ** 2).sum()))
    solver.close()
    return out


def stream_counts(lib, handle):
    """lbm3d_stream_counts: nodes the bulk kernel streams (bulk, side) and the boundary kernel's."""
    import ctypes
    v = [ctypes.c_size_t() for _ in range(3)]
    if lib.lbm3d_stream_counts(handle, *[ctypes.byref(x) for x in v]):
        raise RuntimeError(lib.lbm_last_error().decode())
    return {"bulk": v[0].value, "side": v[1].value, "listed": v[2].value}


def cpu_baseline(config, mask, steps, seconds=0.0):
    """The oracle on the same lattice: one warm-up step, then `steps` timed steps, repeated until
    at least `seconds` have passed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import Oracle3D, load
    threads = os.cpu_count() or 1
    load().orc_set_threads(threads)
    oracle = Oracle3D(config, mask)
    oracle.initialize_perturbed(2024, 1e-4)
    oracle.step(1)
    t0 = time.perf_counter()
    done = 0
    while True:
        oracle.step(steps)
        done += steps
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    steps = done
    del oracle
    cells = config.cells
    return {"value": cells * steps / dt / 1e9, "unit": "GLUPS", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{steps} FP64 steps of the same {config.nx}x{config.ny}x{config.nz} lattice after 1 warm-up "
                      f"step ({dt:.1f} s), OpenMP over rows, host CPU {cpu_model()}"}


def cpu_model():
    """The host CPU's model name (lscpu's "Model name", from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for row in f:
                if row.startswith("model name"):
                    return row.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


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


def batched_section(n_lattices, sequential_every, fp64_checks, storage="f32"):
    """The batched wing-section sweep (48 NACA 0012 sections at 96^3 .. 320x160x160) run to
    convergence in one batch, against the same solves one lattice at a time (host-paced force
    reads, same kernels and convergence test), timed on every `sequential_every`-th case and
    scaled by work (cell-steps); the Cd of the `fp64_checks` smallest converged sections against
    the FP64 device lattice over the same samples."""
    from orrery_amd import polar
    cases = polar.wing_cases(n_lattices)
    log(f"batched: voxelising {len(cases)} wing sections on the device")
    lattices = [polar.wing_lattice(c, device=True) for c in cases]
    timing = {}
    batched, tb = polar.run_batched(cases, lattices, storage, log=log, timing=timing)
    kernels = {"step_ms": round(timing["step_ms"], 4), "force_sample_ms": round(timing["force_ms"], 4)}
    for st, g in timing["groups"].items():
        sb = {"f32": 4, "f16s": 2, "bf16s": 2, "f64": 8}[st]
        kernels[st] = {"bulk_ms": round(g["bulk_ms"], 4), "boundary_ms": round(g["boundary_ms"], 4),
                       "bulk_nodes": g["bulk_nodes"], "boundary_nodes": g["boundary_nodes"],
                       "bulk_gbps_algorithmic": round(38 * sb * g["bulk_nodes"] / (g["bulk_ms"] * 1e-3) / 1e9, 1),
                       "bulk_frac_of_8tbs": round(38 * sb * g["bulk_nodes"] / (g["bulk_ms"] * 1e-3) / HBM_PEAK, 4)}
    pick = list(range(0, len(cases), max(1, sequential_every)))
    log(f"batched: one-at-a-time run of {len(pick)} cases")
    sequential, ts = polar.run_sequential([cases[k] for k in pick], [lattices[k] for k in pick], storage, log=log)
    log(f"batched: the same {len(pick)} cases on the reference adapter's schedule (full field every 200 steps)")
    _, ts_ref = polar.run_sequential([cases[k] for k in pick], [lattices[k] for k in pick], storage, log=log,
                                     field_every=200)
    work = {s.case.index: s.case.cells * s.steps for s in batched}
    share = sum(work[k] for k in pick) / sum(work.values())
    seq_total = ts / share
    seq_ref_total = ts_ref / share
    seq_glups = sum(work[k] for k in pick) / ts / 1e9
    by_case = {s.case.index: s for s in batched}
    same = all((s.steps, s.cd, s.cl) == (by_case[s.case.index].steps, by_case[s.case.index].cd,
                                         by_case[s.case.index].cl) for s in sequential)
    done = [s for s in batched if s.converged]
    errors = []
    for s in sorted(done, key=lambda s: s.case.cells)[:fp64_checks]:
        cd64, cl64 = polar.fp64_coefficients(s.case, lattices[s.case.index], s.steps, s.mean_from)
        errors.append({"case": s.case.index, "cells": s.case.cells, "steps": s.steps, "kind": s.kind,
                       "cd": round(s.cd, 6), "cd_fp64": round(cd64, 6), "cd_rel_error": float(f"{abs(s.cd / cd64 - 1):.3e}")})
        log(f"batched: case {s.case.index} ({s.kind}) Cd {s.cd:.6f} vs FP64 {cd64:.6f}")
    kinds = {}
    for s in done:
        kinds[s.kind] = kinds.get(s.kind, 0) + 1
    plain_fixed = sum(c.cells * polar.MAX_STEPS for c in cases) / (seq_glups * 1e9)
    return {
        "workload": f"{len(cases)} NACA 0012 sections, dims in [96^3, 320x160x160], chord 48-96 cells, AoA -4..18 deg, "
                    f"Re 500-2000, LES, interpolated wall, periodic span, {storage}",
        "cells": sum(c.cells for c in cases), "converged": len(done), "converged_kinds": kinds,
        "kernels_first_chunk": kernels,
        "batched_s": round(tb, 2), "converged_solves_per_hour": round(len(done) / tb * 3600.0, 1),
        "sequential_s_estimated": round(seq_total, 2), "sequential_cases_timed": len(pick),
        "sequential_converged_solves_per_hour": round(len(done) / seq_total * 3600.0, 1),
        "speedup": round(seq_total / tb, 3), "same_results_as_sequential": same,
        "glups_batched": round(sum(work.values()) / tb / 1e9, 2), "glups_sequential": round(seq_glups, 2),
        "sequential_reference_schedule_s_estimated": round(seq_ref_total, 2),
        "speedup_vs_reference_schedule": round(seq_ref_total / tb, 3),
        "sequential_fixed_duration_s_estimated": round(plain_fixed, 1),
        "speedup_vs_fixed_duration": round(plain_fixed / tb, 3),
        "lift_slope_per_rad": round(polar.lift_slope(batched), 4),
        "cd_vs_fp64": errors,
        "cd_rel_error_vs_fp64_max": max(e["cd_rel_error"] for e in errors) if errors else None,
        "cd_within_1pct_of_fp64": f"{sum(e['cd_rel_error'] <= 0.01 for e in errors)}/{len(errors)}",
    }


def served_section(n_requests, rate, fp64_checks, sequential_every):
    """The continuous-batching server on `n_requests` seeded Poisson arrivals: throughput,
    latency percentiles, converged solves per hour, the same requests one at a time (every
    `sequential_every`-th, timed alone to the same step count, scaled by work), and the Cd of
    the `fp64_checks` smallest converged requests against the FP64 device lattice over the same
    samples."""
    import numpy as np
    from orrery_amd import Lbm3DSolver, serve
    reqs = serve.requests(n_requests, rate)
    server = serve.Server()
    served, wall = server.run(reqs, log=log)
    lat = np.array([s.latency for s in served])
    queued = np.array([s.queued for s in served])
    first = min(r.arrival for r in reqs)
    last = max(s.finished for s in served)
    done = [s for s in served if s.converged]
    # one request at a time: its lattice alone, host-paced samples, to the step it was served at
    pick = sorted(served, key=lambda s: s.request.id)[::max(1, sequential_every)]
    log(f"served: one-at-a-time run of {len(pick)} requests")
    seq_stepping = 0.0  # stepping and host-paced force reads only, not the lattice builds
    for s in pick:
        config, mask = serve.request_lattice(s.request)
        solver = Lbm3DSolver(config, mask, s.request.storage)
        solver.initialize_perturbed(server.seed, 1e-4)
        t1 = time.perf_counter()
        while solver.steps() < s.steps:
            solver.step(serve.SAMPLE_EVERY)
            solver.force()
        seq_stepping += time.perf_counter() - t1
        solver.close()
    work_all = sum(s.request.cells * s.steps for s in served)
    work_pick = sum(s.request.cells * s.steps for s in pick)
    seq_total = seq_stepping * work_all / work_pick
    span = last - first
    errors = []
    for s in sorted(done, key=lambda s: s.request.cells)[:fp64_checks]:
        ref = serve.fp64_reference_cd(s.request, s.steps, mean_from=s.mean_from)
        errors.append({"request": s.request.id, "cells": s.request.cells, "storage": s.request.storage,
                       "steps": s.steps, "kind": s.kind, "cd": round(s.cd, 6), "cd_fp64": round(ref, 6),
                       "cd_rel_error": float(f"{abs(s.cd / ref - 1):.3e}")})
        log(f"served: request {s.request.id} {s.request.storage} ({s.kind}) Cd {s.cd:.6f} vs FP64 {ref:.6f}")
    return {
        "workload": f"{n_requests} requests, Poisson lambda {rate}/s, bodies {'/'.join(serve.BODIES)}, dims in "
                    "[64^3, 384x192x192], AoA 0-30 deg, Re 200-3000, f32 or f16s per request, HBM budget 240 GB",
        "served": len(served), "converged": len(done), "wall_s": round(wall, 2),
        "requests_per_s": round(len(served) / span, 4),
        "converged_solves_per_hour": round(len(done) / span * 3600.0, 1),
        "latency_p50_s": round(float(np.percentile(lat, 50)), 3), "latency_p99_s": round(float(np.percentile(lat, 99)), 3),
        "queued_p50_s": round(float(np.percentile(queued, 50)), 3), "queued_p99_s": round(float(np.percentile(queued, 99)), 3),
        "gpu_busy_fraction": round(server.busy / wall, 3), "offered_per_s": rate,
        "sequential_requests_timed": len(pick), "work_share_timed": round(work_pick / work_all, 4),
        "sequential_s_estimated": round(seq_total, 2),
        "sequential_converged_solves_per_hour": round(len(done) / seq_total * 3600.0, 1),
        "speedup": round(seq_total / span, 3),
        "glups_served": round(work_all / span / 1e9, 2), "glups_sequential": round(work_pick / seq_stepping / 1e9, 2),
        "cd_vs_fp64": errors,
        "cd_rel_error_vs_fp64_max": max(e["cd_rel_error"] for e in errors) if errors else None,
        "cd_within_1pct_of_fp64": f"{sum(e['cd_rel_error'] <= 0.01 for e in errors)}/{len(errors)}",
    }


def polar_bench(args):
    """Batched Cd solves per hour of the wing-section sweep alone (the `batched` section of the
    default line), with the comparison table as CSV on request."""
    from orrery_amd import experiment, polar
    section = batched_section(args.lattices, args.sequential_every, args.fp64_checks)
    line = {"metric": "batched Cd solves per hour (wing-section sweep)", "value": section["converged_solves_per_hour"],
            "unit": "solves/h", "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic", "config": {"workload": section["workload"], "parallelism": "single GPU"},
            "batched": section}
    if args.csv:
        cases = polar.wing_cases(args.lattices)
        log("polar: the comparison table needs the solves; rerunning the batch for the CSV")
        solves, _ = polar.run_batched(cases, [polar.wing_lattice(c, device=True) for c in cases], log=log)
        with open(args.csv, "w", newline="") as f:
            f.write(polar.comparison_csv(solves))
        line["lift_slope_all_angles_per_rad"] = experiment.compare_by_geometry(polar.records(solves))[0].lift_slope_per_rad
    print(json.dumps(line), flush=True)
    return 0


D2Q9_SAMPLE = 10


def d2q9_channel():
    """400 x 160 channel, cylinder D 32 centred at (80, 80), inlet u 0.05, Re 100, zero-gradient
    outlet, slip top and bottom, halfway bounce-back."""
    from orrery_amd import Lbm2DConfig, Outlet, Wall, WallKind, XBoundary, circle_mask, tau_of_viscosity
    u, d, re = 0.05, 32.0, 100.0
    config = Lbm2DConfig(nx=400, ny=160, tau=tau_of_viscosity(u * d / re), x_boundary=XBoundary.inflow_outflow,
                         outlet=Outlet.extrapolation, bottom=Wall(WallKind.slip), top=Wall(WallKind.slip), inlet_ux=u)
    return config, circle_mask(400, 160, 80.0, 80.0, 0.5 * d), u, d


def d2q9_run(solver, steps, u, d):
    """Steps with a force sample every D2Q9_SAMPLE steps; returns (time, Cd, Cl) and the wall time."""
    import numpy as np
    q = 0.5 * u * u * d
    t, cd, cl = [], [], []
    t0 = time.perf_counter()
    for k in range(steps // D2Q9_SAMPLE):
        solver.step(D2Q9_SAMPLE)
        f = solver.force()
        t.append(float((k + 1) * D2Q9_SAMPLE))
        cd.append(f.x / q)
        cl.append(f.y / q)
    return (np.array(t), np.array(cd), np.array(cl)), time.perf_counter() - t0


def d2q9_summary(series, u, d, steps):
    """Cd mean and Cl rms over the second half of the configured steps; the Strouhal number of the
    lift oscillation after them (None when no settled oscillation is found)."""
    import numpy as np
    from orrery_amd import history
    t, cd, cl = series
    half = t <= steps
    late = (t > steps // 2) & half
    osc = history.dominant_oscillation(t, cl, float(steps))
    return {"cd": float(history.mean_since(t[half], cd[half], steps // 2)),
            "cl_rms": float(np.sqrt((cl[late] ** 2).mean())),
            "strouhal": float(osc.frequency * d / u) if osc else None,
            "shedding_cycles": int(osc.cycles) if osc else 0}


def d2q9_bench(args):
    """The plain D2Q9 cylinder channel on the device (FP32 and FP64 storage) and on the host
    (the oracle), from the same seeded start, with a force sample every 10 steps."""
    from orrery_amd import Lbm2DSolver
    config, mask, u, d = d2q9_channel()
    steps = args.d2q9_steps
    total = 2 * steps  # the Strouhal number is measured after the configured steps
    cells = config.nx * config.ny
    out = {}
    for storage in ("f32", "f64"):
        s = Lbm2DSolver(config, mask, storage)
        s.initialize_perturbed(2024, 1e-4)
        log(f"d2q9: {storage}, {total} steps with a force sample every {D2Q9_SAMPLE}")
        series, wall = d2q9_run(s, total, u, d)
        out[storage] = dict(d2q9_summary(series, u, d, steps), mlups_sampled=cells * total / wall / 1e6,
                            wall_s=wall)
        # Stepping alone: graph replays against launches one by one, and the one-launch step
        # against the bulk and boundary kernels launched separately.
        for name, graphs, schedule in (("mlups_graphs", True, "fused"), ("mlups_eager", False, "fused"),
                                       ("mlups_two_kernels_graphs", True, "serial")):
            s.set_graphs(graphs)
            s.set_schedule(schedule)
            s.step(args.steps)  # captures the graph the timed call replays
            s.synchronize()
            t0 = time.perf_counter()
            s.step(args.steps)
            s.synchronize()
            out[storage][name] = cells * args.steps / (time.perf_counter() - t0) / 1e6
        s.set_schedule("fused")
        s.close()
    cpu = None
    if args.cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle import Oracle2D, load
        threads = os.cpu_count() or 1
        load().orc_set_threads(threads)
        o = Oracle2D(config, mask)
        o.initialize_perturbed(2024, 1e-4)
        log(f"d2q9: the oracle, {total} steps on {threads} threads")
        series, wall = d2q9_run(o, total, u, d)
        ref = d2q9_summary(series, u, d, steps)
        cpu = {"value": round(cells * total / wall / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
               "sample": f"the whole run: {total} FP64 steps of the 400x160 channel with force samples ({wall:.1f} s)",
               **{k: v for k, v in ref.items()}}
        for r in out.values():
            r["cd_rel_error_vs_oracle"] = abs(r["cd"] / ref["cd"] - 1.0)
            if r["strouhal"] is not None and ref["strouhal"] is not None:
                r["strouhal_rel_error_vs_oracle"] = abs(r["strouhal"] / ref["strouhal"] - 1.0)
    f32 = out["f32"]
    line = {
        "metric": "D2Q9 MLUPS (cylinder channel, force sampled every 10 steps) with Cd, Cl and Strouhal number",
        "value": round(f32["mlups_sampled"], 2), "unit": "MLUPS", "n_gpus": 1, "steps": total, "warmup": 0,
        "ms_per_step": round(f32["wall_s"] / total * 1e3, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "D2Q9 400x160 channel, cylinder D32 at (80,80), Re 100, u 0.05, zero-gradient outlet, "
                               f"halfway walls, seeded 1e-4 perturbation; Cd and Cl over steps {steps // 2}-{steps}, "
                               f"Strouhal number over steps {steps}-{total}", "parallelism": "single GPU"},
        "storages": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                     for k, v in out.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def sphere_channel():
    """256 x 128 x 128 channel around a voxelised sphere of diameter 32 cells centred at
    (64, 64, 64): inlet u 0.05, Re 200, Smagorinsky C 0.17, convective outlet, slip sides,
    halfway bounce-back."""
    from orrery_amd import Lbm3DConfig, Smagorinsky, Wall, WallKind, XBoundary, geometry, tau_of_viscosity
    from orrery_amd.polar import place_body
    u, d, re = 0.05, 32.0, 200.0
    config = Lbm3DConfig(nx=256, ny=128, nz=128, tau=tau_of_viscosity(u * d / re), x_boundary=XBoundary.inflow_outflow,
                         y_low=Wall(WallKind.slip), y_high=Wall(WallKind.slip), z_low=Wall(WallKind.slip),
                         z_high=Wall(WallKind.slip), inlet_ux=u, smagorinsky=Smagorinsky(True, 0.17))
    mesh = geometry.sphere_mesh(0.5 * d, geometry.RevolutionResolution(64, 0.05)).transformed(
        None, (0.0, 0.0, 0.0), (64.0, 64.0, 64.0))
    return config, place_body(mesh, 256, 128, 128, interpolated=False, device=True), u, d


def sphere_bench(args):
    """The sphere channel: f32 and f64 storage on the device from the same seeded start, a force
    sample every 10 steps, Cd over the last 1000 steps; GLUPS and the bulk kernel's HBM rate; the
    oracle on the host for a bounded prefix of the same run, compared with the f64 lattice."""
    import ctypes
    import numpy as np
    from orrery_amd import Lbm3DSolver, abi, autotune, history
    config, mask, u, d = sphere_channel()
    steps = args.sphere_steps
    q = 0.5 * u * u * np.pi * (0.5 * d) ** 2
    lib = abi.load()
    out = {}
    for storage in ("f32", "f64"):
        s = Lbm3DSolver(config, mask, storage)
        s.initialize_perturbed(2024, 1e-4)
        if args.tune_steps > 0:
            autotune.tune(s, args.tune_steps)
        log(f"sphere: {storage}, {steps} steps with a force sample every 10")
        t, cd = [], []
        t0 = time.perf_counter()
        for k in range(steps // 10):
            s.step(10)
            t.append(float((k + 1) * 10))
            cd.append(s.force().force[0] / q)
        wall = time.perf_counter() - t0
        t, cd = np.array(t), np.array(cd)
        ms = ctypes.c_double()
        for rc in (lib.lbm3d_time_steps(s.handle, args.steps, 0, ctypes.byref(ms), None, None),):
            if rc:
                raise RuntimeError(lib.lbm_last_error().decode())
        step_ms = ms.value / args.steps
        kernels = (ctypes.c_double * 6)()
        if lib.lbm3d_time_kernels(s.handle, 100, kernels):
            raise RuntimeError(lib.lbm_last_error().decode())
        counts = [ctypes.c_size_t() for _ in range(4)]
        lib.lbm3d_counts(s.handle, *[ctypes.byref(c) for c in counts])
        split = stream_counts(lib, s.handle)
        per_launch = split["bulk"] + split["side"] / 2
        bulk_ms = (kernels[0] + kernels[1]) / 2
        sb = 4 if storage == "f32" else 8
        out[storage] = {"cd": float(history.mean_since(t, cd, float(steps - 1000))), "cd_last": float(cd[-1]),
                        "glups": config.cells / (step_ms * 1e-3) / 1e9, "ms_per_step": step_ms,
                        "sampled_run_s": wall, "bulk_ms": bulk_ms, "side_ms_b_step": kernels[3],
                        "boundary_ms": (kernels[4] + kernels[5]) / 2,
                        "bulk_gbps": 2 * 19 * sb * per_launch / (bulk_ms * 1e-3) / 1e9,
                        "bulk_nodes": counts[0].value, "side_nodes": split["side"], "solid_nodes": counts[2].value,
                        "links": counts[3].value}
        s.close()
    out["f32"]["cd_rel_error_vs_f64"] = abs(out["f32"]["cd"] / out["f64"]["cd"] - 1.0)
    cpu = None
    if args.cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle import Oracle3D, load
        threads = os.cpu_count() or 1
        load().orc_set_threads(threads)
        n = args.cpu_steps
        log(f"sphere: the oracle, {n} steps on {threads} threads, against the f64 lattice")
        o = Oracle3D(config, mask)
        o.initialize_perturbed(2024, 1e-4)
        t0 = time.perf_counter()
        o.step(n)
        dt = time.perf_counter() - t0
        g = Lbm3DSolver(config, mask, "f64")
        g.initialize_perturbed(2024, 1e-4)
        g.step(n)
        fo, fg = o.force().force[0], g.force().force[0]
        mo, mg = o.macroscopic(), g.macroscopic()
        fluid = np.asarray(mo.solid) == 0
        dv = max(float(np.abs(getattr(mo, a)[fluid] - getattr(mg, a)[fluid]).max())
                 for a in ("velocity_x", "velocity_y", "velocity_z"))
        g.close()
        cpu = {"value": round(config.cells * n / dt / 1e9, 4), "unit": "GLUPS", "cores": threads, "kind": "port",
               "sample": f"{n} FP64 steps of the same lattice ({dt:.1f} s) from the same seeded start",
               "drag_rel_diff_vs_f64_device": abs(fg / fo - 1.0), "velocity_max_abs_diff_vs_f64_device": dv}
    f32 = out["f32"]
    line = {
        "metric": "D3Q19 GLUPS and Cd (sphere channel 256x128x128, Re 200, LES)", "value": round(f32["glups"], 3),
        "unit": "GLUPS", "n_gpus": 1, "steps": args.steps, "warmup": 0, "ms_per_step": round(f32["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "D3Q19 256x128x128 channel, voxelised sphere D32 at (64,64,64), u 0.05, Re 200, "
                               "Smagorinsky C 0.17, convective outlet, slip sides, halfway walls, seeded 1e-4 "
                               f"perturbation; Cd over steps {steps - 1000}-{steps}", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(f32["bulk_gbps"], 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(f32["bulk_gbps"] * 1e9 / HBM_PEAK, 4), "traffic": None,
                     "kernel": "d3q19_bulk<StoreF32>"},
        "storages": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                     for k, v in out.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def serve_bench(args):
    """The continuous-batching server alone (the `served` section of the default line)."""
    section = served_section(args.requests, 4.0, args.fp64_checks, args.served_sequential_every)
    line = {"metric": "served Cd requests per second (continuous batching, Poisson arrivals)",
            "value": section["requests_per_s"], "unit": "requests/s", "n_gpus": 1, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32/f16s", "data": "synthetic",
            "config": {"workload": section["workload"], "parallelism": "single GPU"}, "served": section}
    print(json.dumps(line), flush=True)
    return 0


# The reference's published numbers for its own GPU benchmarks (AMD Radeon RX 9060 XT, OpenGL
# compute, binary32 shifted populations, A-B ping-pong): docs/performance/
# results-windows-msvc-release.json:370-420 and docs/lbm.md:536-538.
REFERENCE_GPU = {
    "d3q19_128_bgk_steps_s": 681.9, "d3q19_128_les_steps_s": 676.7,
    "d2q9_2048x1024_bgk_steps_s": 1521.0, "d2q9_2048x1024_les_steps_s": 1475.0,
    "adapter_sphere_re100_s": 16.4, "adapter_sphere_re1000_s": 202.0,
}


def reference_sphere_channel(les):
    """sphere_channel(128, 128, 128, 24, 2.5, 0.05, 300) of tests/aero/lbm3d/lbm3d_test_support.h:
    248-270 (a quarter node off the axis, sphere_mask), Smagorinsky {les, 0.1}: the setup of the
    reference's D3Q19 GPU benchmark (tests/bench/aero/bench_lbm_gpu.cpp:89-108)."""
    from orrery_amd import Lbm3DConfig, Smagorinsky, Wall, WallKind, XBoundary, sphere_mask, tau_of_viscosity
    n, d, u, re = 128, 24.0, 0.05, 300.0
    config = Lbm3DConfig(nx=n, ny=n, nz=n, tau=tau_of_viscosity(u * d / re), x_boundary=XBoundary.inflow_outflow,
                         y_low=Wall(WallKind.slip), y_high=Wall(WallKind.slip), z_low=Wall(WallKind.slip),
                         z_high=Wall(WallKind.slip), inlet_ux=u, smagorinsky=Smagorinsky(les, 0.1))
    centre = (2.5 * d, 0.5 * (n - 1) + 0.25, 0.5 * (n - 1) + 0.25)
    return config, sphere_mask(n, n, n, centre, 0.5 * d)


def reference_cylinder_channel(les):
    """cylinder_setup(2048, 1024, 64, 400, 512.5, 0.05, 1000) of tests/aero/lbm/lbm_test_support.h:
    154-167 (convective outlet, slip sides), Smagorinsky {les, 0.1}: the setup of the reference's
    D2Q9 GPU benchmark (bench_lbm_gpu.cpp:111-133)."""
    from orrery_amd import Lbm2DConfig, Smagorinsky, Wall, WallKind, XBoundary, circle_mask, tau_of_viscosity
    d, u, re = 64.0, 0.05, 1000.0
    config = Lbm2DConfig(nx=2048, ny=1024, tau=tau_of_viscosity(u * d / re), x_boundary=XBoundary.inflow_outflow,
                         bottom=Wall(WallKind.slip), top=Wall(WallKind.slip), inlet_ux=u,
                         smagorinsky=Smagorinsky(les, 0.1))
    return config, circle_mask(2048, 1024, 400.0, 512.5, 0.5 * d)


def steps_per_second(solver, count=1000, warmup=50):
    """The reference's method (bench_lbm_gpu.cpp:53-65): warm-up steps, the device drained, then
    steady-clock time around `count` steps and a drain."""
    solver.step(warmup)
    solver.synchronize()
    t0 = time.perf_counter()
    solver.step(count)
    solver.synchronize()
    return count / (time.perf_counter() - t0)


def adapter_sphere_solve(reynolds, cells, upstream, downstream, width, duration, average):
    """One adapter solve of the reference's sphere validation (tests/aero_validation/lbm3d/
    test_sphere.cpp, body_case.h): setup (voxelisation included) and every step, wall clock."""
    import math
    from orrery_amd import geometry
    from orrery_amd.aero import (AeroGeometry, FlowConditions, LbmAeroSolver3D, LbmSolver3DOptions, LesMode,
                                 ReferenceValues)
    geo = AeroGeometry(geometry.sphere_mesh(0.5, geometry.RevolutionResolution(64, 0.05)),
                       ReferenceValues(area=0.25 * math.pi, length=1.0))
    cond = FlowConditions(speed=0.1, density=1000.0)
    cond.viscosity = cond.density * cond.speed / reynolds
    opt = LbmSolver3DOptions(cells_per_length=cells, upstream=upstream, downstream=downstream, width=width,
                             duration=duration, average_window=average, les=LesMode.off, steps_per_call=2000)
    t0 = time.perf_counter()
    s = LbmAeroSolver3D(opt)
    s.setup(geo, cond)
    r = s.run()
    wall = time.perf_counter() - t0
    c = s.info().config
    s.close()
    return {"seconds": wall, "cd": r.cd, "iterations": int(r.iterations), "grid": f"{c.nx}x{c.ny}x{c.nz}",
            "schiller_naumann": 24.0 / reynolds * (1.0 + 0.15 * reynolds ** 0.687)}


def reference_bench(args):
    """The reference's own GPU benchmarks and adapter solves, run here, beside its published
    numbers on its GPU: D3Q19 128^3 sphere channel and D2Q9 2048 x 1024 cylinder (BGK and
    Smagorinsky, steps per second over 1000 steps), and the Re 100 / Re 1000 sphere solves of its
    validation (seconds per solve)."""
    from orrery_amd import Lbm2DSolver, Lbm3DSolver
    out = {}
    for les in (False, True):
        tag = "les" if les else "bgk"
        print(f"bench: reference D3Q19 128^3 sphere channel ({tag})", file=sys.stderr, flush=True)
        config, mask = reference_sphere_channel(les)
        s = Lbm3DSolver(config, mask, "f32")
        rate = steps_per_second(s)
        out[f"d3q19_128_{tag}"] = {"steps_s": rate, "mlups": rate * config.cells / 1e6,
                                   "reference_steps_s": REFERENCE_GPU[f"d3q19_128_{tag}_steps_s"],
                                   "ratio": rate / REFERENCE_GPU[f"d3q19_128_{tag}_steps_s"]}
        s.close()
        print(f"bench: reference D2Q9 2048x1024 cylinder ({tag})", file=sys.stderr, flush=True)
        config2, mask2 = reference_cylinder_channel(les)
        s2 = Lbm2DSolver(config2, mask2, "f32")
        rate2 = steps_per_second(s2)
        out[f"d2q9_2048x1024_{tag}"] = {"steps_s": rate2, "mlups": rate2 * config2.cells / 1e6,
                                        "reference_steps_s": REFERENCE_GPU[f"d2q9_2048x1024_{tag}_steps_s"],
                                        "ratio": rate2 / REFERENCE_GPU[f"d2q9_2048x1024_{tag}_steps_s"]}
        s2.close()
    for re, case in ((100.0, (16.0, 4.0, 8.0, 6.0, 30.0, 10.0)), (1000.0, (48.0, 2.0, 4.0, 4.0, 20.0, 10.0))):
        print(f"bench: reference adapter sphere solve, Re {re:g}", file=sys.stderr, flush=True)
        r = adapter_sphere_solve(re, *case)
        key = f"adapter_sphere_re{int(re)}"
        r["reference_seconds"] = REFERENCE_GPU[f"{key}_s"]
        r["ratio"] = REFERENCE_GPU[f"{key}_s"] / r["seconds"]
        out[key] = r
    head = out["d3q19_128_bgk"]
    line = {
        "metric": "the reference's GPU benchmark: D3Q19 128^3 sphere channel steps/s (D 24, Re 300, BGK)",
        "value": round(head["steps_s"], 1), "unit": "steps/s", "n_gpus": 1, "steps": 1000, "warmup": 50,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": round(head["ratio"], 2), "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "tests/bench/aero/bench_lbm_gpu.cpp cases and the sphere validation solves of "
                               "tests/aero_validation/lbm3d, at the reference's sizes", "parallelism": "single GPU"},
        "baseline_hardware": "AMD Radeon RX 9060 XT (the reference's published runs)",
        "cases": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                  for k, v in out.items()},
    }
    print(json.dumps(line), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="oracle steps per chunk of the CPU baseline (0: no CPU baseline)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="the CPU baseline repeats its chunk until this many seconds have passed")
    ap.add_argument("--storages", default="f32,f16s,bf16s")
    ap.add_argument("--tune-steps", type=int, default=30,
                    help="steps per candidate of the launch autotuner before timing (0: the default launch)")
    ap.add_argument("--groups-per-cu", default="0,0", metavar="A,B",
                    help="with --tune-steps 0: occupancy caps of the bulk kernel, A step and B step (0: none)")
    ap.add_argument("--compare-steps", type=int, default=2000,
                    help="step at which every storage's force and velocity are compared with the f64 storage's "
                         "(0: no comparison, no f64 run)")
    ap.add_argument("--workload", choices=["channel", "polar", "serve", "d2q9", "sphere", "reference"],
                    default="channel",
                    help="channel: the 512x256x256 lattice (the headline line); polar: the batched wing-section "
                         "sweep, batched against one at a time; serve: the continuous-batching server; d2q9: the "
                         "400x160 cylinder channel; sphere: the 256x128x128 sphere channel; reference: the "
                         "reference's own GPU benchmarks and adapter solves beside its published numbers")
    ap.add_argument("--d2q9-steps", type=int, default=20000, help="d2q9: configured steps (the run is twice as long)")
    ap.add_argument("--sphere-steps", type=int, default=5000, help="sphere: steps of the sampled runs")
    ap.add_argument("--lattices", type=int, default=48,
                    help="wing sections of the batched sweep (the `batched` section; 0 leaves it out)")
    ap.add_argument("--sequential-every", type=int, default=8,
                    help="time the one-at-a-time baseline of the sweep on every k-th section")
    ap.add_argument("--requests", type=int, default=16,
                    help="requests of the served workload (the `served` section; 0 leaves it out)")
    ap.add_argument("--served-sequential-every", type=int, default=4,
                    help="time every k-th served request alone for the one-at-a-time baseline")
    ap.add_argument("--fp64-checks", type=int, default=8,
                    help="converged solves (the smallest) whose Cd is checked against the FP64 device lattice")
    ap.add_argument("--csv", default="", help="polar: write the comparison table (CSV) of the batched solves here")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device with a plain copy kernel (3 launches), the "
                         "known-size reference for the HBM byte counters of a profiler run")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU; --gpus {args.gpus} is not supported "
              f"(independent lattices shard across GPUs by HBM budget, one lattice stays on one GPU)",
              file=sys.stderr)
        return 2

    if args.workload == "polar":
        return polar_bench(args)
    if args.workload == "serve":
        return serve_bench(args)
    if args.workload == "d2q9":
        return d2q9_bench(args)
    if args.workload == "sphere":
        return sphere_bench(args)
    if args.workload == "reference":
        return reference_bench(args)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
    config, mask = workload()
    per_kernel_steps = min(200, max(20, args.steps // 5))
    compare = max(args.compare_steps, args.warmup + args.steps + per_kernel_steps) if args.compare_steps else 0
    storages = [k for k in args.storages.split(",") if k]
    if compare and "f64" not in storages:
        storages.insert(0, "f64")
    storages.sort(key=lambda k: k != "f64")
    results = {}
    ref = None
    for storage in storages:
        print(f"bench: timing {storage} ({args.warmup} warm-up, {args.steps} timed steps"
              + (f", compared with f64 at step {compare})" if compare else ")"), file=sys.stderr, flush=True)
        results[storage] = time_storage(config, mask, storage, args.steps, args.warmup, per_kernel_steps, compare,
                                        ref, args.tune_steps,
                                        groups=tuple(int(g) for g in args.groups_per_cu.split(",")))
        if storage == "f64" and compare:
            ref = {"force_x": results[storage]["force_x"], "fields": results[storage].pop("fields")}
    ref = None
    main_key = "f32" if "f32" in results else next(k for k in results if k != "f64")
    main_r = results[main_key]
    cells = config.cells
    glups = cells / (main_r["ms_per_step"] * 1e-3) / 1e9
    bytes_per_node = 2 * 19 * {"f32": 4, "f16s": 2, "bf16s": 2, "f64": 8}[main_key]
    fused = main_r["launch"]["schedule"] == "fused"
    if fused:
        # one kernel per step: its time is the step time, over every node
        achieved = bytes_per_node * cells / (main_r["ms_per_step"] * 1e-3)
    else:
        achieved = bytes_per_node * main_r["bulk_kernel_nodes"] / (main_r["bulk_ms"] * 1e-3)
    cpu = None
    if args.cpu_steps > 0:
        print("bench: timing the CPU baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(config, mask, args.cpu_steps, args.cpu_seconds)
    traffic = measured_traffic() if main_key == "f32" else None  # the counter runs cover the f32 kernel
    line = {
        "metric": "D3Q19 GLUPS (FP32 A-A, BGK+Smagorinsky) and % HBM roofline; batched Cd solves/hour, p50/p99 "
                  "latency @1% Cd error",
        "value": round(glups, 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": main_key,
        "data": "synthetic",
        "config": {
            "workload": "D3Q19 512x256x256 channel, voxelised Apollo-type capsule D64 pitched 20 deg (Re 1000, "
                        "u 0.05, Smagorinsky C 0.17), velocity inlet, convective outlet, slip sides, halfway "
                        f"bounce-back, {main_key} A-A in place",
            "cells": cells,
            "bulk_nodes": main_r["bulk_nodes"],
            "boundary_nodes": main_r["boundary_nodes"],
            "side_nodes": main_r["side_nodes"],
            "listed_boundary_nodes": main_r["listed_nodes"],
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
            "traffic_over_algorithmic": round(traffic["bytes_per_launch"] / (bytes_per_node * main_r["bulk_kernel_nodes"]), 4)
            if traffic else None,
            "kernel": ("d3q19_fused<%s>" if fused else "d3q19_bulk<%s>") % {"f32": "StoreF32", "f16s": "StoreF16",
                                                                            "bf16s": "StoreBF16", "f64": "StoreF64"}[main_key],
            "launch": main_r["launch"],
            "bulk_ms": round(main_r["bulk_ms"], 4),
            "bulk_ms_a_b": main_r["bulk_ms_a_b"],
            "nodes_per_launch": main_r["bulk_kernel_nodes"],
            "side_ms_b_step": round(main_r["side_ms_b"], 4),
            "boundary_ms": round(main_r["boundary_ms"], 4),
            "step_roofline_frac": round(bytes_per_node * cells / (main_r["ms_per_step"] * 1e-3) / HBM_PEAK, 4),
        },
        "cpu_baseline": cpu,
        "batched": None,
        "served": None,
        "storage_variants": {
            k: {"glups": round(cells / (r["ms_per_step"] * 1e-3) / 1e9, 3),
                "bytes_per_node": 2 * 19 * {"f32": 4, "f16s": 2, "bf16s": 2, "f64": 8}[k],
                "hbm_frac_bulk": round(2 * 19 * {"f32": 4, "f16s": 2, "bf16s": 2, "f64": 8}[k] * r["bulk_kernel_nodes"]
                                       / (r["bulk_ms"] * 1e-3) / HBM_PEAK, 4),
                "ms_per_step": round(r["ms_per_step"], 4), "launch": r["launch"],
                **{e: float(f"{r[e]:.3e}") for e in ("force_x_rel_error_vs_f64", "velocity_l2_vs_f64") if e in r},
                **({"compared_at_step": r["compared_at_step"]} if "compared_at_step" in r else {})}
            for k, r in results.items()
        },
    }
    # The metric's second half: batched and served converged Cd solves, measured in this run.
    if args.lattices > 0:
        t0 = time.perf_counter()
        log(f"batched: the {args.lattices}-section wing sweep")
        line["batched"] = batched_section(args.lattices, args.sequential_every, args.fp64_checks)
        line["batched"]["section_s"] = round(time.perf_counter() - t0, 1)
    if args.requests > 0:
        t0 = time.perf_counter()
        log(f"served: {args.requests} requests at 4/s")
        line["served"] = served_section(args.requests, 4.0, args.fp64_checks, args.served_sequential_every)
        line["served"]["section_s"] = round(time.perf_counter() - t0, 1)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
