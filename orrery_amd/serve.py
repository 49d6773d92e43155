"""A continuous-batching server for lattice Boltzmann drag requests: requests arrive over time,
their lattices are built on the host while the device runs, admitted into one batch of resident
lattices as soon as the HBM budget has room (not first-in-first-out alone: a request that fits
goes in while a larger one waits), advanced together, and retired as each converges.

A request is a body (sphere, cone, cylinder, capsule or wedge: the reference's primitives,
src/geom/primitives.cpp) at an angle of attack and Reynolds number in a channel of its own
size, in binary32 or 16-bit storage. Converged, after one flow-through: the mean drag coefficient
of the last 50 samples (500 steps) within 1e-4 of the 50 before (a steady wake), or the forces of
the last half of the run settled periodic shedding by the reference's adapter test
(forces_settled_3d, src/aero/lbm3d/lbm_aero_solver3d.cpp:60-83), or the mean coefficients of the two
quarters that make up the last half of the run within 0.5% of each other (an irregular wake whose
time average has settled; history.judge_solve).
"""
from __future__ import annotations

import concurrent.futures
import dataclasses
import math
import time

import numpy as np

from . import autotune
from . import geometry as geo
from . import history
from .batch import LatticeBatch
from .config import Lbm3DConfig, Outlet, Smagorinsky, SolidWall, Sponge, Wall, WallKind, XBoundary
from .polar import place_body
from .solver import LbmError, Lbm3DSolver

U_LATTICE = 0.05
SAMPLE_EVERY = 10
WINDOW = 50
TOLERANCE = 1e-4
MAX_STEPS = 20000
BODIES = ("sphere", "cone", "cylinder", "capsule", "wedge")


@dataclasses.dataclass
class Request:
    id: int
    arrival: float  # seconds after the start
    body: str
    nx: int
    ny: int
    nz: int
    aoa_deg: float
    reynolds: float
    storage: str

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def length(self) -> int:
        """Body length in cells."""
        return max(8, min(self.nx // 5, min(self.ny, self.nz) // 3))

    @property
    def min_steps(self) -> int:
        steps = int(math.ceil(self.nx / U_LATTICE))
        return -(-steps // (WINDOW * SAMPLE_EVERY)) * WINDOW * SAMPLE_EVERY


def requests(n: int, rate: float = 4.0, seed: int = 11, dims=((64, 64, 64), (384, 192, 192))) -> list[Request]:
    """n requests with Poisson arrivals of `rate` per second (exponential gaps), bodies, dims,
    angles (0-30 degrees), Reynolds numbers (200-3000) and storage (f32 or f16s) drawn from a
    seeded generator."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    lo, hi = dims
    for k in range(n):
        t += float(rng.exponential(1.0 / rate))
        out.append(Request(k, t, BODIES[int(rng.integers(0, len(BODIES)))], int(rng.integers(lo[0], hi[0] + 1)),
                           int(rng.integers(lo[1], hi[1] + 1)), int(rng.integers(lo[2], hi[2] + 1)),
                           float(rng.uniform(0.0, 30.0)), float(rng.uniform(200.0, 3000.0)),
                           "f32" if rng.random() < 0.5 else "f16s"))
    return out


def body_mesh(body: str, length: float) -> geo.TriangleMesh:
    """The body, `length` cells long along +x, centred on the origin."""
    L = float(length)
    if body == "sphere":
        return geo.sphere_mesh(0.5 * L, geo.RevolutionResolution(32, 0.1))
    if body == "cone":
        # apex forward: the axis (z, base at 0) turned onto -x
        m = geo.cone_mesh(0.35 * L, L, geo.RevolutionResolution(32, 0.1))
        return m.transformed(geo.rotation_y(-0.5 * math.pi), (0.0, 0.0, 0.5 * L))
    if body == "cylinder":
        m = geo.cylinder_mesh(0.25 * L, L, geo.RevolutionResolution(32, 0.1))
        return m.transformed(geo.rotation_y(0.5 * math.pi), (0.0, 0.0, 0.5 * L))
    if body == "capsule":
        shape = geo.CapsuleShape()
        m = geo.capsule_mesh(shape, geo.RevolutionResolution(32, 0.05))
        lo, hi = m.bounds()
        s = L / float(max(hi - lo))
        # heat shield (the bottom, z = 0) facing the flow: the axis turned onto +x, scaled
        r = geo.rotation_y(0.5 * math.pi) * s
        return m.transformed(r, (0.0, 0.0, 0.5 * (lo[2] + hi[2])))
    if body == "wedge":
        m = geo.wedge_mesh(L, math.radians(12.0), L)
        return m.transformed(None, (0.5 * L, 0.0, 0.0))
    raise ValueError(f"unknown body {body!r}")


def request_lattice(req: Request) -> tuple[Lbm3DConfig, "geo.SolidMask3D"]:
    L = req.length
    mesh = body_mesh(req.body, L)
    to = (round(0.3 * req.nx) + 0.25, 0.5 * (req.ny - 1) + 0.25, 0.5 * (req.nz - 1) + 0.25)
    mesh = mesh.transformed(geo.rotation_y(math.radians(req.aoa_deg)), (0.0, 0.0, 0.0), to)
    mask = place_body(mesh, req.nx, req.ny, req.nz, interpolated=True)
    tau = 0.5 + 3.0 * U_LATTICE * L / req.reynolds
    config = Lbm3DConfig(nx=req.nx, ny=req.ny, nz=req.nz, tau=tau, x_boundary=XBoundary.inflow_outflow,
                         outlet=Outlet.convective, y_low=Wall(WallKind.slip), y_high=Wall(WallKind.slip),
                         z_low=Wall(WallKind.slip), z_high=Wall(WallKind.slip), inlet_ux=U_LATTICE,
                         smagorinsky=Smagorinsky(True, 0.1), sponge=Sponge(0.1, 0, req.nx // 10, 0),
                         solid_wall=SolidWall.interpolated)
    return config, mask


def drag_coefficient(req: Request, force: np.ndarray) -> np.ndarray:
    """Cd = F_x / (rho u^2 L^2 / 2), reference area L^2."""
    return force[:, 0] / (0.5 * U_LATTICE * U_LATTICE * req.length * req.length)


def judge(req: Request, time_steps: np.ndarray, force: np.ndarray) -> history.Verdict:
    """Convergence verdict of a request's force history: drag x, lift z, side force y, all over
    the reference area L^2."""
    q = 0.5 * U_LATTICE * U_LATTICE * req.length * req.length
    return history.judge_solve(time_steps, force[:, 0] / q, force[:, 2] / q, force[:, 1] / q, req.min_steps,
                               WINDOW, TOLERANCE)


@dataclasses.dataclass
class Served:
    request: Request
    admitted: float
    finished: float
    steps: int
    cd: float
    converged: bool
    kind: str = ""          # "steady" or "settled" when converged
    mean_from: float = 0.0  # first step of the samples the Cd mean covers

    @property
    def latency(self) -> float:
        return self.finished - self.request.arrival

    @property
    def queued(self) -> float:
        return self.admitted - self.request.arrival


class Server:
    """Runs a list of requests against the wall clock: lattices are prepared on `workers` host
    threads as requests arrive; the device batch advances `chunk` steps at a time.

    Admission: a built request enters the batch when the HBM budget has room for its lattice and
    the lattices in the batch hold fewer than `max_active_cells` nodes (one step of that many
    nodes already saturates HBM, so more resident lattices only share the same bandwidth and
    finish later); with `order="shortest"` the waiting request with the least work (nodes times
    its minimum steps) goes first, except that one waiting longer than `max_wait` seconds goes
    before any other. `max_active_cells=None` and `order="arrival"` admit everything that fits,
    oldest first."""

    def __init__(self, budget_bytes: int = 240 * 10 ** 9, chunk: int = WINDOW * SAMPLE_EVERY, workers: int = 4,
                 seed: int = 5, tune_steps: int = 10, max_active_cells: int | None = 48_000_000,
                 order: str = "shortest", max_wait: float = 30.0, max_steps: int = MAX_STEPS):
        self.budget = budget_bytes
        self.chunk = chunk
        self.workers = workers
        self.seed = seed
        self.tune_steps = tune_steps  # launch autotuner, once per storage format (0: default shapes)
        self.max_active_cells = max_active_cells
        self.order = order
        self.max_wait = max_wait
        self.max_steps = max_steps  # a request leaves unconverged at this many steps
        self.busy = 0.0  # seconds inside batch.run

    def run(self, reqs: list[Request], log=None) -> tuple[list[Served], float]:
        batch = LatticeBatch(budget_bytes=self.budget, sample_every=SAMPLE_EVERY)
        pool = concurrent.futures.ThreadPoolExecutor(self.workers)
        pending = sorted(reqs, key=lambda r: r.arrival)
        building = {}   # id -> future of (config, mask)
        waiting = []    # built, not admitted, in arrival order
        active = {}     # id -> (request, lattice, admitted)
        done = []
        tuned = set()
        t0 = time.perf_counter()
        try:
            while pending or building or waiting or active:
                now = time.perf_counter() - t0
                while pending and pending[0].arrival <= now:
                    r = pending.pop(0)
                    building[r.id] = (r, pool.submit(request_lattice, r))
                for rid in [k for k, (_, fut) in building.items() if fut.done()]:
                    r, fut = building.pop(rid)
                    waiting.append((r, fut.result()))
                now = time.perf_counter() - t0
                if self.order == "shortest":
                    waiting.sort(key=lambda item: (now - item[0].arrival <= self.max_wait,
                                                   item[0].cells * item[0].min_steps, item[0].arrival))
                else:
                    waiting.sort(key=lambda item: item[0].arrival)
                active_cells = sum(r.cells for r, _, _ in active.values())
                for item in list(waiting):
                    r, (config, mask) = item
                    if self.max_active_cells is not None and active and \
                            active_cells + r.cells > self.max_active_cells:
                        break  # keeps the order: nothing passes the request at the head
                    try:
                        lat = batch.add(config, mask, r.storage, self.seed, 1e-4)
                    except LbmError as e:
                        if "budget" in str(e):
                            continue
                        raise
                    active[r.id] = (r, lat, time.perf_counter() - t0)
                    active_cells += r.cells
                    waiting.remove(item)
                if not active:
                    if pending and not building and not waiting:
                        time.sleep(max(0.0, min(0.05, pending[0].arrival - (time.perf_counter() - t0))))
                    else:
                        time.sleep(0.002)
                    continue
                storages = {r.storage for r, _, _ in active.values()}
                if self.tune_steps > 0 and not storages <= tuned:
                    chosen = autotune.autotune_batch(batch, self.tune_steps)
                    tuned |= set(chosen)
                    if log:
                        log(f"serve: launch shapes {chosen}")
                t1 = time.perf_counter()
                batch.run(self.chunk)
                self.busy += time.perf_counter() - t1
                for rid in list(active):
                    r, lat, admitted = active[rid]
                    t, f = lat.history()
                    v = judge(r, t, f)
                    cd = drag_coefficient(r, f)
                    if v.converged or t[-1] >= self.max_steps or not np.isfinite(cd[-1]):
                        done.append(Served(r, admitted, time.perf_counter() - t0, int(t[-1]),
                                           v.drag if v.converged else float(cd[-WINDOW:].mean()), v.converged,
                                           v.kind, v.start))
                        batch.remove(lat)
                        del active[rid]
                        if log:
                            log(f"serve: request {r.id} ({r.body} {r.nx}x{r.ny}x{r.nz} {r.storage}) done at step "
                                f"{int(t[-1])}, latency {done[-1].latency:.2f} s, {len(active)} resident")
        finally:
            pool.shutdown(wait=True)
            batch.close()
        return done, time.perf_counter() - t0


def fp64_reference_cd(req: Request, steps: int, seed: int = 5, mean_from: float | None = None) -> float:
    """The request's lattice alone in FP64 storage (the device FP64 oracle) to `steps`; Cd
    averaged over the samples at step >= `mean_from` (default: the last window)."""
    config, mask = request_lattice(req)
    s = Lbm3DSolver(config, mask, "f64")
    s.initialize_perturbed(seed, 1e-4)
    times = []
    forces = []
    while s.steps() < steps:
        s.step(SAMPLE_EVERY)
        times.append(float(s.steps()))
        forces.append(s.force().force)
    s.close()
    cd = drag_coefficient(req, np.array(forces))
    if mean_from is None:
        return float(cd[-WINDOW:].mean())
    return history.mean_since(np.array(times), cd, mean_from)


def serve_bench(n: int, rate: float = 4.0, log=None, check: int = 3, order: str = "shortest",
                max_active_cells: int | None = 48_000_000) -> dict:
    reqs = requests(n, rate)
    server = Server(order=order, max_active_cells=max_active_cells)
    served, wall = server.run(reqs, log=log)
    lat = np.array([s.latency for s in served])
    queued = np.array([s.queued for s in served])
    first = min(r.arrival for r in reqs)
    last = max(s.finished for s in served)
    # Cd error against the device FP64 lattice run to the same step: the smallest converged
    # requests (a steady wake; an unsteady one diverges from any rerun in its shedding phase),
    # else the smallest.
    errors = []
    pool = [s for s in served if s.converged] or served
    for s in sorted(pool, key=lambda s: s.request.cells)[:check]:
        ref = fp64_reference_cd(s.request, s.steps, mean_from=s.mean_from if s.converged else None)
        errors.append(abs(s.cd / ref - 1.0) if ref != 0 else abs(s.cd))
        if log:
            log(f"serve: request {s.request.id} {s.request.storage} Cd {s.cd:.6f} vs FP64 {ref:.6f}")
    return {
        "metric": "served drag requests per second (continuous batching, Poisson arrivals)",
        "value": round(len(served) / (last - first), 4), "unit": "requests/s", "n_gpus": 1, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32/f16s", "data": "synthetic",
        "config": {"workload": f"{n} requests, lambda {rate}/s, bodies {'/'.join(BODIES)}, dims in [64^3, 384x192x192], "
                               "AoA 0-30 deg, Re 200-3000, HBM budget 240 GB", "parallelism": "single GPU"},
        "latency_p50_s": round(float(np.percentile(lat, 50)), 3), "latency_p99_s": round(float(np.percentile(lat, 99)), 3),
        "queued_p50_s": round(float(np.percentile(queued, 50)), 3),
        "queued_p99_s": round(float(np.percentile(queued, 99)), 3),
        "offered_per_s": rate, "gpu_busy_fraction": round(server.busy / wall, 3), "wall_s": round(wall, 2),
        "admission": {"order": order, "max_active_cells": max_active_cells},
        "converged": sum(s.converged for s in served), "cd_rel_error_vs_fp64_max": max(errors) if errors else None,
        "cd_checked": [{"request": s.request.id, "storage": s.request.storage, "cells": s.request.cells,
                        "converged": s.converged} for s in sorted(pool, key=lambda s: s.request.cells)[:check]],
    }


__all__ = ["Request", "requests", "body_mesh", "request_lattice", "Server", "Served", "serve_bench",
           "fp64_reference_cd", "drag_coefficient"]
