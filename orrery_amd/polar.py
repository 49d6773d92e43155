"""An angle-of-attack sweep of NACA 0012 wing sections, solved as one batch of resident lattices
or one lattice at a time: the batched Cd workload.

Each case is an infinite-span section (the span runs through periodic y sides) of chord c
cells in a channel with a velocity inlet, a convective outlet, slip z sides and an absorbing
layer before the outlet, at its own Reynolds number and angle of attack. A case has converged,
after at least two flow-throughs, when the mean drag coefficient of the last 50 samples (500
steps) differs from that of the 50 before by less than 1e-4 (a steady wake), or when the forces
of the last half of the run are settled periodic shedding by the reference's adapter test
(forces_settled_3d, src/aero/lbm3d/lbm_aero_solver3d.cpp:60-83), or when the mean coefficients of
the two quarters that make up the last half of the run agree within 0.5% (an irregular wake
whose time average has settled; history.judge_solve). The lift slope is the least-squares
slope of Cl against the angle of attack over the attached range, as in the reference's
comparison of solvers (src/app/aero/aero_comparison.cpp:101-164).
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from . import autotune
from . import geometry as geo
from . import history
from .batch import LatticeBatch
from .config import Lbm3DConfig, Outlet, Smagorinsky, SolidWall, Sponge, Wall, WallKind, XBoundary
from .masks import SolidMask3D, fluid_mask_3d
from .solver import Lbm3DSolver

U_LATTICE = 0.05
SAMPLE_EVERY = 10
WINDOW = 50          # samples in a mean (500 steps)
TOLERANCE = 1e-4     # |Cd(t) - Cd(t - 500)|
MAX_STEPS = 20000


@dataclasses.dataclass
class WingCase:
    index: int
    nx: int
    ny: int
    nz: int
    chord: int
    aoa_deg: float
    reynolds: float

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def min_steps(self) -> int:
        """Two flow-throughs of the channel, in whole sampling windows."""
        steps = int(math.ceil(2.0 * self.nx / U_LATTICE))
        return -(-steps // (WINDOW * SAMPLE_EVERY)) * WINDOW * SAMPLE_EVERY


def wing_cases(n: int = 48, seed: int = 2024, aoa_range=(-4.0, 18.0)) -> list[WingCase]:
    """n cases with seeded dims uniform over [96^3, 320 x 160 x 160], chords uniform over
    [48, 96] cells (at most nx / 3), Reynolds numbers uniform over [500, 2000], and angles of
    attack evenly spaced over `aoa_range` degrees."""
    rng = np.random.default_rng(seed)
    angles = np.linspace(aoa_range[0], aoa_range[1], n) if n > 1 else np.array([aoa_range[0]])
    cases = []
    for k in range(n):
        nx = int(rng.integers(96, 321))
        ny = int(rng.integers(96, 161))
        nz = int(rng.integers(96, 161))
        chord = int(min(rng.integers(48, 97), nx // 3))
        cases.append(WingCase(k, nx, ny, nz, chord, float(angles[k]), float(rng.uniform(500.0, 2000.0))))
    return cases


def place_body(mesh: geo.TriangleMesh, nx: int, ny: int, nz: int, interpolated: bool = True,
               margin: int = 4, device: bool = False) -> SolidMask3D:
    """Voxelises `mesh` (lattice units: node (i, j, k) sits at (i, j, k)) over its bounding box
    plus `margin` nodes, clipped to the lattice, and places it in a fluid mask; with
    `interpolated`, the signed distance within three cells becomes the mask's distance. With
    `device`, the voxels and distances are computed on the GPU (the same values bit for bit)."""
    lo, hi = mesh.bounds()
    dims = (nx, ny, nz)
    first = [max(0, int(math.floor(lo[a])) - margin) for a in range(3)]
    last = [min(dims[a] - 1, int(math.ceil(hi[a])) + margin) for a in range(3)]
    count = [last[a] - first[a] + 1 for a in range(3)]
    spec = geo.GridSpec(tuple(first[a] - 0.5 for a in range(3)), 1.0, *count)
    vox = geo.voxelize(mesh, spec, device=device)
    mask = fluid_mask_3d(nx, ny, nz)
    geo.place_voxels(mask, vox, first)
    if interpolated:
        geo.place_distance(mask, geo.signed_distance_band(mesh, vox, 3.0, device=device), first)
    return mask


def wing_lattice(case: WingCase, interpolated: bool = True, device: bool = False) -> tuple[Lbm3DConfig, SolidMask3D]:
    """The case's lattice; with `device` the wing is voxelised on the GPU (the same mask and
    distances bit for bit)."""
    c = case.chord
    # A section longer than the periodic span, so no cell sees the end caps within the
    # three-cell distance band.
    wing = geo.naca_wing(0.0, 0.0, 0.12, float(c), float(case.ny + 16), max(16, 2 * c), 1)
    pivot = (0.25 * c, 0.0, 0.0)
    to = (round(0.3 * case.nx) + 0.25, 0.5 * (case.ny - 1), 0.5 * (case.nz - 1) + 0.25)
    mesh = wing.transformed(geo.rotation_y(math.radians(case.aoa_deg)), pivot, to)
    mask = place_body(mesh, case.nx, case.ny, case.nz, interpolated, device=device)
    tau = 0.5 + 3.0 * U_LATTICE * c / case.reynolds
    config = Lbm3DConfig(nx=case.nx, ny=case.ny, nz=case.nz, tau=tau, x_boundary=XBoundary.inflow_outflow,
                         outlet=Outlet.convective, y_low=Wall(WallKind.periodic), y_high=Wall(WallKind.periodic),
                         z_low=Wall(WallKind.slip), z_high=Wall(WallKind.slip), inlet_ux=U_LATTICE,
                         smagorinsky=Smagorinsky(True, 0.1), sponge=Sponge(0.1, 0, case.nx // 10, 0),
                         solid_wall=SolidWall.interpolated if interpolated else SolidWall.halfway)
    return config, mask


def coefficients(case: WingCase, force: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(Cd, Cl) of lattice forces [n, >=3]: F / (rho u^2 c span / 2), span = ny."""
    q = 0.5 * U_LATTICE * U_LATTICE * case.chord * case.ny
    return force[:, 0] / q, force[:, 2] / q


def judge(case: WingCase, time_steps: np.ndarray, force: np.ndarray) -> history.Verdict:
    """The convergence verdict of a case's force history (lattice forces [n, >=3]); the side
    force is the spanwise (y) component."""
    q = 0.5 * U_LATTICE * U_LATTICE * case.chord * case.ny
    return history.judge_solve(time_steps, force[:, 0] / q, force[:, 2] / q, force[:, 1] / q, case.min_steps,
                               WINDOW, TOLERANCE)


def converged(case: WingCase, time_steps: np.ndarray, force: np.ndarray) -> bool:
    return judge(case, time_steps, force).converged


@dataclasses.dataclass
class Solve:
    case: WingCase
    steps: int
    cd: float
    cl: float
    converged: bool
    finished_at: float  # seconds from the start of the run
    kind: str = ""      # "steady", "settled" or "stationary" when converged
    mean_from: float = 0.0  # first step of the samples the coefficients average


def run_batched(cases: list[WingCase], lattices, storage: str = "f32", budget_bytes: int = 240 * 10 ** 9,
                chunk: int = WINDOW * SAMPLE_EVERY, seed: int = 7, log=None,
                tune_steps: int = 10, timing: dict | None = None) -> tuple[list[Solve], float]:
    """Every case resident in one batch (admission and, with `tune_steps`, the launch autotuner
    included in the time), advanced `chunk` steps per launch sequence; converged cases leave
    the batch. With `timing`, the first chunk runs with per-kernel HIP events and the dict
    receives their times (LatticeBatch.run_timed). `lattices` holds the (config, mask) of each case. Returns the solves and the wall
    time."""
    t0 = time.perf_counter()
    batch = LatticeBatch(budget_bytes=budget_bytes, sample_every=SAMPLE_EVERY)
    active = {}
    for case, (config, mask) in zip(cases, lattices):
        active[case.index] = (case, batch.add(config, mask, storage, seed, 1e-4))
    if tune_steps > 0:
        chosen = autotune.autotune_batch(batch, tune_steps)
        if log:
            log(f"batched: launch shapes {chosen}")
    done = []
    first = True
    while active:
        if first and timing is not None:
            # the first chunk with each kernel between HIP events (the same results as run())
            timing.update(batch.run_timed(chunk))
        else:
            batch.run(chunk)
        first = False
        for key in list(active):
            case, lat = active[key]
            t, f = lat.history()
            v = judge(case, t, f)
            if v.converged or t[-1] >= MAX_STEPS:
                cd, cl = (v.drag, v.lift) if v.converged else [float(c[-WINDOW:].mean()) for c in coefficients(case, f)]
                done.append(Solve(case, int(t[-1]), cd, cl, v.converged, time.perf_counter() - t0, v.kind,
                                  v.start if v.converged else float(t[max(0, len(t) - WINDOW)])))
                batch.remove(lat)
                del active[key]
                if log:
                    log(f"batched: case {case.index} done at step {int(t[-1])} ({len(active)} left)")
    wall = time.perf_counter() - t0
    batch.close()
    return done, wall


def run_sequential(cases: list[WingCase], lattices, storage: str = "f32", seed: int = 7,
                   log=None, field_every: int = 0) -> tuple[list[Solve], float]:
    """One plain lattice at a time (its creation included in the time), its force read back
    every sample (host-paced). With `field_every`, the whole macroscopic field is also read
    back every `field_every` steps, as the reference's adapter does on every step() call
    (LbmAeroSolver3D::step, src/aero/lbm3d/lbm_aero_solver3d.cpp:387-490; 200 steps per call by
    default)."""
    done = []
    t0 = time.perf_counter()
    for case, (config, mask) in zip(cases, lattices):
        s = Lbm3DSolver(config, mask, storage)
        s.initialize_perturbed(seed, 1e-4)
        times = []
        forces = []
        while True:
            s.step(SAMPLE_EVERY)
            times.append(float(s.steps()))
            forces.append(s.force().force)
            if field_every > 0 and s.steps() % field_every == 0:
                s.macroscopic()
            if len(times) % WINDOW == 0:
                t = np.array(times)
                f = np.array(forces)
                v = judge(case, t, f)
                if v.converged or t[-1] >= MAX_STEPS:
                    cd, cl = (v.drag, v.lift) if v.converged else [float(c[-WINDOW:].mean())
                                                                   for c in coefficients(case, f)]
                    done.append(Solve(case, int(t[-1]), cd, cl, v.converged, time.perf_counter() - t0, v.kind,
                                      v.start if v.converged else float(t[max(0, len(t) - WINDOW)])))
                    if log:
                        log(f"sequential: case {case.index} done at step {int(t[-1])}")
                    break
        s.close()
    return done, time.perf_counter() - t0


def fp64_coefficients(case: WingCase, lattice, steps: int, mean_from: float, seed: int = 7) -> tuple[float, float]:
    """The case's lattice alone in FP64 storage (the device FP64 oracle) from the same seeded
    start to `steps`, its (Cd, Cl) averaged over the samples at step >= `mean_from`."""
    from . import history as hist
    config, mask = lattice
    s = Lbm3DSolver(config, mask, "f64")
    s.initialize_perturbed(seed, 1e-4)
    times, forces = [], []
    while s.steps() < steps:
        s.step(SAMPLE_EVERY)
        times.append(float(s.steps()))
        forces.append(s.force().force)
    s.close()
    cd, cl = coefficients(case, np.array(forces))
    t = np.array(times)
    return hist.mean_since(t, cd, mean_from), hist.mean_since(t, cl, mean_from)


def records(solves: list[Solve], storage: str = "f32") -> list:
    """The solves as aero experiment records for the comparison tables: a NACA 0012 section of
    unit chord (the span is periodic), the angle of attack, and the lattice result (Cd, Cl, the
    steps run, convergence). All cases share the geometry, so they form one table."""
    from . import experiment as ex
    out = []
    for s in solves:
        e = ex.AeroExperiment(name=f"case {s.case.index:02d} ({storage})")
        e.geometry = ex.GeometryRef(shape=ex.AeroShape.naca_section, naca="0012",
                                    dimensions=ex.ShapeDimensions(chord_m=1.0))
        e.conditions.angle_of_attack_deg = s.case.aoa_deg
        e.solver = ex.AeroSolverSettings(solver=ex.SolverKind.lbm, reference_area=ex.ReferenceAreaKind.planform,
                                         reference_length_m=1.0, lbm_cells_per_length=float(s.case.chord))
        e.result = ex.AeroResultRecord(solver=ex.SolverKind.lbm, cd=s.cd, cl=s.cl, reynolds=s.case.reynolds,
                                       reference_area_m2=1.0, reference_length_m=1.0, iterations=s.steps,
                                       converged=s.converged)
        out.append(e)
    return out


def comparison_csv(solves: list[Solve], storage: str = "f32") -> str:
    """The sweep's comparison table as CSV (comparison_csv, src/app/aero/aero_comparison.cpp:166-181)."""
    from . import experiment as ex
    tables = ex.compare_by_geometry(records(solves, storage))
    return "".join(ex.comparison_csv(t) for t in tables)


def lift_slope(solves: list[Solve], attached=(-4.0, 8.0)) -> float:
    """Least-squares dCl/dalpha per radian over the cases with alpha in `attached` degrees."""
    pts = [(math.radians(s.case.aoa_deg), s.cl) for s in solves if attached[0] <= s.case.aoa_deg <= attached[1]]
    if len(pts) < 2:
        return float("nan")
    a = np.array(pts)
    x = a[:, 0] - a[:, 0].mean()
    return float((x * (a[:, 1] - a[:, 1].mean())).sum() / (x * x).sum())


__all__ = ["WingCase", "wing_cases", "wing_lattice", "place_body", "run_batched", "run_sequential", "lift_slope",
           "coefficients", "converged", "judge", "Solve", "records", "comparison_csv", "fp64_coefficients"]
