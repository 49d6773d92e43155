"""Running aero experiment records on the lattice Boltzmann adapter: the record in, the record
with its result out, and a solver thread driven by a command queue.

- prepare_geometry restates build_primitive (src/app/aero/aero_geometry.cpp:84-188): the body of a
  primitive record, faced into the free stream and centred on its bounds, with its reference
  quantities (src/geom/reference_quantities.cpp: wetted area, silhouette areas on a 512-cell
  raster, extent along the flow, volume centroid) and moment centre;
- run_experiment restates solve + record_result (src/app/aero/aero_solve.cpp:166-261) for the
  lattice Boltzmann solvers (3D bodies on LbmAeroSolver3D, 2D NACA sections on
  LbmAeroSolver2D): options from the record's settings, steps until the solver finishes, and
  the AeroResultRecord (coefficients, Reynolds and Mach numbers, reference values, iterations,
  convergence, warning codes, residuals, force history, Strouhal number of a section);
- ExperimentRunner restates the command protocol of AeroRunner (src/app/aero/aero_runner.h):
  start, pause, steps per frame, run to convergence, one frame, stop; progress published after
  each change;
- serve_request is the JSON request/response form: an experiment record in, the same record
  with its result out.

The records' other solvers (panel, Newtonian, "auto") and imported models are not part of this
library: such records are refused with a message.
"""
from __future__ import annotations

import dataclasses
import enum
import math
import queue
import threading
import time
from typing import Optional

import numpy as np

from . import geometry as geo
from .aero import AeroError, AeroGeometry, LbmAeroSolver2D, LbmAeroSolver3D, LbmSolver3DOptions, LbmSolverOptions, ReferenceValues
from .experiment import (AeroExperiment, AeroResultRecord, AeroShape, GeometryRef, GeometrySource, ReferenceAreaKind,
                         SolverKind, parse_aero_experiment, save_aero_experiment)

REVOLUTION = geo.RevolutionResolution(48, 0.08)  # kRevolutionAround, kRevolutionArcStep
NACA_STATIONS = 48
NACA_SPAN_DIVISIONS = 12
SECTION_PREVIEW_SPAN = 0.6  # kSectionPreviewSpan: span of a section's preview wing, chords
REFERENCE_RASTER = 512
STEPS_PER_CALL_2D = 50
STEPS_PER_CALL_3D = 20
MAX_CELLS_3D = 1 << 22


class RunnerError(RuntimeError):
    pass


# ---- geometry of a primitive record --------------------------------------------------------

@dataclasses.dataclass
class ReferenceQuantities:
    wetted_area: float
    frontal_area: float
    planform_area: float
    reference_length: float
    volume: float
    centre_of_mass: tuple


@dataclasses.dataclass
class PreparedGeometry:
    ref: GeometryRef
    mesh: geo.TriangleMesh
    quantities: ReferenceQuantities
    moment_centre: tuple
    label: str
    section: Optional[np.ndarray] = None  # (x, z) points of a 2D section, metres
    two_dimensional: bool = False


def _rotated(mesh: geo.TriangleMesh, rows) -> geo.TriangleMesh:
    return geo.TriangleMesh(mesh.vertices @ np.asarray(rows, dtype=np.float64).T, mesh.triangles)


AXIS_Z_TO_Y = ((1, 0, 0), (0, 0, 1), (0, -1, 0))         # -90 degrees about x
AXIS_Z_TO_MINUS_X = ((0, 0, -1), (0, 1, 0), (1, 0, 0))   # -90 degrees about y
AXIS_Z_TO_X = ((0, 0, 1), (0, 1, 0), (-1, 0, 0))         # +90 degrees about y


def _centre_on_bounds(mesh: geo.TriangleMesh) -> geo.TriangleMesh:
    lo, hi = mesh.bounds()
    return geo.TriangleMesh(mesh.vertices - (lo + hi) * 0.5, mesh.triangles)


def _edge(a, b, u, v):
    return (b[0] - a[0]) * (v - a[1]) - (b[1] - a[1]) * (u - a[0])


def projected_area(mesh: geo.TriangleMesh, direction, resolution: int = REFERENCE_RASTER) -> float:
    """Silhouette area along `direction`: cell-centre coverage of the projected triangles on a
    raster of `resolution` cells along the larger side (reference_quantities.cpp:62-131)."""
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    ad = np.abs(d)
    helper = (np.array([1.0, 0.0, 0.0]) if ad[0] <= ad[1] and ad[0] <= ad[2]
              else (np.array([0.0, 1.0, 0.0]) if ad[1] <= ad[2] else np.array([0.0, 0.0, 1.0])))
    u = np.cross(d, helper)
    u = u / np.linalg.norm(u)
    v = np.cross(d, u)
    p = np.stack([mesh.vertices @ u, mesh.vertices @ v], axis=1)
    lo, hi = p.min(axis=0), p.max(axis=0)
    side = max(hi[0] - lo[0], hi[1] - lo[1])
    if not side > 0.0:
        return 0.0
    h = side / resolution
    n = [min(max(int(math.ceil(e / h)), 1), resolution) for e in (hi[0] - lo[0], hi[1] - lo[1])]
    covered = np.zeros((n[1], n[0]), dtype=bool)

    def cells(a, b, origin, count):
        first = math.ceil((a - origin) / h - 0.5)
        last = math.floor((b - origin) / h - 0.5)
        top = count - 1.0
        return int(min(max(first, 0.0), top)), int(min(max(last, -1.0), top) + 1.0)

    for t in mesh.triangles:
        a, b, c = p[t[0]], p[t[1]], p[t[2]]
        if _edge(a, b, c[0], c[1]) == 0.0:
            continue
        i0, i1 = cells(min(a[0], b[0], c[0]), max(a[0], b[0], c[0]), lo[0], n[0])
        j0, j1 = cells(min(a[1], b[1], c[1]), max(a[1], b[1], c[1]), lo[1], n[1])
        if i1 <= i0 or j1 <= j0:
            continue
        cu = lo[0] + (np.arange(i0, i1) + 0.5) * h
        cv = lo[1] + (np.arange(j0, j1) + 0.5) * h
        U, V = np.meshgrid(cu, cv)
        e0, e1, e2 = _edge(a, b, U, V), _edge(b, c, U, V), _edge(c, a, U, V)
        inside = ((e0 >= 0) & (e1 >= 0) & (e2 >= 0)) | ((e0 <= 0) & (e1 <= 0) & (e2 <= 0))
        covered[j0:j1, i0:i1] |= inside
    return float(covered.sum()) * h * h


def reference_quantities(mesh: geo.TriangleMesh, resolution: int = REFERENCE_RASTER) -> ReferenceQuantities:
    """reference_quantities (reference_quantities.cpp:133-158), flow along +x and lift along +z."""
    v = mesh.vertices
    t = mesh.triangles.astype(np.int64)
    lo, hi = mesh.bounds()
    o = (lo + hi) * 0.5
    a, b, c = v[t[:, 0]] - o, v[t[:, 1]] - o, v[t[:, 2]] - o
    v6 = np.einsum("ij,ij->i", a, np.cross(b, c))
    volume6 = float(v6.sum())
    centroid = o + ((a + b + c) * v6[:, None]).sum(axis=0) / (4.0 * volume6) if volume6 != 0.0 else o
    return ReferenceQuantities(
        wetted_area=mesh.surface_area(),
        frontal_area=projected_area(mesh, (1.0, 0.0, 0.0), resolution),
        planform_area=projected_area(mesh, (0.0, 0.0, 1.0), resolution),
        reference_length=float(hi[0] - lo[0]),
        volume=volume6 / 6.0,
        centre_of_mass=tuple(float(x) for x in centroid))


def _metres(value: float) -> str:
    return f"{value:.3g} m"


def _capsule_shape(name: str, diameter: float) -> geo.CapsuleShape:
    base = (geo.CapsuleShape(5.03, 6.04, 0.25, 32.5 * math.pi / 180.0, 1.8) if name == "orion"
            else geo.CapsuleShape.apollo())
    k = diameter / base.max_diameter
    return geo.CapsuleShape(base.max_diameter * k, base.nose_radius * k, base.shoulder_radius * k,
                            base.afterbody_half_angle, base.top_diameter * k)


def prepare_geometry(ref: GeometryRef) -> PreparedGeometry:
    """The body of a primitive record (build_primitive, aero_geometry.cpp:84-188)."""
    if ref.source != GeometrySource.primitive:
        raise RunnerError("this geometry is a user model: import it with the reference's tools")
    d = ref.dimensions
    naca = False
    s = ref.shape
    if s == AeroShape.sphere:
        mesh = geo.sphere_mesh(0.5 * d.diameter_m, REVOLUTION)
        label = f"Sphere, d {_metres(d.diameter_m)}"
    elif s == AeroShape.cube:
        mesh = geo.box_mesh(d.side_m, d.side_m, d.side_m)
        label = f"Cube, side {_metres(d.side_m)}"
    elif s == AeroShape.cylinder:
        mesh = _rotated(geo.cylinder_mesh(0.5 * d.diameter_m, d.length_m, REVOLUTION), AXIS_Z_TO_Y)
        label = f"Cylinder (cross-flow), d {_metres(d.diameter_m)}, span {_metres(d.length_m)}"
    elif s == AeroShape.cone:
        mesh = _rotated(geo.cone_mesh(0.5 * d.diameter_m, d.length_m, REVOLUTION), AXIS_Z_TO_MINUS_X)
        label = f"Cone, d {_metres(d.diameter_m)}, length {_metres(d.length_m)}"
    elif s == AeroShape.flat_plate:
        mesh = geo.box_mesh(d.chord_m, d.span_m, d.thickness_m)
        label = f"Flat plate, {_metres(d.chord_m)} x {_metres(d.span_m)}"
    elif s == AeroShape.wedge:
        mesh = geo.wedge_mesh(d.length_m, d.half_angle_deg * math.pi / 180.0, d.span_m)
        label = f"Wedge, {_metres(d.length_m)} long"
    elif s == AeroShape.capsule:
        mesh = _rotated(geo.capsule_mesh(_capsule_shape(ref.capsule, d.diameter_m), REVOLUTION), AXIS_Z_TO_X)
        label = ("Orion-type capsule, d " if ref.capsule == "orion" else "Apollo-type capsule, d ") + _metres(d.diameter_m)
    elif s == AeroShape.naca_wing:
        digits = ref.naca
        mesh = geo.naca_wing(int(digits[0]) / 100.0, int(digits[1]) / 10.0, int(digits[2:]) / 100.0, d.chord_m,
                             d.span_m, NACA_STATIONS, NACA_SPAN_DIVISIONS)
        label = f"NACA {digits} wing, chord {_metres(d.chord_m)}, span {_metres(d.span_m)}"
        naca = True
    elif s == AeroShape.naca_section:
        # naca_section_geometry (src/aero/panel2d.cpp:145-160): the section scaled by the chord,
        # moment centre at the quarter chord; a short wing of the section for display.
        digits = ref.naca
        m, p, t = int(digits[0]) / 100.0, int(digits[1]) / 10.0, int(digits[2:]) / 100.0
        section = geo.naca_section(m, p, t, NACA_STATIONS) * d.chord_m
        mesh = geo.naca_wing(m, p, t, d.chord_m, SECTION_PREVIEW_SPAN * d.chord_m, NACA_STATIONS, 1)
        label = f"NACA {digits} section, chord {_metres(d.chord_m)}"
        return PreparedGeometry(ref, mesh, reference_quantities(mesh), (0.25 * d.chord_m, 0.0, 0.0), label,
                                section=section, two_dimensional=True)
    else:
        raise RunnerError(f"the {s.value} primitive is not built by this library (sphere, cube, cylinder, cone, "
                          "flat_plate, wedge, capsule and naca_wing are)")
    lo, hi = mesh.bounds()
    mesh = _centre_on_bounds(mesh)
    q = reference_quantities(mesh)
    if naca:
        moment = (float(lo[0] - 0.5 * (lo[0] + hi[0]) + 0.25 * d.chord_m), 0.0, 0.0)
    else:
        moment = q.centre_of_mass
    return PreparedGeometry(ref, mesh, q, moment, label)


def reference_values(prepared: PreparedGeometry, settings) -> ReferenceValues:
    """aero_geometry.cpp:239-255: per unit span for a section (chord x 1 m, the chord as the
    length), else the record's reference area of the body."""
    if prepared.two_dimensional:
        chord = prepared.ref.dimensions.chord_m
        length = settings.reference_length_m if settings.reference_length_m is not None else chord
        return ReferenceValues(chord, length, prepared.moment_centre)
    q = prepared.quantities
    area = {ReferenceAreaKind.frontal: q.frontal_area, ReferenceAreaKind.planform: q.planform_area,
            ReferenceAreaKind.wetted: q.wetted_area}[settings.reference_area]
    length = settings.reference_length_m if settings.reference_length_m is not None else q.reference_length
    return ReferenceValues(area, length, prepared.moment_centre)


# ---- one run -------------------------------------------------------------------------------

def _check_runnable(e: AeroExperiment) -> None:
    if e.solver.solver is not None and e.solver.solver != SolverKind.lbm:
        raise RunnerError(f"solver '{e.solver.solver.value}' is not part of this library (lattice Boltzmann only)")
    if e.solver.solver is None:
        raise RunnerError("solver 'auto' picks among solvers this library does not have: name 'lbm'")


def solver_options(e: AeroExperiment, storage: str = "f32",
                   steps_per_call: int = STEPS_PER_CALL_3D) -> LbmSolver3DOptions:
    """make_solver for the 3D lattice (aero_solve.cpp:98-108)."""
    o = LbmSolver3DOptions(storage=storage)
    if e.solver.lbm_cells_per_length > 0.0:
        o.cells_per_length = e.solver.lbm_cells_per_length
    if e.solver.lbm_duration > 0.0:
        o.duration = e.solver.lbm_duration
    o.steps_per_call = steps_per_call
    o.max_cells = MAX_CELLS_3D
    return o


def solver_options_2d(e: AeroExperiment, storage: str = "f32",
                      steps_per_call: int = STEPS_PER_CALL_2D) -> LbmSolverOptions:
    """make_solver for the 2D lattice (aero_solve.cpp:91-98)."""
    o = LbmSolverOptions(storage=storage)
    if e.solver.lbm_cells_per_length > 0.0:
        o.cells_per_length = e.solver.lbm_cells_per_length
    if e.solver.lbm_duration > 0.0:
        o.duration = e.solver.lbm_duration
    o.steps_per_call = steps_per_call
    return o


def steps_per_call_for(two_dimensional: bool, limit: Optional[int] = None) -> int:
    """lbm_steps_per_call_for (aero_solve.cpp:57-70): the per-call batch, or its largest divisor
    of a step limit so that a limited run stops on it."""
    batch = STEPS_PER_CALL_2D if two_dimensional else STEPS_PER_CALL_3D
    if not limit:
        return batch
    for n in range(batch, 1, -1):
        if limit % n == 0:
            return n
    return 1


def make_solver(prepared: PreparedGeometry, e: AeroExperiment, storage: str = "f32",
                steps_per_call: Optional[int] = None):
    """The adapter of a record's geometry with its options (make_solver, aero_solve.cpp:72-111)."""
    if prepared.two_dimensional:
        per_call = steps_per_call if steps_per_call else STEPS_PER_CALL_2D
        return LbmAeroSolver2D(solver_options_2d(e, storage, per_call))
    per_call = steps_per_call if steps_per_call else STEPS_PER_CALL_3D
    return LbmAeroSolver3D(solver_options(e, storage, per_call))


def aero_geometry(prepared: PreparedGeometry, reference: ReferenceValues) -> AeroGeometry:
    """to_aero_geometry (aero_geometry.cpp:257-268): the section of a 2D record, else the mesh."""
    if prepared.two_dimensional:
        return AeroGeometry(reference=reference, section=prepared.section)
    return AeroGeometry(prepared.mesh, reference)


def _finite(x: float) -> bool:
    return math.isfinite(x)


def record_result(solver, e: AeroExperiment, reference: ReferenceValues) -> AeroResultRecord:
    """record_result (aero_solve.cpp:166-211) of a finished or interrupted run."""
    r = solver.result()
    flow = e.conditions.flow()
    t, cd, cl = solver.history()
    two_d = isinstance(solver, LbmAeroSolver2D)
    warnings = [w.code for w in r.warnings]
    if two_d and e.conditions.sideslip_deg != 0.0:
        warnings.append("workspace.sideslip_2d")  # summarize_flow (aero_flow.cpp:82-85)
    out = AeroResultRecord(solver=SolverKind.lbm, cd=r.cd, cl=r.cl, cy=r.cy, cm=r.cm,
                           reynolds=flow.reynolds(reference.length), mach=flow.mach(),
                           reference_area_m2=reference.area, reference_length_m=reference.length,
                           iterations=int(r.iterations), converged=bool(r.converged),
                           strouhal=solver.strouhal() if two_d else None,
                           warnings=warnings, residuals=[float(x) for x in r.residuals],
                           history_step=[float(x) for x in t], history_cd=[float(x) for x in cd],
                           history_cl=[float(x) for x in cl])
    for key in ("cd", "cl", "cy", "cm"):
        if not _finite(getattr(out, key)):
            setattr(out, key, 0.0)
            out.converged = False
    for series in (out.residuals, out.history_cd, out.history_cl):
        for k, v in enumerate(series):
            if not _finite(v):
                series[k] = 0.0
    return out


def run_experiment(e: AeroExperiment, storage: str = "f32", max_lattice_steps: Optional[int] = None,
                   steps_per_call: Optional[int] = None) -> AeroExperiment:
    """solve (aero_solve.cpp:226-261): the record with its result."""
    _check_runnable(e)
    prepared = prepare_geometry(e.geometry)
    reference = reference_values(prepared, e.solver)
    if steps_per_call is None:
        steps_per_call = steps_per_call_for(prepared.two_dimensional, max_lattice_steps)
    solver = make_solver(prepared, e, storage, steps_per_call)
    try:
        solver.setup(aero_geometry(prepared, reference), e.conditions.flow())
        finished = False
        while not finished:
            finished = solver.step()
            if max_lattice_steps is not None and solver.result().iterations >= max_lattice_steps:
                break
        out = dataclasses.replace(e, result=record_result(solver, e, reference))
    except (AeroError, RuntimeError) as err:
        raise RunnerError(f"lbm: {err}") from err
    finally:
        solver.close()
    return out


def serve_request(text: str, storage: str = "f32", max_lattice_steps: Optional[int] = None) -> str:
    """A JSON experiment record in, the record with its result out (the reference's file
    format, src/scene/aero_experiment.cpp)."""
    return save_aero_experiment(run_experiment(parse_aero_experiment(text), storage, max_lattice_steps))


# ---- the solver thread and its commands (aero_runner.h) ------------------------------------

@dataclasses.dataclass
class StartRun:
    experiment: AeroExperiment
    to_convergence: bool = False


@dataclasses.dataclass
class SetPaused:
    paused: bool = True


@dataclasses.dataclass
class SetStepsPerFrame:
    steps: int = 100


@dataclasses.dataclass
class SetToConvergence:
    enabled: bool = True


@dataclasses.dataclass
class StepFrame:
    pass


@dataclasses.dataclass
class StopRun:
    pass


class RunState(enum.Enum):
    idle = "idle"
    setting_up = "setting up"
    running = "running"
    paused = "paused"
    finished = "finished"
    failed = "failed"


@dataclasses.dataclass
class RunStatus:
    state: RunState = RunState.idle
    run: int = 0
    to_convergence: bool = False
    steps_per_frame: int = 100
    message: str = ""
    step_ms: float = 0.0
    steps_per_second: float = 0.0


@dataclasses.dataclass
class Progress:
    status: RunStatus
    record: Optional[AeroExperiment] = None  # the experiment with its latest result


class ExperimentRunner:
    """A solver thread fed by send(); wake() asks for one frame of steps in the per-frame mode
    (steps_per_frame lattice steps), the run-to-convergence mode steps back to back. latest()
    returns the newest published progress."""

    def __init__(self, storage: str = "f32"):
        self._storage = storage
        self._commands: queue.Queue = queue.Queue()
        self._frames = threading.Semaphore(0)
        self._lock = threading.Lock()
        self._progress = Progress(RunStatus())
        self._status = RunStatus()
        self._solver = None  # LbmAeroSolver3D or LbmAeroSolver2D
        self._per_call = STEPS_PER_CALL_3D
        self._experiment: Optional[AeroExperiment] = None
        self._reference: Optional[ReferenceValues] = None
        self._stop = threading.Event()
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self._thread.start()

    # any thread
    def send(self, command) -> None:
        self._commands.put(command)
        self._frames.release()

    def wake(self) -> None:
        self._frames.release()

    def latest(self) -> Progress:
        with self._lock:
            return self._progress

    def wait(self, states, timeout: float = 600.0) -> Progress:
        """Blocks until the published state is one of `states` (tests and scripts)."""
        end = time.monotonic() + timeout
        while time.monotonic() < end:
            p = self.latest()
            if p.status.state in states:
                return p
            time.sleep(0.01)
        raise TimeoutError(f"runner still {self.latest().status.state.value}")

    def close(self) -> None:
        self._stop.set()
        self._frames.release()
        self._thread.join()
        self._drop()

    # solver thread
    def _publish(self, with_result: bool) -> None:
        record = None
        if with_result and self._solver is not None and self._experiment is not None:
            record = dataclasses.replace(self._experiment,
                                         result=record_result(self._solver, self._experiment, self._reference))
        elif self._progress.record is not None and self._progress.status.run == self._status.run:
            record = self._progress.record
        with self._lock:
            self._progress = Progress(dataclasses.replace(self._status), record)

    def _drop(self) -> None:
        if self._solver is not None:
            self._solver.close()
        self._solver = None

    def _fail(self, what: str) -> None:
        self._drop()
        self._status.state = RunState.failed
        self._status.message = what
        self._publish(False)

    def _start(self, cmd: StartRun) -> None:
        self._drop()
        self._status = RunStatus(RunState.setting_up, self._status.run + 1, cmd.to_convergence,
                                 cmd.experiment.solver.steps_per_frame)
        self._experiment = cmd.experiment
        self._publish(False)
        try:
            _check_runnable(cmd.experiment)
            prepared = prepare_geometry(cmd.experiment.geometry)
            self._reference = reference_values(prepared, cmd.experiment.solver)
            self._per_call = STEPS_PER_CALL_2D if prepared.two_dimensional else STEPS_PER_CALL_3D
            self._solver = make_solver(prepared, cmd.experiment, self._storage)
        except Exception as err:  # refusals before the solver exists end the run with their message
            self._fail(str(err))
            return
        try:
            self._solver.setup(aero_geometry(prepared, self._reference), cmd.experiment.conditions.flow())
        except Exception as err:  # setup refusals (AeroRunner::start, aero_runner.cpp:205-211)
            self._fail(f"Lattice Boltzmann setup failed: {err}")
            return
        self._status.state = RunState.running
        self._publish(False)

    def _apply(self, cmd) -> None:
        if isinstance(cmd, StartRun):
            self._start(cmd)
        elif isinstance(cmd, SetPaused):
            if self._status.state in (RunState.running, RunState.paused):
                self._status.state = RunState.paused if cmd.paused else RunState.running
                self._publish(False)
        elif isinstance(cmd, SetStepsPerFrame):
            self._status.steps_per_frame = max(1, int(cmd.steps))
            self._publish(False)
        elif isinstance(cmd, SetToConvergence):
            self._status.to_convergence = bool(cmd.enabled)
            self._publish(False)
        elif isinstance(cmd, StepFrame):
            if self._status.state in (RunState.running, RunState.paused):
                self._frame()
                if self._status.state == RunState.running:
                    self._status.state = RunState.paused
                self._publish(True)
        elif isinstance(cmd, StopRun):
            self._drop()
            self._status.state = RunState.idle
            self._publish(False)

    def _one_call(self) -> bool:
        """One solver call; False once the run cannot continue."""
        t0 = time.perf_counter()
        before = self._solver.result().iterations
        try:
            finished = self._solver.step()
        except Exception as err:
            self._fail(str(err))
            return False
        dt = time.perf_counter() - t0
        steps = self._solver.result().iterations - before
        self._status.step_ms = dt * 1e3
        self._status.steps_per_second = steps / dt if dt > 0 else 0.0
        if finished:
            self._status.state = RunState.finished
            self._publish(True)
            return False
        return True

    def _frame(self) -> None:
        calls = max(1, -(-self._status.steps_per_frame // self._per_call))
        for _ in range(calls):
            if not self._one_call():
                return

    def _loop(self) -> None:
        while not self._stop.is_set():
            try:
                while True:
                    self._apply(self._commands.get_nowait())
            except queue.Empty:
                pass
            if self._status.state == RunState.running and self._status.to_convergence:
                if self._one_call() and self._status.state == RunState.running:
                    self._publish(True)
                continue
            self._frames.acquire()
            if self._stop.is_set():
                break
            if self._commands.empty() and self._status.state == RunState.running:
                self._frame()
                if self._status.state == RunState.running:
                    self._publish(True)


__all__ = ["prepare_geometry", "reference_quantities", "projected_area", "reference_values", "run_experiment",
           "serve_request", "record_result", "solver_options", "solver_options_2d", "steps_per_call_for",
           "make_solver", "aero_geometry", "ExperimentRunner", "StartRun", "SetPaused",
           "SetStepsPerFrame", "SetToConvergence", "StepFrame", "StopRun", "RunState", "RunStatus", "Progress",
           "PreparedGeometry", "ReferenceQuantities", "RunnerError"]
