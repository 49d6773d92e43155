"""The lattices behind Orrery's aerodynamic solver interface (IAeroSolver): LbmAeroSolver3D and
LbmAeroSolver2D and their inputs and outputs, named as in the reference
(src/aero/lbm3d/include/orrery/aero/lbm3d/lbm_aero_solver3d.h,
src/aero/lbm/include/orrery/aero/lbm/lbm_aero_solver.h,
src/aero/include/orrery/aero/{aero_solver,flow_conditions,aero_result}.h).

The work is done by the library's adapters (lbm_aero3d_* and lbm_aero2d_* in include/orrery_lbm.h):
the setups follow the reference's setup() operation for operation; the per-call outputs are
computed on the device.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np

from . import abi
from .config import Lbm2DConfig, Lbm3DConfig, Outlet, SolidWall, Wall, WallKind, XBoundary
from .geometry import TriangleMesh
from .masks import SolidMask2D, SolidMask3D


class AeroError(ValueError):
    pass


class LesMode:
    automatic = 0
    on = 1
    off = 2


@dataclasses.dataclass
class FlowConditions:
    """Free-stream state in SI units (flow_conditions.h:14-55)."""

    speed: float = 0.0
    density: float = 1.225
    viscosity: float = 1.789e-5
    temperature: float = 288.15
    angle_of_attack: float = 0.0
    sideslip: float = 0.0
    gamma: float = 1.4
    gas_constant: float = 287.05

    def speed_of_sound(self) -> float:
        return math.sqrt(self.gamma * self.gas_constant * self.temperature)

    def mach(self) -> float:
        return self.speed / self.speed_of_sound()

    def reynolds(self, length: float) -> float:
        return self.density * self.speed * length / self.viscosity

    def dynamic_pressure(self) -> float:
        return 0.5 * self.density * self.speed * self.speed

    def drag_direction(self) -> np.ndarray:
        cb = math.cos(self.sideslip)
        return np.array([math.cos(self.angle_of_attack) * cb, math.sin(self.sideslip),
                         math.sin(self.angle_of_attack) * cb])

    def lift_direction(self) -> np.ndarray:
        v = np.cross(self.drag_direction(), [0.0, 1.0, 0.0])
        n = math.hypot(*v)
        return v / n if n > 0 else v

    def side_direction(self) -> np.ndarray:
        return np.cross(self.lift_direction(), self.drag_direction())

    def to_c(self) -> abi.CFlowConditions:
        return abi.CFlowConditions(*(float(getattr(self, f.name)) for f in dataclasses.fields(self)))


@dataclasses.dataclass
class ReferenceValues:
    area: float = 1.0
    length: float = 1.0
    moment_centre: tuple = (0.0, 0.0, 0.0)

    def to_c(self) -> abi.CReferenceValues:
        return abi.CReferenceValues(self.area, self.length, (ctypes.c_double * 3)(*self.moment_centre))


@dataclasses.dataclass
class AeroGeometry:
    """AeroGeometry (aero_solver.h): a closed body mesh for the 3D solver, or a closed section
    [n, 2] of (x, z) points for the 2D solver, with the reference values."""

    mesh: TriangleMesh | None = None
    reference: ReferenceValues = dataclasses.field(default_factory=ReferenceValues)
    section: np.ndarray | None = None


@dataclasses.dataclass
class LbmSolver3DOptions:
    """LbmSolver3DOptions (lbm_aero_solver3d.h:45-81); `storage` picks the lattice's population
    format (no reference counterpart)."""

    cells_per_length: float = 16.0
    lattice_velocity: float = 0.05
    upstream: float = 5.0
    downstream: float = 10.0
    width: float = 6.0
    sides: int = WallKind.slip
    wall: int = SolidWall.interpolated
    les: int = LesMode.automatic
    smagorinsky_constant: float = 0.1
    duration: float = 30.0
    average_window: float = 8.0
    start_ramp: float = 3.0
    sponge_strength: float = 0.1
    sponge_inlet: float = 0.0
    sponge_outlet: float = 2.0
    sponge_sides: float = 0.0
    steps_per_call: int = 200
    sample_every: int = 10
    max_cells: int = 0
    storage: str = "f32"

    def to_c(self) -> abi.CAeroOptions:
        return abi.CAeroOptions(self.cells_per_length, self.lattice_velocity, self.upstream, self.downstream,
                                self.width, int(self.sides), int(self.wall), int(self.les), self.smagorinsky_constant,
                                self.duration, self.average_window, self.start_ramp, self.sponge_strength,
                                self.sponge_inlet, self.sponge_outlet, self.sponge_sides, self.steps_per_call,
                                self.sample_every, self.max_cells, abi.STORAGE[self.storage])


@dataclasses.dataclass
class ValidityWarning:
    code: str
    message: str


@dataclasses.dataclass
class SurfaceField:
    points: np.ndarray
    normals: np.ndarray
    areas: np.ndarray
    cp: np.ndarray


@dataclasses.dataclass
class AeroResult:
    cd: float = 0.0
    cl: float = 0.0
    cy: float = 0.0
    cm: float = 0.0
    force: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    moment: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    surface: SurfaceField | None = None
    residuals: list = dataclasses.field(default_factory=list)
    iterations: int = 0
    converged: bool = False
    warnings: list = dataclasses.field(default_factory=list)

    def has_warning(self, code: str) -> bool:
        return any(w.code == code for w in self.warnings)


@dataclasses.dataclass
class LatticePlan:
    """What setup decided: lattice configuration, placement and step counts."""

    config: Lbm3DConfig
    origin: tuple
    centre: tuple
    first: tuple
    count: tuple
    axes: np.ndarray
    dx: float
    dt: float
    reynolds: float
    lattice_velocity: float
    cells_per_length: float
    total_steps: int
    window_samples: int
    ramp_steps: float
    blockage: float
    surface_area: float
    solid_nodes: int


def _config_from_c(c: abi.CConfig3D) -> Lbm3DConfig:
    from .config import Smagorinsky, Sponge
    wall = lambda w: Wall(WallKind(w.kind), w.velocity_x)  # noqa: E731
    return Lbm3DConfig(nx=c.nx, ny=c.ny, nz=c.nz, tau=c.tau, x_boundary=XBoundary(c.x_boundary), outlet=Outlet(c.outlet),
                       y_low=wall(c.y_low), y_high=wall(c.y_high), z_low=wall(c.z_low), z_high=wall(c.z_high),
                       inlet_ux=c.inlet_ux, inlet_uy=c.inlet_uy, inlet_uz=c.inlet_uz, body_force_x=c.body_force_x,
                       body_force_y=c.body_force_y, body_force_z=c.body_force_z,
                       smagorinsky=Smagorinsky(bool(c.smagorinsky_enabled), c.smagorinsky_constant),
                       sponge=Sponge(c.sponge_strength, c.sponge_inlet, c.sponge_outlet, c.sponge_sides),
                       solid_wall=SolidWall(c.solid_wall))


class LbmAeroSolver3D:
    """IAeroSolver over the D3Q19 lattice (lbm_aero_solver3d.h:102-166)."""

    kind = "lattice_boltzmann"

    def __init__(self, options: LbmSolver3DOptions | None = None):
        self.options = options or LbmSolver3DOptions()
        self._lib = abi.load()
        h = ctypes.c_void_p()
        o = self.options.to_c()
        self._check(self._lib.lbm_aero3d_create(ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._result = AeroResult()
        self._set_up = False

    def _check(self, rc: int) -> None:
        if rc != abi.LBM_OK:
            raise AeroError(self._lib.lbm_aero3d_last_error().decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.lbm_aero3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, fn, geometry: AeroGeometry, conditions: FlowConditions) -> None:
        if geometry.mesh is None:
            raise AeroError("lbm3d: invalid mesh: mesh has no vertices or no triangles")
        handle = geometry.mesh._handle()
        with handle as m:
            r = geometry.reference.to_c()
            c = conditions.to_c()
            self._check(fn(self._h, m, ctypes.byref(r), ctypes.byref(c)))
        self.geometry = geometry
        self.conditions = conditions

    def plan(self, geometry: AeroGeometry, conditions: FlowConditions) -> LatticePlan:
        """setup() without the device lattice: the lattice, mask and warnings, host only."""
        self._set_up = False
        self._call(self._lib.lbm_aero3d_plan, geometry, conditions)
        self._result = AeroResult(warnings=self._warnings())
        return self.info()

    def setup(self, geometry: AeroGeometry, conditions: FlowConditions) -> None:
        self._set_up = False
        self._call(self._lib.lbm_aero3d_setup, geometry, conditions)
        self._set_up = True
        self._result = AeroResult(warnings=self._warnings())

    def step(self) -> bool:
        if not self._set_up:
            raise RuntimeError("lbm3d: step() called before setup()")
        done = ctypes.c_int32()
        self._check(self._lib.lbm_aero3d_step(self._h, ctypes.byref(done)))
        self._result = self._read_result()
        return bool(done.value)

    def run(self, max_calls: int = 10 ** 6) -> AeroResult:
        for _ in range(max_calls):
            if self.step():
                return self._result
        raise RuntimeError("lbm3d: run did not finish")

    def result(self) -> AeroResult:
        return self._result

    def _warnings(self) -> list:
        n = ctypes.c_size_t()
        self._check(self._lib.lbm_aero3d_warning_count(self._h, ctypes.byref(n)))
        out = []
        for k in range(n.value):
            code = ctypes.create_string_buffer(64)
            msg = ctypes.create_string_buffer(1024)
            self._check(self._lib.lbm_aero3d_warning(self._h, k, code, 64, msg, 1024))
            out.append(ValidityWarning(code.value.decode(), msg.value.decode()))
        return out

    def _read_result(self) -> AeroResult:
        r = abi.CAeroResult()
        self._check(self._lib.lbm_aero3d_get_result(self._h, ctypes.byref(r)))
        n = ctypes.c_size_t()
        self._check(self._lib.lbm_aero3d_surface(self._h, None, None, None, None, 0, ctypes.byref(n)))
        pts = np.empty((n.value, 3))
        nrm = np.empty((n.value, 3))
        areas = np.empty(n.value)
        cp = np.empty(n.value)
        self._check(self._lib.lbm_aero3d_surface(self._h, abi.ptr(pts, ctypes.c_double), abi.ptr(nrm, ctypes.c_double),
                                                 abi.ptr(areas, ctypes.c_double), abi.ptr(cp, ctypes.c_double),
                                                 n.value, ctypes.byref(n)))
        self._check(self._lib.lbm_aero3d_residuals(self._h, None, 0, ctypes.byref(n)))
        res = np.empty(n.value)
        self._check(self._lib.lbm_aero3d_residuals(self._h, abi.ptr(res, ctypes.c_double), n.value,
                                                   ctypes.byref(n)))
        return AeroResult(cd=r.cd, cl=r.cl, cy=r.cy, cm=r.cm, force=np.array(r.force[:]), moment=np.array(r.moment[:]),
                          surface=SurfaceField(pts, nrm, areas, cp), residuals=res.tolist(), iterations=r.iterations,
                          converged=bool(r.converged), warnings=self._warnings())

    def history(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(time, drag, lift) of the samples: CoefficientHistory."""
        t, d, l, _ = self._history()
        return t, d, l

    def side_history(self) -> np.ndarray:
        return self._history()[3]

    def _history(self):
        n = ctypes.c_size_t()
        self._check(self._lib.lbm_aero3d_history(self._h, None, None, None, None, 0, ctypes.byref(n)))
        arrays = [np.empty(n.value) for _ in range(4)]
        self._check(self._lib.lbm_aero3d_history(self._h, *[abi.ptr(a, ctypes.c_double) for a in arrays], n.value,
                                                 ctypes.byref(n)))
        return arrays

    def moving_average(self) -> tuple[float, float]:
        """CoefficientHistory::moving_average (force_history.cpp:20-34)."""
        t, d, l = self.history()
        n = min(self.info().window_samples, len(t))
        if n == 0:
            return 0.0, 0.0
        sd = 0.0
        sl = 0.0
        for k in range(len(t) - n, len(t)):
            sd += d[k]
            sl += l[k]
        return sd / n, sl / n

    def instantaneous(self) -> tuple[float, float, float]:
        c = (ctypes.c_double * 3)()
        self._check(self._lib.lbm_aero3d_instantaneous(self._h, c))
        return tuple(c[:])

    def info(self) -> LatticePlan:
        i = abi.CAeroInfo()
        self._check(self._lib.lbm_aero3d_get_info(self._h, ctypes.byref(i)))
        return LatticePlan(config=_config_from_c(i.config), origin=tuple(i.origin[:]), centre=tuple(i.centre[:]),
                           first=tuple(i.first[:]), count=tuple(i.count[:]), axes=np.array(i.axes[:]).reshape(3, 3),
                           dx=i.dx, dt=i.dt, reynolds=i.reynolds, lattice_velocity=i.lattice_velocity,
                           cells_per_length=i.cells_per_length, total_steps=i.total_steps,
                           window_samples=i.window_samples, ramp_steps=i.ramp_steps, blockage=i.blockage,
                           surface_area=i.surface_area, solid_nodes=i.solid_nodes)

    def mask(self) -> SolidMask3D:
        info = self.info()
        c = info.config
        solid = np.empty(c.cells, dtype=np.uint8)
        distance = np.empty(c.cells) if c.solid_wall == SolidWall.interpolated else None
        self._check(self._lib.lbm_aero3d_mask(self._h, abi.ptr(solid, ctypes.c_uint8), abi.ptr(distance, ctypes.c_double)))
        return SolidMask3D(c.nx, c.ny, c.nz, solid, distance)

    def total_steps(self) -> int:
        return self.info().total_steps

    def blockage(self) -> float:
        return self.info().blockage

    def lattice_handle(self) -> ctypes.c_void_p:
        h = ctypes.c_void_p()
        self._check(self._lib.lbm_aero3d_lattice(self._h, ctypes.byref(h)))
        return h

    def velocity_field(self) -> np.ndarray:
        """SI velocity on the lattice [nz, ny, nx, 3] (zero angle of attack and sideslip only)."""
        c = self.info().config
        out = np.empty((c.cells, 3))
        self._check(self._lib.lbm_aero3d_velocity(self._h, abi.ptr(out, ctypes.c_double)))
        return out.reshape(c.nz, c.ny, c.nx, 3)


# ---- the 2D adapter ------------------------------------------------------------------------

@dataclasses.dataclass
class LbmSolverOptions:
    """LbmSolverOptions (lbm_aero_solver.h:38-52); `storage` picks the lattice's population format
    (no reference counterpart)."""

    cells_per_length: float = 40.0
    lattice_velocity: float = 0.05
    upstream: float = 8.0
    downstream: float = 16.0
    height: float = 16.0
    les: int = LesMode.automatic
    smagorinsky_constant: float = 0.1
    duration: float = 60.0
    average_window: float = 6.0
    steps_per_call: int = 500
    sample_every: int = 10
    storage: str = "f32"

    def to_c(self) -> abi.CAero2DOptions:
        return abi.CAero2DOptions(self.cells_per_length, self.lattice_velocity, self.upstream, self.downstream,
                                  self.height, int(self.les), self.smagorinsky_constant, self.duration,
                                  self.average_window, self.steps_per_call, self.sample_every,
                                  abi.STORAGE[self.storage])


@dataclasses.dataclass
class LatticePlan2D:
    """What the 2D setup decided: lattice configuration, placement and step counts."""

    config: Lbm2DConfig
    origin: tuple
    centre: tuple
    dx: float
    dt: float
    reynolds: float
    lattice_velocity: float
    cells_per_length: float
    total_steps: int
    window_samples: int
    blockage: float
    solid_nodes: int


def _config2_from_c(c: abi.CConfig2D) -> Lbm2DConfig:
    from .config import Smagorinsky
    wall = lambda w: Wall(WallKind(w.kind), w.velocity_x)  # noqa: E731
    return Lbm2DConfig(nx=c.nx, ny=c.ny, tau=c.tau, x_boundary=XBoundary(c.x_boundary), outlet=Outlet(c.outlet),
                       bottom=wall(c.bottom), top=wall(c.top), inlet_ux=c.inlet_ux, inlet_uy=c.inlet_uy,
                       body_force_x=c.body_force_x, body_force_y=c.body_force_y,
                       smagorinsky=Smagorinsky(bool(c.smagorinsky_enabled), c.smagorinsky_constant))


class LbmAeroSolver2D:
    """IAeroSolver over the D2Q9 lattice (lbm_aero_solver.h:54-106): a closed section in the x-z
    plane, per unit span."""

    kind = "lattice_boltzmann"

    def __init__(self, options: LbmSolverOptions | None = None):
        self.options = options or LbmSolverOptions()
        self._lib = abi.load()
        h = ctypes.c_void_p()
        o = self.options.to_c()
        self._check(self._lib.lbm_aero2d_create(ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._result = AeroResult()
        self._set_up = False

    def _check(self, rc: int) -> None:
        if rc != abi.LBM_OK:
            raise AeroError(self._lib.lbm_aero2d_last_error().decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.lbm_aero2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, fn, geometry: AeroGeometry, conditions: FlowConditions) -> None:
        section = np.zeros((0, 2)) if geometry.section is None else geometry.section
        pts = np.ascontiguousarray(np.asarray(section, dtype=np.float64).reshape(-1, 2))
        r = geometry.reference.to_c()
        c = conditions.to_c()
        self._check(fn(self._h, abi.ptr(pts, ctypes.c_double), len(pts), ctypes.byref(r), ctypes.byref(c)))
        self.geometry = geometry
        self.conditions = conditions

    def plan(self, geometry: AeroGeometry, conditions: FlowConditions) -> LatticePlan2D:
        """setup() without the device lattice: the lattice, mask and warnings, host only."""
        self._set_up = False
        self._call(self._lib.lbm_aero2d_plan, geometry, conditions)
        self._result = AeroResult(warnings=self._warnings())
        return self.info()

    def setup(self, geometry: AeroGeometry, conditions: FlowConditions) -> None:
        self._set_up = False
        self._call(self._lib.lbm_aero2d_setup, geometry, conditions)
        self._set_up = True
        self._result = AeroResult(warnings=self._warnings())

    def step(self) -> bool:
        if not self._set_up:
            raise RuntimeError("lbm: step() called before setup()")
        done = ctypes.c_int32()
        self._check(self._lib.lbm_aero2d_step(self._h, ctypes.byref(done)))
        self._result = self._read_result()
        return bool(done.value)

    def run(self, max_calls: int = 10 ** 7) -> AeroResult:
        for _ in range(max_calls):
            if self.step():
                return self._result
        raise RuntimeError("lbm: run did not finish")

    def result(self) -> AeroResult:
        return self._result

    def _warnings(self) -> list:
        n = ctypes.c_size_t()
        self._check(self._lib.lbm_aero2d_warning_count(self._h, ctypes.byref(n)))
        out = []
        for k in range(n.value):
            code = ctypes.create_string_buffer(64)
            msg = ctypes.create_string_buffer(1024)
            self._check(self._lib.lbm_aero2d_warning(self._h, k, code, 64, msg, 1024))
            out.append(ValidityWarning(code.value.decode(), msg.value.decode()))
        return out

    def _read_result(self) -> AeroResult:
        r = abi.CAeroResult()
        self._check(self._lib.lbm_aero2d_get_result(self._h, ctypes.byref(r)))
        n = ctypes.c_size_t()
        self._check(self._lib.lbm_aero2d_surface(self._h, None, None, None, None, 0, ctypes.byref(n)))
        pts = np.empty((n.value, 3))
        nrm = np.empty((n.value, 3))
        areas = np.empty(n.value)
        cp = np.empty(n.value)
        self._check(self._lib.lbm_aero2d_surface(self._h, abi.ptr(pts, ctypes.c_double), abi.ptr(nrm, ctypes.c_double),
                                                 abi.ptr(areas, ctypes.c_double), abi.ptr(cp, ctypes.c_double),
                                                 n.value, ctypes.byref(n)))
        self._check(self._lib.lbm_aero2d_residuals(self._h, None, 0, ctypes.byref(n)))
        res = np.empty(n.value)
        self._check(self._lib.lbm_aero2d_residuals(self._h, abi.ptr(res, ctypes.c_double), n.value,
                                                   ctypes.byref(n)))
        return AeroResult(cd=r.cd, cl=r.cl, cy=r.cy, cm=r.cm, force=np.array(r.force[:]), moment=np.array(r.moment[:]),
                          surface=SurfaceField(pts, nrm, areas, cp), residuals=res.tolist(), iterations=r.iterations,
                          converged=bool(r.converged), warnings=self._warnings())

    def history(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(time, drag, lift) of the samples: CoefficientHistory."""
        n = ctypes.c_size_t()
        self._check(self._lib.lbm_aero2d_history(self._h, None, None, None, 0, ctypes.byref(n)))
        arrays = [np.empty(n.value) for _ in range(3)]
        self._check(self._lib.lbm_aero2d_history(self._h, *[abi.ptr(a, ctypes.c_double) for a in arrays], n.value,
                                                 ctypes.byref(n)))
        return tuple(arrays)

    def moving_average(self) -> tuple[float, float]:
        """CoefficientHistory::moving_average (force_history.cpp:20-34)."""
        t, d, l = self.history()
        n = min(self.info().window_samples, len(t))
        if n == 0:
            return 0.0, 0.0
        sd = 0.0
        sl = 0.0
        for k in range(len(t) - n, len(t)):
            sd += d[k]
            sl += l[k]
        return sd / n, sl / n

    def instantaneous(self) -> tuple[float, float]:
        c = (ctypes.c_double * 2)()
        self._check(self._lib.lbm_aero2d_instantaneous(self._h, c))
        return tuple(c[:])

    def strouhal(self) -> float | None:
        """Shedding frequency f L / U of the lift over the averaging window, None when the lift
        shows no oscillation (lbm_aero_solver.cpp:249-263)."""
        st = ctypes.c_double()
        found = ctypes.c_int32()
        self._check(self._lib.lbm_aero2d_strouhal(self._h, ctypes.byref(st), ctypes.byref(found)))
        return st.value if found.value else None

    def info(self) -> LatticePlan2D:
        i = abi.CAero2DInfo()
        self._check(self._lib.lbm_aero2d_get_info(self._h, ctypes.byref(i)))
        return LatticePlan2D(config=_config2_from_c(i.config), origin=tuple(i.origin[:]), centre=tuple(i.centre[:]),
                             dx=i.dx, dt=i.dt, reynolds=i.reynolds, lattice_velocity=i.lattice_velocity,
                             cells_per_length=i.cells_per_length, total_steps=i.total_steps,
                             window_samples=i.window_samples, blockage=i.blockage, solid_nodes=i.solid_nodes)

    def mask(self) -> SolidMask2D:
        c = self.info().config
        solid = np.empty(c.cells, dtype=np.uint8)
        self._check(self._lib.lbm_aero2d_mask(self._h, abi.ptr(solid, ctypes.c_uint8)))
        return SolidMask2D(c.nx, c.ny, solid)

    def total_steps(self) -> int:
        return self.info().total_steps

    def lattice_handle(self) -> ctypes.c_void_p:
        h = ctypes.c_void_p()
        self._check(self._lib.lbm_aero2d_lattice(self._h, ctypes.byref(h)))
        return h

    def velocity_field(self) -> np.ndarray:
        """SI velocity on the lattice [ny, nx, 3] = (u, 0, w) (zero angle of attack only)."""
        c = self.info().config
        out = np.empty((c.cells, 3))
        self._check(self._lib.lbm_aero2d_velocity(self._h, abi.ptr(out, ctypes.c_double)))
        return out.reshape(c.ny, c.nx, 3)


__all__ = ["FlowConditions", "ReferenceValues", "AeroGeometry", "LbmSolver3DOptions", "LbmAeroSolver3D", "AeroResult",
           "SurfaceField", "ValidityWarning", "LesMode", "LatticePlan", "AeroError", "LbmSolverOptions",
           "LbmAeroSolver2D", "LatticePlan2D"]
