"""Problem description of the LBM solvers: the Python mirror of the reference's config types.

Names and defaults follow src/aero/lbm/include/orrery/aero/lbm/lbm2d.h:43-78 (XBoundary, Outlet,
WallKind, Wall, Smagorinsky, Lbm2DConfig) and src/aero/lbm3d/include/orrery/aero/lbm3d/
lbm3d.h:61-103 (Sponge, SolidWall, Lbm3DConfig).
"""
from __future__ import annotations

import dataclasses
import enum

from . import abi


class XBoundary(enum.IntEnum):
    periodic = 0
    inflow_outflow = 1


class Outlet(enum.IntEnum):
    convective = 0
    extrapolation = 1


class WallKind(enum.IntEnum):
    periodic = 0
    slip = 1
    no_slip = 2


class SolidWall(enum.IntEnum):
    halfway = 0
    interpolated = 1


@dataclasses.dataclass
class Wall:
    kind: WallKind = WallKind.slip
    velocity_x: float = 0.0


@dataclasses.dataclass
class Smagorinsky:
    enabled: bool = False
    constant: float = 0.1


@dataclasses.dataclass
class Sponge:
    strength: float = 0.0
    inlet: int = 0
    outlet: int = 0
    sides: int = 0


def _wall(w: Wall) -> abi.CWall:
    return abi.CWall(int(w.kind), float(w.velocity_x))


@dataclasses.dataclass
class Lbm3DConfig:
    nx: int = 0
    ny: int = 0
    nz: int = 0
    tau: float = 0.6
    x_boundary: XBoundary = XBoundary.inflow_outflow
    outlet: Outlet = Outlet.convective
    y_low: Wall = dataclasses.field(default_factory=Wall)
    y_high: Wall = dataclasses.field(default_factory=Wall)
    z_low: Wall = dataclasses.field(default_factory=Wall)
    z_high: Wall = dataclasses.field(default_factory=Wall)
    inlet_ux: float = 0.0
    inlet_uy: float = 0.0
    inlet_uz: float = 0.0
    body_force_x: float = 0.0
    body_force_y: float = 0.0
    body_force_z: float = 0.0
    smagorinsky: Smagorinsky = dataclasses.field(default_factory=Smagorinsky)
    sponge: Sponge = dataclasses.field(default_factory=Sponge)
    solid_wall: SolidWall = SolidWall.halfway

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def to_c(self) -> abi.CConfig3D:
        return abi.CConfig3D(
            self.nx, self.ny, self.nz, float(self.tau), int(self.x_boundary), int(self.outlet),
            _wall(self.y_low), _wall(self.y_high), _wall(self.z_low), _wall(self.z_high),
            float(self.inlet_ux), float(self.inlet_uy), float(self.inlet_uz),
            float(self.body_force_x), float(self.body_force_y), float(self.body_force_z),
            int(bool(self.smagorinsky.enabled)), float(self.smagorinsky.constant),
            float(self.sponge.strength), int(self.sponge.inlet), int(self.sponge.outlet),
            int(self.sponge.sides), int(self.solid_wall))

    def copy(self) -> "Lbm3DConfig":
        return dataclasses.replace(
            self, y_low=dataclasses.replace(self.y_low), y_high=dataclasses.replace(self.y_high),
            z_low=dataclasses.replace(self.z_low), z_high=dataclasses.replace(self.z_high),
            smagorinsky=dataclasses.replace(self.smagorinsky),
            sponge=dataclasses.replace(self.sponge))


@dataclasses.dataclass
class Lbm2DConfig:
    nx: int = 0
    ny: int = 0
    tau: float = 0.6
    x_boundary: XBoundary = XBoundary.inflow_outflow
    outlet: Outlet = Outlet.convective
    bottom: Wall = dataclasses.field(default_factory=Wall)
    top: Wall = dataclasses.field(default_factory=Wall)
    inlet_ux: float = 0.0
    inlet_uy: float = 0.0
    body_force_x: float = 0.0
    body_force_y: float = 0.0
    smagorinsky: Smagorinsky = dataclasses.field(default_factory=Smagorinsky)

    @property
    def cells(self) -> int:
        return self.nx * self.ny

    def to_c(self) -> abi.CConfig2D:
        return abi.CConfig2D(
            self.nx, self.ny, float(self.tau), int(self.x_boundary), int(self.outlet),
            _wall(self.bottom), _wall(self.top), float(self.inlet_ux), float(self.inlet_uy),
            float(self.body_force_x), float(self.body_force_y),
            int(bool(self.smagorinsky.enabled)), float(self.smagorinsky.constant))


def viscosity_of_tau(tau: float) -> float:
    """(tau - 1/2) / 3 (lbm2d.h:89-91)."""
    return (tau - 0.5) * (1.0 / 3.0)


def tau_of_viscosity(nu: float) -> float:
    """nu / c_s^2 + 1/2 (lbm2d.h:92-94)."""
    return nu / (1.0 / 3.0) + 0.5
