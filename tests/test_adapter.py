"""The C++ drop-in (include/orrery_lbm_solver.hpp): HipLbm3D implements the reference's
Lbm3DSolver interface over the C ABI. Compiled against the reference's own headers where the
reference tree is present; its refusals carry the reference's validate() messages, compared
with the reference's validate compiled from its sources (oracle/_ref)."""
import subprocess

import pytest

import adapter_build
import refcheck
from orrery_amd import Lbm3DConfig, Wall, WallKind, SolidWall, fluid_mask_3d

pytestmark = pytest.mark.skipif(not (adapter_build.possible() and refcheck.available()),
                                reason="needs the reference's headers and oracle/_ref")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    return adapter_build.build(str(tmp_path_factory.mktemp("adapter") / "adapter_check"))


def config(what):
    c = Lbm3DConfig(nx=16, ny=8, nz=8, tau=0.6, inlet_ux=0.05)
    m = fluid_mask_3d(16, 8, 8)
    if what == "tau":
        c.tau = 0.5
    elif what == "mach":
        c.inlet_ux = 0.2
    elif what == "small":
        c.nz = 2
        m = fluid_mask_3d(16, 8, 2)
    elif what == "distance":
        c.solid_wall = SolidWall.interpolated
    elif what == "periodic":
        c.y_low = Wall(WallKind.periodic)
    return c, m


@pytest.mark.parametrize("what", ["tau", "mach", "small", "distance", "periodic"])
def test_refusals_are_the_references(driver, what):
    out = subprocess.run([driver, what], check=True, capture_output=True, text=True).stdout.strip()
    name, message = out.split("\t")
    assert name == what
    assert message == refcheck.validate3(*config(what))


def test_mask_shape_is_checked_before_the_pointer_crosses(driver):
    out = subprocess.run([driver, "size"], check=True, capture_output=True, text=True).stdout.strip()
    assert out == "size\tlbm3d: the solid mask does not match the lattice size"


@pytest.mark.parametrize("what", ["speed", "sides", "mach"])
def test_aero_solver_wrapper_refuses_like_the_adapter(driver, what):
    """HipLbmAeroSolver3D (include/orrery_lbm_aero.hpp) behind the reference's IAeroSolver: its
    setup refusals are the adapter's (the reference's texts)."""
    out = subprocess.run([driver, "aero-" + what], check=True, capture_output=True, text=True).stdout.strip()
    name, message = out.split("\t")
    want = {"speed": "lbm3d: speed, density and viscosity must be positive and finite, angles finite",
            "sides": "lbm3d: the channel sides are free slip or periodic",
            "mach": "lattice units: lattice Mach number 0.103923 (u_lat 0.06) is not below 0.1; lower the lattice "
                    "velocity (weak compressibility, APPROXIMATION A3)"}[what]
    assert name == "aero-" + what and message == want


@pytest.mark.parametrize("what", ["speed", "section", "mach"])
def test_2d_aero_solver_wrapper_refuses_like_the_adapter(driver, what):
    """HipLbmAeroSolver2D (include/orrery_lbm_aero.hpp) behind the reference's IAeroSolver: its
    setup refusals are the 2D adapter's (the reference's texts, lbm_aero_solver.cpp:45-80)."""
    out = subprocess.run([driver, "aero2d-" + what], check=True, capture_output=True, text=True).stdout.strip()
    name, message = out.split("\t")
    want = {"speed": "lbm: speed, density and viscosity must be positive and finite, angles finite",
            "section": "lbm: the 2D solver needs a section of at least 3 points",
            "mach": "lattice units: lattice Mach number 0.138564 (u_lat 0.08) is not below 0.1; lower the lattice "
                    "velocity (weak compressibility, APPROXIMATION A3)"}[what]
    assert name == "aero2d-" + what and message == want


@pytest.mark.parametrize("what", ["tau", "mach", "periodic", "inlet"])
def test_2d_lattice_refusals_are_the_references(driver, what):
    """make_hip_solver for Lbm2DConfig (HipLbm2D, include/orrery_lbm_solver.hpp) refuses with
    the reference's validate() messages (oracle/_ref)."""
    from orrery_amd import Lbm2DConfig, fluid_mask
    from orrery_amd.config import Wall, WallKind
    out = subprocess.run([driver, "lattice2d-" + what], check=True, capture_output=True, text=True).stdout.strip()
    name, message = out.split("\t")
    c = Lbm2DConfig(nx=16, ny=8, tau=0.6, inlet_ux=0.05)
    m = fluid_mask(16, 8)
    if what == "tau":
        c.tau = 0.5
    elif what == "mach":
        c.inlet_ux = 0.2
    elif what == "periodic":
        c.top = Wall(WallKind.periodic)
    elif what == "inlet":
        m.solid[16 * 3] = 1
    assert name == "lattice2d-" + what
    assert message == refcheck.validate2(c, m) and message


def test_2d_lattice_mask_shape_is_checked_before_the_pointer_crosses(driver):
    out = subprocess.run([driver, "lattice2d-size"], check=True, capture_output=True, text=True).stdout.strip()
    assert out == "lattice2d-size\tlbm2d: the solid mask does not match the lattice size"
