"""The C++ drop-in on the device: the driver tests/adapter_check.cpp steps lattices through the
reference's own interfaces (Lbm3DSolver, Lbm2DSolver, IAeroSolver) implemented by HipLbm3D,
HipLbm2D and HipLbmAeroSolver3D, and its results equal, bit for bit, the same runs made through
the C ABI from Python. The driver is compiled against the reference's headers by the package
build (tests/adapter_build.py); it reads no reference file when it runs."""
import os
import subprocess

import numpy as np
import pytest

import adapter_build
from orrery_amd import (Lbm2DConfig, Lbm2DSolver, Lbm3DConfig, Lbm3DSolver, Smagorinsky, SolidMask2D,
                        SolidMask3D)
from orrery_amd.aero import AeroGeometry, FlowConditions, LbmAeroSolver3D, LbmSolver3DOptions, ReferenceValues
from orrery_amd.geometry import TriangleMesh

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(adapter_build.DRIVER),
                                 reason="the C++ driver is built only where the reference's headers are present")]


def drive(*args):
    print(f"case: adapter_check {' '.join(args)}", flush=True)
    out = subprocess.run([adapter_build.DRIVER, *args], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    fields = out.stdout.strip().split("\t")
    assert fields[0] == args[0], out.stdout
    return fields[1:]


def pattern(k, period):
    return (np.asarray(k) % period).astype(np.float64) / float(period)


def test_lattice_through_the_reference_interface_equals_the_c_abi(tmp_path):
    path = str(tmp_path / "run3d.bin")
    fields = drive("run3d", path)
    c = Lbm3DConfig(nx=40, ny=20, nz=18, tau=0.6, inlet_ux=0.05, smagorinsky=Smagorinsky(True, 0.17))
    z, y, x = np.meshgrid(np.arange(c.nz), np.arange(c.ny), np.arange(c.nx), indexing="ij")
    solid = ((x - 12) ** 2 + (y - 10) ** 2 + (z - 9) ** 2 <= 16).astype(np.uint8).ravel()
    mask = SolidMask3D(c.nx, c.ny, c.nz, solid)
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    planes = np.stack([1.0 + 1e-3 * pattern(7 * x + 3 * y + 5 * z, 11), 0.05 + 1e-3 * pattern(x + 2 * y, 5),
                       1e-3 * pattern(3 * x + z, 7), -5e-4 * pattern(y + z, 3)])
    s = Lbm3DSolver(c, mask, "f32")
    s.initialize(planes)
    s.step(7)
    s.set_inlet_velocity((0.045, 0.002, -0.001))
    s.step(33)
    f = s.force()
    want = [s.steps(), *f.force, *f.moment, s.mass(), *s.inlet_velocity(), len(s.links().q)]
    got = [int(fields[0])] + [float(v) for v in fields[1:11]] + [int(fields[11])]
    print(f"measured: drop-in Fx {got[1]!r}, C ABI Fx {want[1]!r}", flush=True)
    assert got == want
    data = np.fromfile(path, dtype=np.float64)
    cells = c.cells
    pops = s.populations()
    m = s.macroscopic()
    assert data.size == 19 * cells + 4 * cells
    assert np.array_equal(data[:19 * cells], pops)
    rest = data[19 * cells:].reshape(4, cells)
    for k, a in enumerate(("density", "velocity_x", "velocity_y", "velocity_z")):
        assert np.array_equal(rest[k], getattr(m, a)), a
    s.close()


def test_2d_lattice_through_the_reference_interface_equals_the_c_abi(tmp_path):
    path = str(tmp_path / "run2d.bin")
    fields = drive("run2d", path)
    c = Lbm2DConfig(nx=48, ny=24, tau=0.58, inlet_ux=0.04)
    solid = np.zeros(c.nx * c.ny, dtype=np.uint8)
    for yy in range(9, 14):
        solid[10 + c.nx * yy:15 + c.nx * yy] = 1
    y, x = np.meshgrid(np.arange(c.ny), np.arange(c.nx), indexing="ij")
    x, y = x.ravel(), y.ravel()
    planes = np.stack([1.0 + 1e-3 * pattern(5 * x + 3 * y, 13), 0.04 + 1e-3 * pattern(x + y, 7),
                       1e-3 * pattern(2 * x + y, 5)])
    s = Lbm2DSolver(c, SolidMask2D(c.nx, c.ny, solid), "f32")
    s.initialize(planes)
    s.step(9)
    s.step(31)
    f = s.force()
    want = [s.steps(), f.x, f.y, f.moment, s.mass()]
    got = [int(fields[0])] + [float(v) for v in fields[1:5]]
    assert got == want
    data = np.fromfile(path, dtype=np.float64)
    cells = c.nx * c.ny
    assert data.size == 9 * cells + 3 * cells
    assert np.array_equal(data[:9 * cells], s.populations())
    m = s.macroscopic()
    rest = data[9 * cells:].reshape(3, cells)
    for k, a in enumerate(("density", "velocity_x", "velocity_y")):
        assert np.array_equal(rest[k], getattr(m, a)), a
    s.close()


def test_aero_solver_through_the_reference_interface_equals_the_c_abi():
    fields = drive("solve3d")
    h = 0.01
    vertices = [(-h, -h, -h), (h, -h, -h), (-h, h, -h), (h, h, -h), (-h, -h, h), (h, -h, h), (-h, h, h), (h, h, h)]
    triangles = [(0, 4, 6), (0, 6, 2), (1, 3, 7), (1, 7, 5), (0, 1, 5), (0, 5, 4),
                 (2, 6, 7), (2, 7, 3), (0, 2, 3), (0, 3, 1), (4, 5, 7), (4, 7, 6)]
    geo = AeroGeometry(TriangleMesh(np.array(vertices), np.array(triangles)), ReferenceValues(area=0.0004, length=0.02))
    cond = FlowConditions(speed=0.1, density=1000.0, viscosity=0.1)
    o = LbmSolver3DOptions(cells_per_length=8.0, lattice_velocity=0.05, upstream=2.0, downstream=3.0, width=3.0,
                           duration=2.0, average_window=1.0, steps_per_call=40)
    s = LbmAeroSolver3D(o)
    s.setup(geo, cond)
    calls = 1
    while not s.step():
        calls += 1
    r = s.result()
    want = [calls, r.iterations, r.cd, r.cl, r.cy, r.cm, r.force[0], len(r.surface.cp)]
    got = [int(fields[0]), int(fields[1])] + [float(v) for v in fields[2:7]] + [int(fields[7])]
    print(f"measured: drop-in Cd {got[2]!r}, C ABI Cd {want[2]!r} after {got[1]} steps", flush=True)
    assert got == want
    s.close()
