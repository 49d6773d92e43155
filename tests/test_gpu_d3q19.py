"""D3Q19 device solver against the oracle, through the C ABI.

The reference pins its GPU path with GPU-vs-CPU comparisons at relative L2 < 1e-4
(tests/aero/lbm3d/test_lbm3d_gpu.cpp:197-324); the same cases run here for the binary32
storage. The binary64 storage runs the CPU solver's arithmetic on the device, so it must agree
with the oracle to rounding (relative 1e-12) on every boundary rule, which pins the in-place
A-A addressing and the boundary stash exactly.
"""
import math

import numpy as np
import pytest

from cases import (MAGIC_TAU, couette3d_config, duct_config, l2_relative, periodic_box, poiseuille3d_config,
                   sphere_channel, uniform_state3d, uniform_state_error, velocity_l2, wall)
from oracle import Oracle3D
from orrery_amd import (Lbm3DSolver, LbmError, Outlet, Smagorinsky, SolidWall, Sponge, WallKind, XBoundary, box_mask,
                        boundary_links, fluid_mask_3d)

pytestmark = pytest.mark.gpu


def compare(config, mask, steps, storage="f32", label="", init=None):
    cpu = Oracle3D(config, mask)
    gpu = Lbm3DSolver(config, mask, storage)
    if init is not None:
        cpu.initialize(init)
        gpu.initialize(init)
    cpu.step(steps)
    gpu.step(steps)
    c = cpu.macroscopic()
    g = gpu.macroscopic()
    out = {
        "velocity": velocity_l2(g, c),
        "populations": l2_relative(gpu.populations(), cpu.populations()),
    }
    fc = cpu.force()
    fg = gpu.force()
    out["force_x"] = abs(fg.force[0] - fc.force[0]) / abs(fc.force[0]) if fc.force[0] != 0 else abs(fg.force[0])
    out["mass"] = abs(gpu.mass() - cpu.mass()) / cpu.mass()
    print(f"measured: GPU {storage} vs oracle {label} after {steps} steps: {out}", flush=True)
    return out, cpu, gpu


def test_analytic_flows_agree():
    # test_lbm3d_gpu.cpp:197-209
    for cfg, mask, label in [
        (poiseuille3d_config(4, 32, 4, 0.8, 1e-6), fluid_mask_3d(4, 32, 4), "Poiseuille"),
        (duct_config(4, 20, 20, 0.8, 1e-6), fluid_mask_3d(4, 20, 20), "duct"),
        (couette3d_config(4, 4, 16, 0.8, 0.02), fluid_mask_3d(4, 4, 16), "Couette"),
    ]:
        out, _, _ = compare(cfg, mask, 4000, label=label)
        assert out["velocity"] < 1e-4


SPHERE = sphere_channel(64, 32, 32, 8.0, 2.5, 0.05, 20.0)


def sphere_variants():
    base_c, mask = SPHERE
    les = base_c.copy()
    les.smagorinsky = Smagorinsky(True, 0.1)
    interp = base_c.copy()
    interp.solid_wall = SolidWall.interpolated
    extrap = base_c.copy()
    extrap.outlet = Outlet.extrapolation
    extrap.y_low = wall(WallKind.periodic)
    extrap.y_high = wall(WallKind.periodic)
    return {"bgk": base_c, "les": les, "interpolated": interp, "extrapolation_periodic_y": extrap}, mask


@pytest.mark.parametrize("variant", ["bgk", "les", "interpolated", "extrapolation_periodic_y"])
def test_laminar_sphere_channel_f32(variant):
    # test_lbm3d_gpu.cpp:211-238: Re 20, D 8, 64x32x32, 800 steps, every metric < 1e-4
    variants, mask = sphere_variants()
    out, _, _ = compare(variants[variant], mask, 800, "f32", variant)
    assert out["velocity"] < 1e-4
    assert out["populations"] < 1e-4
    assert out["force_x"] < 1e-4


@pytest.mark.parametrize("variant", ["bgk", "les", "interpolated", "extrapolation_periodic_y"])
def test_laminar_sphere_channel_f64_matches_the_oracle_to_rounding(variant):
    variants, mask = sphere_variants()
    out, cpu, gpu = compare(variants[variant], mask, 201, "f64", variant)
    assert out["velocity"] < 1e-12
    assert out["populations"] < 1e-13
    assert out["force_x"] < 1e-11
    assert out["mass"] < 1e-13
    fc, fg = cpu.force(), gpu.force()
    for a in range(3):
        assert fg.moment[a] == pytest.approx(fc.moment[a], rel=1e-10, abs=1e-12)


def test_every_wall_kind_f64_matches_the_oracle_to_rounding():
    # Moving no-slip walls on z, slip on y, a sphere with the interpolated wall, LES, a
    # sponge at the inlet and the outlet, and the convective outlet: all rules at once, odd
    # and even step counts.
    c, mask = sphere_channel(40, 18, 16, 6.0, 2.0, 0.04, 30.0)
    c.z_low = wall(WallKind.no_slip, 0.01)
    c.z_high = wall(WallKind.no_slip, -0.02)
    c.smagorinsky = Smagorinsky(True, 0.15)
    c.solid_wall = SolidWall.interpolated
    c.sponge = Sponge(0.1, 4, 6, 3)
    for steps in (1, 2, 77):
        out, cpu, gpu = compare(c, mask, steps, "f64", f"all walls {steps}")
        assert out["populations"] < 1e-13
        assert out["velocity"] < 1e-12
        assert out["force_x"] < 1e-10


def _two_spheres(nx, ny, nz, a, ra, b, rb):
    from orrery_amd.masks import mask_from_sdf_3d

    def sdf(x, y, z):
        da = np.sqrt((x - a[0]) ** 2 + (y - a[1]) ** 2 + (z - a[2]) ** 2) - ra
        db = np.sqrt((x - b[0]) ** 2 + (y - b[1]) ** 2 + (z - b[2]) ** 2) - rb
        return np.minimum(da, db)

    return mask_from_sdf_3d(nx, ny, nz, sdf)


@pytest.mark.parametrize("case", ["periodic_seams", "face_adjacent"])
def test_boundary_sources_and_lattice_stores_f64_match_the_oracle(case):
    """Boundary nodes read other boundary nodes' outputs from the stash and store into the
    lattice only what a lattice reader loads. The readers that decide it meet here: bodies whose
    short interpolated links (q < 1/2, far reads) sit next to the periodic seams, the inlet and
    outlet faces and the moving walls, with stash sources through every periodic wrap; odd and
    even step counts, against the oracle to rounding."""
    if case == "periodic_seams":
        c = periodic_box(20, 14, 12, 0.56)
        c.z_low = wall(WallKind.slip)
        c.z_high = wall(WallKind.no_slip, 0.02)
        c.body_force_x = 2e-5
        mask = _two_spheres(20, 14, 12, (1.3, 0.7, 6.2), 3.1, (17.6, 12.4, 3.3), 2.2)
    else:
        c, _ = sphere_channel(30, 14, 12, 6.0, 2.0, 0.04, 30.0)
        c.z_low = wall(WallKind.no_slip, 0.01)
        c.y_low = wall(WallKind.periodic)
        c.y_high = wall(WallKind.periodic)
        mask = _two_spheres(30, 14, 12, (2.4, 1.2, 5.6), 1.45, (27.2, 6.7, 1.4), 1.7)
    c.smagorinsky = Smagorinsky(True, 0.12)
    c.solid_wall = SolidWall.interpolated
    assert (boundary_links(c, mask).q < 0.5).sum() > 0  # far reads are exercised
    for steps in (1, 2, 61):
        out, cpu, gpu = compare(c, mask, steps, "f64", f"{case} {steps}")
        assert out["populations"] < 1e-13
        assert out["velocity"] < 1e-12
        assert out["force_x"] < 1e-10


@pytest.mark.parametrize("sides", ["slip_slip", "slip_periodic", "periodic_slip"])
def test_side_nodes_next_to_boundary_nodes_f64_match_the_oracle(sides):
    """Side nodes (the y and z face nodes streamed in place through slip mirrors and periodic
    wraps, by the bulk kernel in A steps and the side kernel in B steps) next to listed boundary
    nodes: bodies cut by the faces, so that boundary nodes pull from side nodes through the
    lattice (mirrored or wrapped) and store their outputs where side-node readers load them.
    Odd and even step counts, the FP64 storage against the oracle to rounding."""
    y, z = {"slip_slip": (WallKind.slip, WallKind.slip), "slip_periodic": (WallKind.slip, WallKind.periodic),
            "periodic_slip": (WallKind.periodic, WallKind.slip)}[sides]
    c, _ = sphere_channel(34, 16, 14, 6.0, 2.0, 0.04, 30.0)
    c.y_low, c.y_high, c.z_low, c.z_high = wall(y), wall(y), wall(z), wall(z)
    c.smagorinsky = Smagorinsky(True, 0.12)
    c.solid_wall = SolidWall.interpolated
    c.sponge = Sponge(0.1, 3, 5, 2)
    # one sphere through the y = 0 face, one through the z = nz - 1 face
    mask = _two_spheres(34, 16, 14, (9.3, 0.6, 6.2), 2.6, (22.4, 8.3, 13.2), 2.4)
    for steps in (1, 2, 53):
        out, cpu, gpu = compare(c, mask, steps, "f64", f"sides {sides} {steps}")
        bulk, side, listed = gpu.stream_counts()
        counts = gpu.counts()
        # inlet nodes off the faces stream with the bulk nodes
        assert side > 0 and bulk > counts.bulk_nodes
        assert bulk + side + listed == counts.bulk_nodes + counts.boundary_nodes
        assert out["populations"] < 1e-13
        assert out["velocity"] < 1e-12
        assert out["force_x"] < 1e-10
        # the same populations summed in another order (device block partials, oracle rows):
        # 1.4e5 terms of FP64
        assert out["mass"] < 1e-11


def test_side_nodes_of_a_slip_channel():
    """In a channel with slip sides, every face node away from the inlet and outlet faces and
    from the body is a side node, and every inlet node off the side faces streams in place."""
    c, mask = SPHERE
    gpu = Lbm3DSolver(c, mask, "f32")
    bulk, side, listed = gpu.stream_counts()
    nx, ny, nz = c.nx, c.ny, c.nz
    face_rows = 2 * nz + 2 * (ny - 2)
    assert side == face_rows * (nx - 2)  # the sphere does not reach the faces
    # the inlet nodes off the faces stream like bulk nodes; the outlet face, the inlet nodes on
    # the faces and the fluid neighbours of the sphere are listed
    inlet = (ny - 2) * (nz - 2)
    assert bulk == gpu.counts().bulk_nodes + inlet
    assert listed > ny * nz + 2 * (ny + nz) - 4
    assert bulk + side + listed == c.cells - gpu.counts().solid_nodes
    gpu.close()


def test_perturbed_start_is_identical_on_device_and_host():
    c, mask = sphere_channel(32, 16, 16, 6.0, 2.0, 0.05, 50.0)
    cpu = Oracle3D(c, mask)
    gpu = Lbm3DSolver(c, mask, "f64")
    cpu.initialize_perturbed(1234, 1e-4)
    gpu.initialize_perturbed(1234, 1e-4)
    assert np.array_equal(cpu.populations(), gpu.populations())
    cpu.step(10)
    gpu.step(10)
    assert l2_relative(gpu.populations(), cpu.populations()) < 1e-14


def test_unusable_wall_distances_agree():
    # test_lbm3d_gpu.cpp:240-268
    c, mask = sphere_channel(48, 24, 24, 8.0, 2.0, 0.05, 20.0)
    c.solid_wall = SolidWall.interpolated
    centre = 2.0 * 8.0
    x = np.tile(np.arange(48, dtype=np.float64), 24 * 24)
    front = x < centre
    d = mask.distance.copy()
    d[front & (mask.solid == 0)] = math.inf
    d[~front & (mask.solid != 0)] = 1.0
    mask.distance = d
    out, _, _ = compare(c, mask, 400, "f32", "unusable distances")
    assert out["velocity"] < 1e-4
    assert out["populations"] < 1e-4
    assert out["force_x"] < 1e-4


@pytest.mark.parametrize("rho", [0.7, 2.0])
def test_co_moving_walls_f32(rho):
    # test_lbm3d_gpu.cpp:270-289
    c = periodic_box(4, 8, 8, 0.7)
    u = 0.03
    for side in ("y_low", "y_high", "z_low", "z_high"):
        setattr(c, side, wall(WallKind.no_slip, u))
    g = Lbm3DSolver(c, fluid_mask_3d(4, 8, 8))
    g.initialize(uniform_state3d(c, rho, u))
    g.step(300)
    assert uniform_state_error(g.macroscopic(), rho, (u, 0.0, 0.0), u) < 1e-4


def test_call_size_independence_is_bitwise():
    c, mask = sphere_channel(32, 16, 16, 6.0, 2.0, 0.05, 50.0)
    a = Lbm3DSolver(c, mask)
    b = Lbm3DSolver(c, mask)
    a.step(150)
    b.step(100)
    b.step(50)
    assert np.array_equal(a.populations(), b.populations())
    assert a.force() == b.force()


@pytest.mark.parametrize("storage,x_periodic", [("f32", False), ("f16s", False), ("f64", False), ("f32", True),
                                                ("f16s", True)])
def test_schedules_and_launch_shapes_are_bitwise_identical(storage, x_periodic):
    """One fused kernel per step, the two kernels in turn (under every bulk launch shape and
    occupancy cap), and the two kernels on two streams give the same state bit for bit. 70 x 21 x 19 cells is not a
    multiple of 64, so the last wave of the fused kernel is partial; with 4 cells per thread
    some runs hold the x = nx - 1 node of one row and the x = 0 node of the next. With periodic
    x, the x-face nodes pull across the row ends. The populations are stored in blocks of 2048
    cells (the plane holds 1470), whose edges fall at varying x: the B-step runs of 2 and 4 cells
    that an x shift carries across a block edge, moved cell by cell, meet the fused kernel's
    per-cell addresses."""
    c, mask = sphere_channel(70, 21, 19, 7.0, 2.0, 0.05, 60.0)
    c.smagorinsky = Smagorinsky(True, 0.17)
    c.solid_wall = SolidWall.interpolated
    if x_periodic:
        c.x_boundary = XBoundary.periodic
        c.body_force_x = 1e-6
    runs = []
    for schedule, block, cpt, caps in [("fused", 0, 0, (0, 0)), ("serial", 256, 1, (0, 0)), ("serial", 128, 2, (0, 0)),
                                       ("serial", 64, 4, (0, 0)), ("streams", 192, 1, (0, 0)),
                                       ("serial", 256, 1, (4, 3)), ("streams", 256, 2, (2, 5))]:
        s = Lbm3DSolver(c, mask, storage)
        s.set_schedule(schedule)
        s.set_launch(block, cpt)
        s.set_occupancy(*caps)
        assert s.occupancy() == caps
        s.initialize_perturbed(5, 1e-4)
        s.step(41)
        runs.append((schedule, block, cpt, caps, s.populations(), s.force()))
    for schedule, block, cpt, caps, f, force in runs[1:]:
        print(f"{storage}: {schedule} block {block} cpt {cpt} workgroups per CU {caps}", flush=True)
        assert np.array_equal(f, runs[0][4]) and force == runs[0][5]


def test_ramped_inlet_agrees():
    c, mask = sphere_channel(32, 16, 16, 6.0, 2.0, 0.05, 50.0)
    cpu = Oracle3D(c, mask)
    gpu = Lbm3DSolver(c, mask, "f64")
    for n in range(5):
        v = (0.05 * (1 - math.cos(math.pi * n / 5)) / 2, 0.0, 0.0)
        cpu.set_inlet_velocity(v)
        gpu.set_inlet_velocity(v)
        cpu.step(7)
        gpu.step(7)
    assert l2_relative(gpu.populations(), cpu.populations()) < 1e-13


def test_momentum_balance_on_device():
    # test_lbm3d_flows.cpp:170-188 on the device (binary32): force = body force x fluid cells
    c = periodic_box(4, 16, 4, 0.9)
    c.body_force_x = 2e-6
    g = Lbm3DSolver(c, box_mask(4, 16, 4, (0.0, 0.0, 0.0), (3.0, 0.0, 3.0)))
    g.step(6000)
    f = g.force()
    assert f.force[0] == pytest.approx(2e-6 * 240, rel=2e-3)


def shear_wave(c, amplitude):
    """rho 1, u_x = amplitude sin(2 pi y / ny), u_y = u_z = 0 on every cell."""
    planes = np.zeros((4, c.cells))
    y = (np.arange(c.cells) // c.nx) % c.ny
    planes[0] = 1.0
    planes[1] = amplitude * np.sin(2.0 * math.pi * y / c.ny)
    return planes


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_fully_periodic_box_decays_like_the_shear_wave(storage):
    # Every side periodic, no solid: the nodes on the seams (whose pulls wrap) are boundary nodes,
    # the interior bulk nodes. The shear wave decays as exp(-nu k^2 t), nu = (tau - 1/2) / 3,
    # k = 2 pi / ny, up to the lattice's O(k^2) dispersion; binary64 agrees with the oracle to
    # rounding.
    c = periodic_box(8, 32, 6, 0.8)
    amplitude, steps = 1e-3, 400
    planes = shear_wave(c, amplitude)
    out, cpu, gpu = compare(c, fluid_mask_3d(8, 32, 6), steps, storage, "shear wave", planes)
    counts = gpu.counts()
    interior = (c.nx - 2) * (c.ny - 2) * (c.nz - 2)
    assert counts.bulk_nodes == interior and counts.boundary_nodes == c.cells - interior
    assert counts.solid_nodes == 0 and counts.links == 0
    k = 2.0 * math.pi / c.ny
    want = amplitude * math.exp(-(c.tau - 0.5) / 3.0 * k * k * steps)
    m = gpu.macroscopic()
    y = (np.arange(c.cells) // c.nx) % c.ny
    got = np.asarray(m.velocity_x)[y == 8].mean()  # sin = 1 at y = ny / 4
    print(f"measured: shear wave {storage}: amplitude {got:.6e}, exp(-nu k^2 t) {want:.6e}", flush=True)
    assert got == pytest.approx(want, rel=1e-2)
    if storage == "f64":
        assert out["velocity"] < 1e-12 and out["populations"] < 1e-12
    else:
        assert out["velocity"] < 1e-4  # the reference's GPU-vs-CPU bar (relative L2)


def test_mass_conservation_f64():
    from test_oracle_flows import mass_case
    c, mask, planes = mass_case()
    g = Lbm3DSolver(c, mask, "f64")
    g.initialize(planes)
    before = g.mass()
    g.step(1500)
    assert abs(g.mass() - before) / before < 1e-12


def test_refusals_happen_before_allocation():
    c = periodic_box(8, 8, 8, 0.5 + 1e-9)
    with pytest.raises(LbmError, match="rounds to 1/2 in binary32"):
        Lbm3DSolver(c, fluid_mask_3d(8, 8, 8))
    Lbm3DSolver(c, fluid_mask_3d(8, 8, 8), "f64")  # binary64 keeps the viscosity
    from orrery_amd import set_memory_limit
    set_memory_limit(1024)
    try:
        with pytest.raises(LbmError, match="are available; use fewer cells"):
            Lbm3DSolver(periodic_box(16, 16, 16, 0.6), fluid_mask_3d(16, 16, 16))
    finally:
        set_memory_limit(0)


# The velocity field is held tightly (measured f16s 4.2e-5, bf16s 3.3e-4). A force sample reads
# the link populations of one step, each carrying its stochastic rounding: the force error of
# this case moves with random sign between samples, up to 8e-4 (f16s) and 1e-2 (bf16s) over
# steps 100-6400, so its bar is that noise level.
@pytest.mark.parametrize("storage,velocity_limit,force_limit", [("f16s", 2e-4, 5e-3), ("bf16s", 2e-3, 5e-2)])
def test_sixteen_bit_storage_error_against_fp64(storage, velocity_limit, force_limit):
    variants, mask = sphere_variants()
    c = variants["bgk"]
    ref = Lbm3DSolver(c, mask, "f64")
    low = Lbm3DSolver(c, mask, storage)
    ref.step(800)
    low.step(800)
    v = velocity_l2(low.macroscopic(), ref.macroscopic())
    f = abs(low.force().force[0] - ref.force().force[0]) / abs(ref.force().force[0])
    print(f"measured: {storage} vs f64 sphere Re 20 after 800 steps: velocity L2 {v:.3e}, force_x {f:.3e}")
    assert v < velocity_limit
    assert f < force_limit
