"""Batched lattices on the device: every lattice of a batch evolves exactly as the same lattice
alone (bitwise, in every storage format), its force samples equal the single-lattice force, and
admission refuses lattices beyond the HBM budget before allocating."""
import numpy as np
import pytest

from cases import sphere_channel
from orrery_amd import Lbm3DSolver, LbmError, Smagorinsky, SolidWall
from orrery_amd.batch import LatticeBatch

pytestmark = pytest.mark.gpu


def lattices():
    a, ma = sphere_channel(48, 24, 24, 8.0, 2.0, 0.05, 40.0)
    b, mb = sphere_channel(64, 20, 28, 6.0, 3.0, 0.04, 80.0)
    b.smagorinsky = Smagorinsky(True, 0.17)
    c, mc = sphere_channel(40, 30, 18, 7.0, 2.0, 0.05, 30.0)
    c.solid_wall = SolidWall.interpolated
    return [(a, ma, "f32", 11), (b, mb, "f32", 12), (c, mc, "f16s", 13), (a, ma, "bf16s", 14)]


def alone(config, mask, storage, seed, steps):
    s = Lbm3DSolver(config, mask, storage)
    s.initialize_perturbed(seed, 1e-4)
    forces = []
    for _ in range(steps // 10):
        s.step(10)
        forces.append(s.force())
    return s, forces


def test_batch_matches_single_lattices_bitwise():
    steps = 60
    batch = LatticeBatch(sample_every=10)
    members = [(batch.add(c, m, st, seed, 1e-4), c, m, st, seed) for c, m, st, seed in lattices()]
    batch.run(30)
    batch.run(30)
    for lat, c, m, st, seed in members:
        print(f"lattice {lat.id} {c.nx}x{c.ny}x{c.nz} {st}", flush=True)
        single, forces = alone(c, m, st, seed, steps)
        assert lat.steps() == steps
        assert np.array_equal(lat.populations(), single.populations())
        t, f = lat.history()
        assert t.tolist() == [10.0 * (k + 1) for k in range(steps // 10)]
        for k, fs in enumerate(forces):
            assert tuple(f[k, :3]) == fs.force and tuple(f[k, 3:]) == fs.moment


@pytest.mark.parametrize("groups", [3, 5])
def test_occupancy_caps_leave_the_batch_bitwise_unchanged(groups):
    steps = 40
    batch = LatticeBatch(sample_every=10)
    members = [(batch.add(c, m, st, seed, 1e-4), c, m, st, seed) for c, m, st, seed in lattices()]
    for st in ("f32", "f16s", "bf16s"):
        batch.set_occupancy(st, groups)
    batch.run(steps)
    for lat, c, m, st, seed in members:
        single, _ = alone(c, m, st, seed, steps)
        assert np.array_equal(lat.populations(), single.populations())


def test_occupancy_caps_are_checked():
    batch = LatticeBatch(sample_every=10)
    for bad in (-1, 33):
        with pytest.raises(LbmError, match="workgroups per compute unit"):
            batch.set_occupancy("f32", bad)


def test_a_lattice_is_the_same_alone_and_in_the_mix():
    c, m, st, seed = lattices()[1]
    solo = LatticeBatch(sample_every=10)
    x = solo.add(c, m, st, seed, 1e-4)
    solo.run(40)
    mix = LatticeBatch(sample_every=10)
    others = [mix.add(cc, mm, ss, sd, 1e-4) for cc, mm, ss, sd in lattices()]
    mix.run(20)
    y = mix.add(c, m, st, seed, 1e-4)  # admitted while the others are running
    mix.remove(others[0])
    mix.run(40)
    assert np.array_equal(x.populations(), y.populations())
    tx, fx = x.history()
    ty, fy = y.history()
    assert np.array_equal(tx, ty) and np.array_equal(fx, fy)


def test_admission_refuses_beyond_the_budget():
    c, m, st, seed = lattices()[0]
    need = LatticeBatch.required_bytes(c, m, "f32")
    batch = LatticeBatch(budget_bytes=2 * need + need // 2)
    batch.add(c, m, "f32")
    batch.add(c, m, "f32")
    with pytest.raises(LbmError, match="of the budget are free"):
        batch.add(c, m, "f32")
    used, budget = batch.bytes()
    assert used == 2 * need


def test_a_batch_lattice_refuses_single_lattice_stepping():
    """The batch keeps each lattice's A-A parity on the device; stepping, re-initialising,
    retargeting or freeing its handle directly is refused (no device work is done), and the
    lattice still matches its single twin afterwards."""
    import ctypes
    c, m, st, seed = lattices()[0]
    batch = LatticeBatch(sample_every=10)
    lat = batch.add(c, m, st, seed, 1e-4)
    batch.run(20)
    lib = batch._lib
    h = lat._h()
    v = (ctypes.c_double * 3)(0.05, 0.0, 0.0)
    for name, call in (("step", lambda: lib.lbm3d_step(h, 2)), ("destroy", lambda: lib.lbm3d_destroy(h)),
                       ("set_inlet", lambda: lib.lbm3d_set_inlet(h, v)),
                       ("initialize", lambda: lib.lbm3d_initialize_perturbed(h, 1, 1e-4))):
        assert call() != 0, name
        assert b"belongs to a batch" in lib.lbm_last_error(), name
    assert lat.steps() == 20
    batch.run(20)
    single, _ = alone(c, m, st, seed, 40)
    assert np.array_equal(lat.populations(), single.populations())


def test_timed_run_advances_exactly_as_a_run():
    """The per-kernel timed run (eager launches between events) gives the same lattices and
    force histories bit for bit as graph replays, and reports every storage group."""
    steps = 40
    plain = LatticeBatch(sample_every=10)
    timed = LatticeBatch(sample_every=10)
    pairs = []
    for c, m, st, seed in lattices():
        pairs.append((plain.add(c, m, st, seed, 1e-4), timed.add(c, m, st, seed, 1e-4)))
    plain.run(steps)
    report = timed.run_timed(20)
    report2 = timed.run_timed(20)
    assert set(report["groups"]) == {"f32", "f16s", "bf16s"}
    for r in (report, report2):
        assert r["step_ms"] > 0 and r["force_ms"] > 0
        for g in r["groups"].values():
            assert g["bulk_ms"] > 0 and g["boundary_ms"] > 0 and g["bulk_nodes"] > 0
    for a, b in pairs:
        assert b.steps() == steps
        assert np.array_equal(a.populations(), b.populations())
        ta, fa = a.history()
        tb, fb = b.history()
        assert np.array_equal(ta, tb) and np.array_equal(fa, fb)
