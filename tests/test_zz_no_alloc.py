"""lbm3d_sample_fixed, the allocation-free variant of lbm3d_sample: called through the C ABI
on two lattices in a row (the second call with fewer nodes, so stale entries of the shared
device array would show), its outputs equal lbm3d_sample's bit for bit; sizes beyond its fixed
array are refused."""
import ctypes

import numpy as np
import pytest

from cases import sphere_channel
from orrery_amd import Lbm3DSolver, SolidWall, abi

pytestmark = pytest.mark.gpu

MAX_NODES = 1 << 16


def sample(lib, fn, h, cells):
    n = len(cells)
    out = np.full((4, n), np.nan)
    c = np.ascontiguousarray(cells, dtype=np.uint32)
    rc = fn(h, abi.ptr(c, ctypes.c_uint32), n, *[abi.ptr(out[k], ctypes.c_double) for k in range(4)])
    assert rc == 0, lib.lbm_last_error()
    return out


@pytest.mark.parametrize("storage", ["f32", "f16s", "f64"])
def test_fixed_sample_equals_the_allocating_one(storage):
    lib = abi.load()
    big, mbig = sphere_channel(96, 40, 36, 10.0, 2.0, 0.05, 60.0)
    big.solid_wall = SolidWall.interpolated
    small, msmall = sphere_channel(40, 20, 18, 6.0, 2.0, 0.05, 30.0)
    rng = np.random.default_rng(9)
    lattices = []
    for config, mask in ((big, mbig), (small, msmall)):
        s = Lbm3DSolver(config, mask, storage)
        s.initialize_perturbed(3, 1e-4)
        s.step(37)  # an odd count: the post-A-step parity
        lattices.append((config, s))
    # first call: 60000 nodes of the large lattice (bulk, boundary and solid nodes, repeats);
    # second call: 5000 of the small one
    for (config, s), n in zip(lattices, (60000, 5000)):
        cells = rng.integers(0, config.cells, n)
        cells[:64] = np.arange(64)  # the inlet row: boundary nodes
        want = sample(lib, lib.lbm3d_sample, s.handle, cells)
        got = sample(lib, lib.lbm3d_sample_fixed, s.handle, cells)
        assert np.array_equal(got, want, equal_nan=False)
        m = s.macroscopic()
        assert np.array_equal(got[1], m.velocity_x[cells])
    config, s = lattices[0]
    # at the limit and beyond it
    cells = rng.integers(0, config.cells, MAX_NODES)
    assert np.array_equal(sample(lib, lib.lbm3d_sample_fixed, s.handle, cells),
                          sample(lib, lib.lbm3d_sample, s.handle, cells))
    too_many = np.zeros(MAX_NODES + 1, dtype=np.uint32)
    out = np.zeros(MAX_NODES + 1)
    rc = lib.lbm3d_sample_fixed(s.handle, abi.ptr(too_many, ctypes.c_uint32), MAX_NODES + 1,
                                abi.ptr(out, ctypes.c_double), None, None, None)
    assert rc != 0 and b"at most 65536 nodes" in lib.lbm_last_error()
    bad = np.array([config.cells], dtype=np.uint32)
    rc = lib.lbm3d_sample_fixed(s.handle, abi.ptr(bad, ctypes.c_uint32), 1, abi.ptr(out, ctypes.c_double),
                                None, None, None)
    assert rc != 0 and b"outside the lattice" in lib.lbm_last_error()
    for _, s in lattices:
        s.close()
