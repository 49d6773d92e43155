"""Results do not depend on the population layout.

The populations are stored in blocks of 2^k cells with their 19 directions together (k from
the plane size, at most 2^13 cells for 4- and 8-byte formats and 2^14 for 16-bit ones);
ORRERY_LBM_BLOCK_SHIFT, read at each lattice's creation, forces k. Small blocks make many waves
cross a block edge under the B step's shifts (those waves form their offsets lane by lane, and
runs of 2 or 4 cells that straddle an edge are moved cell by cell); large blocks make most waves
take the uniform offsets. The same lattice stepped under every block size must give the same
populations, force and moment, bit for bit, in every storage format and run width.
"""
import numpy as np
import pytest

from cases import sphere_channel, wall

pytestmark = pytest.mark.gpu

SHIFTS = (8, 9, 12, 16)
FORMATS = [("f32", 1), ("f32", 2), ("f32", 4), ("f16s", 2), ("f16s", 1), ("bf16s", 4), ("f64", 1)]


def lattices():
    from orrery_amd import Smagorinsky, SolidWall, WallKind, XBoundary, box_mask
    from orrery_amd.config import Lbm3DConfig
    c, m = sphere_channel(64, 32, 30, 8.0, 2.0, 0.05, 200.0)
    c.smagorinsky = Smagorinsky(True, 0.17)
    c.solid_wall = SolidWall.interpolated
    yield "sphere 64x32x30, slip sides, interpolated wall", c, m
    # a plane of 40 x 24 cells: the y shifts are not multiples of a wave's cells either
    c2 = Lbm3DConfig(nx=40, ny=24, nz=20, tau=0.7, x_boundary=XBoundary.periodic,
                     y_low=wall(WallKind.periodic), y_high=wall(WallKind.periodic),
                     z_low=wall(WallKind.slip), z_high=wall(WallKind.slip), body_force_x=1e-5)
    yield "periodic 40x24x20 box on a slip floor", c2, box_mask(40, 24, 20, (10.0, 8.0, -1.0), (15.0, 14.0, 5.0))


def run(config, mask, storage, cpt, steps=7):
    from orrery_amd import Lbm3DSolver
    s = Lbm3DSolver(config, mask, storage)
    try:
        s.set_launch(256, cpt)
        s.initialize_perturbed(11, 1e-3)
        s.step(steps)
        f = s.force()
        return s.populations(), np.array(list(f.force) + list(f.moment)), s.device_bytes()
    finally:
        s.close()


@pytest.mark.parametrize("storage,cpt", FORMATS)
def test_every_block_size_gives_the_same_bits(storage, cpt, monkeypatch):
    for name, config, mask in lattices():
        print(f"case: {name}, {storage}, {cpt} cells per thread", flush=True)
        monkeypatch.delenv("ORRERY_LBM_BLOCK_SHIFT", raising=False)
        pops, force, nbytes = run(config, mask, storage, cpt)
        sizes = {nbytes}
        for shift in SHIFTS:
            monkeypatch.setenv("ORRERY_LBM_BLOCK_SHIFT", str(shift))
            p, f, b = run(config, mask, storage, cpt)
            sizes.add(b)
            assert np.array_equal(p, pops), (name, storage, cpt, shift)
            assert np.array_equal(f, force), (name, storage, cpt, shift)
        assert len(sizes) > 2  # the forced block sizes took effect (they change the guard cells)
