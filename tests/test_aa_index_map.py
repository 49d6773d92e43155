"""CPU emulation of the in-place (A-A) index map of the D3Q19 bulk kernel, with labels in place of
populations.

The bulk kernel (orrery_amd/csrc/lbm3d_step.h, bulk_run) addresses element e of plane k of the
single population array as below, with n the number of completed steps:
- n odd, A step: node x loads plane i at x and stores its output i to plane opp(i) at x;
- n even, B step: node x loads plane opp(i) at x - c_i (through a base moved back by the guard
  g = nx ny + nx + 1) and stores output i to plane i at x + c_i.
The readback kernels report f*_i(x) from plane opp(i) at x after an even count and from plane
i at x + c_i after an odd one.

The emulation writes a label (node, direction, step) for every output and checks, over several
steps of a lattice whose interior nodes are bulk nodes: every node reads exactly the addresses
it writes (the in-place property), no address is written twice in a step, a bulk node's pull i
is the label its upstream bulk node x - c_i wrote one step before (the streaming), and the
readback map finds every label. A planted mistake in the map (the B step's store without its
shift) breaks the streaming check.
"""
import numpy as np
import pytest

CX = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
CY = [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1]
CZ = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1]
OPP = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17]


class Lattice:
    def __init__(self, nx, ny, nz, shift_b_store=True):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.cells = nx * ny * nz
        self.guard = nx * ny + nx + 1
        self.shift_b_store = shift_b_store
        # plane k, element e (e may run into the guards on either side)
        self.mem = {}

    def off(self, i):
        return CX[i] + self.nx * CY[i] + self.nx * self.ny * CZ[i]

    def addresses(self, x, n):
        """(load, store) addresses (plane, element) of node x in the step after n completed ones."""
        loads, stores = [], []
        for i in range(19):
            if n % 2 == 1:  # A step
                loads.append((i, x))
                stores.append((OPP[i], x))
            else:  # B step: load through the base moved back by the guard
                g = self.guard
                loads.append((OPP[i], (x + g - self.off(i)) - g))
                stores.append((i, x + self.off(i) if self.shift_b_store else x))
        return loads, stores

    def readback(self, x, i, n):
        return (OPP[i], x) if n % 2 == 0 else (i, x + self.off(i))


def interior(lat):
    return [x + lat.nx * (y + lat.ny * z) for z in range(1, lat.nz - 1) for y in range(1, lat.ny - 1)
            for x in range(1, lat.nx - 1)]


def run(lat, steps):
    nodes = interior(lat)
    is_bulk = set(nodes)
    # initial state in the post-A layout (n = 0): f*_i(x) = (x, i, 0) at plane opp(i), element x
    for x in range(lat.cells):
        for i in range(19):
            lat.mem[lat.readback(x, i, 0)] = (x, i, 0)
    failures = []
    for n in range(steps):
        written = {}
        pulled = {}
        for x in nodes:
            loads, stores = lat.addresses(x, n)
            if sorted(loads) != sorted(stores):
                failures.append(("not in place", x, n))
            for i, a in enumerate(loads):
                pulled[(x, i)] = lat.mem.get(a)
            for i, a in enumerate(stores):
                if a in written:
                    failures.append(("written twice", a, n))
                written[a] = (x, i, n + 1)
        for x in nodes:
            for i in range(19):
                up = x - lat.off(i)
                if up in is_bulk and pulled[(x, i)] != (up, i, n):
                    failures.append(("stream", x, i, n, pulled[(x, i)]))
        lat.mem.update(written)
        for x in nodes:
            for i in range(19):
                if lat.mem.get(lat.readback(x, i, n + 1)) != (x, i, n + 1):
                    failures.append(("readback", x, i, n + 1))
    return failures


@pytest.mark.parametrize("dims", [(6, 5, 4), (9, 4, 7)])
def test_the_in_place_map_streams_and_reads_back(dims):
    assert run(Lattice(*dims), 6) == []


def test_a_planted_mistake_is_caught():
    failures = run(Lattice(6, 5, 4, shift_b_store=False), 3)
    assert any(f[0] == "stream" for f in failures)


def test_b_step_sources_stay_inside_the_guards():
    """Every B-step source of any node of the run (bulk or not, loaded before its class is
    known) lies within one guard of the planes: element indices in [-g, cells + g)."""
    for dims in [(6, 5, 4), (16, 3, 3), (3, 16, 5)]:
        lat = Lattice(*dims)
        for x in range(lat.cells):
            loads, stores = lat.addresses(x, 0)
            for _, e in loads + stores:
                assert -lat.guard <= e < lat.cells + lat.guard
