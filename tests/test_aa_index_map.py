"""CPU emulation of the in-place (A-A) index map of the D3Q19 bulk kernel, with labels in place of
populations.

The bulk kernel (orrery_amd/csrc/lbm3d_step.h, bulk_run) addresses population A[k][e] (plane k,
cell e) of the single population array as below, with n the number of completed steps:
- n odd, A step: node x loads plane i at x and stores its output i to plane opp(i) at x;
- n even, B step: node x loads plane opp(i) at x - c_i (cells up to nx ny + nx + 1 outside the
  lattice: the guard cells) and stores output i to plane i at x + c_i.
The readback kernels report f*_i(x) from plane opp(i) at x after an even count and from plane
i at x + c_i after an odd one.

The emulation writes a label (node, direction, step) for every output and checks, over several
steps of a lattice whose interior nodes are bulk nodes: every node reads exactly the addresses
it writes (the in-place property), no address is written twice in a step, a bulk node's pull i
is the label its upstream bulk node x - c_i wrote one step before (the streaming), and the
readback map finds every label. A planted mistake in the map (the B step's store without its
shift) breaks the streaming check. The second part does the same for side nodes (face nodes
streamed in place through slip and periodic sides) next to the listed boundary nodes; the third
checks how the buffer stores A[k][e] (blocks of cells) and the bulk kernel's run accesses to it.
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


# ---- side and inlet nodes and the boundary nodes around them ---------------------------------------
# A lattice with slip or periodic y and z sides and an inflow/outflow x: the rows on the y and z
# faces (away from the x faces) hold side nodes, streamed in place with the B-step address of
# pull i at plane opp(d), element s, (s, d) the source node and direction through the sides; the
# inlet face off the side faces holds inlet nodes, streamed with the plain addresses (their
# pulls from x = -1 land on addresses no other node touches, and the inlet rule replaces them);
# every other face node is a listed boundary node, which reads the outputs of listed sources
# from their stash, those of streamed sources from the lattice (A step: its own cell, plane i;
# B step: plane opp(d) at s), and stores an output in the lattice when its reader is streamed
# (A step: its own cell, plane opp(k); B step: the reader's A-step address, plane d' at r,
# (r, d') the reader and its pull through the sides).
MY = [0, 1, 2, 4, 3, 5, 6, 9, 10, 7, 8, 11, 12, 13, 14, 18, 17, 16, 15]
MZ = [0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16]
SLIP, PERIODIC = "slip", "periodic"


class SideLattice:
    def __init__(self, nx, ny, nz, sides, mirror_side_loads=True):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.cells = nx * ny * nz
        self.sides = sides  # y low, y high, z low, z high
        self.mirror_side_loads = mirror_side_loads
        self.mem = {}
        self.stash = {}
        self.cls = {}
        for x in range(self.cells):
            xx, yy, zz = self.xyz(x)
            face = yy in (0, ny - 1) or zz in (0, nz - 1)
            if xx == 0 and not face:
                self.cls[x] = 4
            elif xx in (0, nx - 1):
                self.cls[x] = 2
            elif face:
                self.cls[x] = 3 if all(self.resolve(x, i)[0] == "node" for i in range(19)) else 2
            else:
                self.cls[x] = 1

    def xyz(self, x):
        return x % self.nx, (x // self.nx) % self.ny, x // (self.nx * self.ny)

    def at(self, a, b, c):
        return a + self.nx * (b + self.ny * c)

    def resolve(self, x, i):
        """(kind, source node, source direction) of pull i: y sides, then z, then x."""
        a, b, c = self.xyz(x)
        sa, sb, sc, d = a - CX[i], b - CY[i], c - CZ[i], i
        if not 0 <= sb < self.ny:
            if self.sides[0 if sb < 0 else 1] == PERIODIC:
                sb %= self.ny
            else:
                sb, d = b, MY[d]
        if not 0 <= sc < self.nz:
            if self.sides[2 if sc < 0 else 3] == PERIODIC:
                sc %= self.nz
            else:
                sc, d = c, MZ[d]
        if not 0 <= sa < self.nx:
            return ("inlet" if sa < 0 else "outlet", None, None)
        return ("node", self.at(sa, sb, sc), d)

    def reader(self, x, k):
        """(reader node, its pull direction) of output k of node x, or None."""
        a, b, c = self.xyz(x)
        ta, tb, tc, d = a + CX[k], b + CY[k], c + CZ[k], k
        if not 0 <= tb < self.ny:
            if self.sides[0 if tb < 0 else 1] == PERIODIC:
                tb %= self.ny
            else:
                tb, d = b, MY[d]
        if not 0 <= tc < self.nz:
            if self.sides[2 if tc < 0 else 3] == PERIODIC:
                tc %= self.nz
            else:
                tc, d = c, MZ[d]
        if not 0 <= ta < self.nx:
            return None
        return self.at(ta, tb, tc), d

    def b_address(self, x, i):
        """B-step load address of pull i of a streamed node (the store of output opp(i))."""
        kind, s, d = self.resolve(x, i)
        if kind == "inlet":  # an inlet node's pull from x = -1: the plain address
            return (OPP[i], x - CX[i] - self.nx * CY[i] - self.nx * self.ny * CZ[i])
        assert kind == "node"
        if self.cls[x] == 3 and not self.mirror_side_loads:
            return (OPP[i], x - CX[i] - self.nx * CY[i] - self.nx * self.ny * CZ[i])
        return (OPP[d], s)

    def readback(self, x, i, n):
        return (OPP[i], x) if n % 2 == 0 else self.b_address(x, OPP[i])


def run_sides(lat, steps):
    streamed = [x for x in range(lat.cells) if lat.cls[x] in (1, 3, 4)]
    listed = [x for x in range(lat.cells) if lat.cls[x] == 2]
    for x in range(lat.cells):
        for i in range(19):
            lat.mem[(OPP[i], x)] = (x, i, 0)
            if lat.cls[x] == 2:
                lat.stash[(x, i)] = (x, i, 0)
    failures = []
    for n in range(steps):
        local = n % 2 == 1
        written, pulled = {}, {}
        read_by, writer = {}, {}  # lattice address -> the node reading / writing it this step
        for x in streamed:
            if local:
                loads = [(i, x) for i in range(19)]
                stores = [(OPP[i], x) for i in range(19)]
            else:
                loads = [lat.b_address(x, i) for i in range(19)]
                stores = [lat.b_address(x, OPP[i]) for i in range(19)]
            if sorted(loads) != sorted(stores):
                failures.append(("not in place", x, n))
            for _, e in loads:
                if not -(lat.nx * lat.ny + lat.nx + 1) <= e < lat.cells + lat.nx * lat.ny + lat.nx + 1:
                    failures.append(("outside the guards", x, n))
            for i, a in enumerate(loads):
                pulled[(x, i)] = lat.mem.get(a)
                read_by.setdefault(a, set()).add(x)
            for k, a in enumerate(stores):
                if a in written:
                    failures.append(("written twice", a, n))
                written[a] = (x, k, n + 1)
                writer[a] = x
        new_stash = {}
        for x in listed:
            for i in range(19):
                kind, s, d = lat.resolve(x, i)
                if kind != "node":
                    continue
                if lat.cls[s] == 2:
                    pulled[(x, i)] = lat.stash[(s, d)]
                else:
                    a = (i, x) if local else (OPP[d], s)
                    pulled[(x, i)] = lat.mem.get(a)
                    read_by.setdefault(a, set()).add(x)
            for k in range(19):
                new_stash[(x, k)] = (x, k, n + 1)
                r = lat.reader(x, k)
                if r is None or lat.cls[r[0]] == 2:
                    continue
                a = (OPP[k], x) if local else (r[1], r[0])
                if a in written:
                    failures.append(("written twice", a, n))
                written[a] = (x, k, n + 1)
                writer[a] = x
        # the kernels of a step run one after another: an address one node writes must not be
        # read by another node in the same step
        for a, w in writer.items():
            if read_by.get(a, {w}) - {w}:
                failures.append(("read and written by different nodes", a, n))
        for x in streamed + listed:
            for i in range(19):
                kind, s, d = lat.resolve(x, i)
                if kind == "node" and pulled[(x, i)] != (s, d, n):
                    failures.append(("stream", x, i, n, pulled[(x, i)], (s, d, n)))
        lat.mem.update(written)
        lat.stash.update(new_stash)
        for x in streamed:
            for i in range(19):
                if lat.mem.get(lat.readback(x, i, n + 1)) != (x, i, n + 1):
                    failures.append(("readback", x, i, n + 1))
    return failures


@pytest.mark.parametrize("sides", [(SLIP, SLIP, SLIP, SLIP), (PERIODIC, PERIODIC, SLIP, SLIP),
                                   (SLIP, SLIP, PERIODIC, PERIODIC), (PERIODIC, PERIODIC, PERIODIC, PERIODIC)])
@pytest.mark.parametrize("dims", [(6, 5, 4), (5, 3, 6)])
def test_side_nodes_stream_in_place_through_slip_and_periodic_sides(dims, sides):
    lat = SideLattice(*dims, sides)
    assert any(c == 3 for c in lat.cls.values()) and any(c == 4 for c in lat.cls.values())
    assert run_sides(lat, 6) == []


def test_side_nodes_without_their_mirrored_addresses_are_caught():
    failures = run_sides(SideLattice(6, 5, 4, (SLIP,) * 4, mirror_side_loads=False), 3)
    assert any(f[0] == "stream" for f in failures)


# ---- the blocked layout of the population buffer ---------------------------------------------------
# The (plane, element) addresses above are stored in blocks of 2^k cells (Lattice3::f,
# orrery_amd/csrc/lbm_device.h): A[dir][c] is element (g >> k) 19 2^k + dir 2^k + (g mod 2^k),
# g = c + guard (guarded_element, lbm3d_step.h). The bulk kernel moves runs of V consecutive cells
# with one access per direction; in the B step a run displaced by an x shift may cross into the
# next block, where its cells are no longer adjacent, and such runs are moved cell by cell
# (load_shifted_all, store_shifted), which a wave decides from uniform values (wave_crosses).
# The emulation checks, on layouts with small blocks (many crossings): the element map is a
# bijection onto the buffer; every access a B-step run makes, the combined access included,
# stays in the buffer; after the fix-ups each cell's load and store element is the per-cell one;
# the per-wave test agrees with a lane-by-lane scan; without the fix-ups the check fails.
# A wave whose 64 V cells stay inside one block under every direction's shift (shift_inside)
# forms its B-step elements as each lane's own element of direction 0 plus a uniform term
# (shift_term); the emulation checks those elements against the per-cell ones, the uniform test
# against a lane scan, and that a planted mistake in the term is caught.


def padded(n, m):
    return (n + m - 1) // m * m


class Blocked:
    """population_layout (lbm_internal.h) with a given block shift."""

    def __init__(self, nx, ny, nz, shift):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.cells = nx * ny * nz
        self.shift = shift
        self.block = 1 << shift
        self.mask = self.block - 1
        self.guard = padded(nx * ny + nx + 1 + 4, self.block)
        self.blocks = (self.cells + 2 * self.guard + self.block - 1) // self.block
        self.elements = 19 * self.block * self.blocks

    def off(self, i):
        return CX[i] + self.nx * CY[i] + self.nx * self.ny * CZ[i]

    def element(self, g, d):
        return (((g >> self.shift) * 19 + d) << self.shift) | (g & self.mask)

    def crosses(self, g, v):
        return v > 1 and (g & self.mask) > self.mask + 1 - v

    def wave_crosses(self, group, dshift, v):
        if v == 1 or dshift & (v - 1) == 0:
            return False
        start = (-v - (dshift & ~(v - 1))) & self.mask
        return start & ~(64 * v - 1) == group

    def shift_inside(self, group, d, v):
        return ((group + d) & self.mask) <= self.mask + 1 - 64 * v

    def shift_term(self, group, d, direction, per_block=18):
        q = (group + d) >> self.shift  # floor, also for a shift back
        return (((q * per_block + direction) << self.shift) + d) % 2**32


def b_step_run(lat, c0, v, fix=True):
    """Elements a run of v cells starting at c0 loads and stores per direction in the B step, as
    the bulk kernel forms them (combined access from the run's first element, then the crossing
    runs' cells one by one), and the elements its combined accesses touch."""
    g0 = c0 + lat.guard
    loads, stores, touched = [], [], []
    for i in range(19):
        g = g0 - lat.off(i)
        e = lat.element(g, OPP[i])
        run = [e + k for k in range(v)]
        touched += run
        if fix and lat.crosses(g, v):
            run = [run[0]] + [lat.element(g + k, OPP[i]) for k in range(1, v)]
            touched += run
        loads.append(run)
        g = g0 + lat.off(i)
        if fix and lat.crosses(g, v):
            run = [lat.element(g + k, i) for k in range(v)]
        else:
            run = [lat.element(g, i) + k for k in range(v)]
        touched += run
        stores.append(run)
    return loads, stores, touched


@pytest.mark.parametrize("dims,shift", [((12, 6, 5), 8), ((20, 7, 6), 8), ((16, 16, 6), 9)])
def test_blocked_layout_is_a_bijection_onto_the_buffer(dims, shift):
    lat = Blocked(*dims, shift)
    seen = np.zeros(lat.elements, dtype=bool)
    for g in range(lat.blocks * lat.block):
        for d in range(19):
            e = lat.element(g, d)
            assert 0 <= e < lat.elements and not seen[e]
            seen[e] = True
    assert seen.all()


@pytest.mark.parametrize("dims,shift", [((12, 6, 5), 8), ((20, 7, 6), 8), ((17, 9, 5), 8), ((16, 16, 6), 9)])
@pytest.mark.parametrize("v", [2, 4])
def test_blocked_runs_address_their_cells_in_the_b_step(dims, shift, v):
    lat = Blocked(*dims, shift)
    crossing = 0
    for c0 in range(0, padded(lat.cells, v), v):
        loads, stores, touched = b_step_run(lat, c0, v)
        assert all(0 <= e < lat.elements for e in touched)
        for i in range(19):
            truth_load = [lat.element(c0 + k + lat.guard - lat.off(i), OPP[i]) for k in range(v)]
            truth_store = [lat.element(c0 + k + lat.guard + lat.off(i), i) for k in range(v)]
            assert loads[i] == truth_load
            assert stores[i] == truth_store
            crossing += lat.crosses(c0 + lat.guard - lat.off(i), v)
    assert crossing > 0  # the case the fix-ups exist for does occur


@pytest.mark.parametrize("v", [2, 4])
def test_the_per_wave_crossing_test_matches_a_lane_scan(v):
    for shift in (8, 9, 11):
        lat = Blocked(4, 4, 4, shift)
        shifts = [s for s in range(-3 * lat.block - 5, 3 * lat.block + 5, 3)] + [1, -1, 2, -2, 3, -3]
        for group in range(0, lat.block, 64 * v):
            for d in shifts:
                scan = any(((group + v * lane + d) & lat.mask) > lat.block - v for lane in range(64))
                assert lat.wave_crosses(group, d, v) == scan, (shift, group, d)


def test_blocked_runs_without_the_fix_ups_are_caught():
    lat = Blocked(12, 6, 5, 8)
    wrong = 0
    for c0 in range(0, lat.cells, 2):
        loads, _, _ = b_step_run(lat, c0, 2, fix=False)
        for i in range(19):
            truth = [lat.element(c0 + k + lat.guard - lat.off(i), OPP[i]) for k in range(2)]
            wrong += loads[i] != truth
    assert wrong > 0


def uniform_term_mismatches(lat, v, per_block=18):
    """Waves inside a block under every shift, and the elements their uniform terms get wrong."""
    inside_waves, wrong = 0, 0
    for c0 in range(0, padded(lat.cells, 64 * v), 64 * v):
        g0 = c0 + lat.guard
        group = g0 & lat.mask & ~(64 * v - 1)
        inside = all(lat.shift_inside(group, -lat.off(i), v) for i in range(1, 19))
        # the load shifts of all directions are their store shifts, and the uniform test is a lane scan
        assert inside == all(lat.shift_inside(group, lat.off(i), v) for i in range(1, 19))
        assert inside == all(((g0 + k - lat.off(i)) >> lat.shift) == ((g0 - lat.off(i)) >> lat.shift)
                             for i in range(19) for k in range(64 * v))
        if not inside:
            continue
        inside_waves += 1
        for k in range(64 * v):
            g = g0 + k
            own = lat.element(g, 0)
            for i in range(19):
                load = (own + lat.shift_term(group, -lat.off(i), OPP[i], per_block)) % 2**32
                store = (own + lat.shift_term(group, lat.off(i), i, per_block)) % 2**32
                wrong += load != lat.element(g - lat.off(i), OPP[i])
                wrong += store != lat.element(g + lat.off(i), i)
    return inside_waves, wrong


# shapes whose inside waves include shifts that move whole blocks (q != 0) at every v, and two
# where they do not
@pytest.mark.parametrize("dims,shift", [((20, 30, 6), 11), ((40, 30, 6), 10), ((24, 40, 5), 10), ((70, 9, 5), 11),
                                        ((12, 6, 5), 10), ((17, 9, 5), 10)])
@pytest.mark.parametrize("v", [1, 2, 4])
def test_waves_inside_a_block_address_by_a_uniform_term(dims, shift, v):
    lat = Blocked(*dims, shift)
    inside_waves, wrong = uniform_term_mismatches(lat, v)
    assert inside_waves > 0 and wrong == 0


def test_a_planted_mistake_in_the_uniform_term_is_caught():
    # 600-cell planes in blocks of 512 cells: the z shifts move whole blocks (q != 0)
    lat = Blocked(20, 30, 6, 9)
    inside_waves, wrong = uniform_term_mismatches(lat, 1, per_block=19)
    assert inside_waves > 0 and wrong > 0
