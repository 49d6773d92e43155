"""Steps replayed from captured HIP graphs against the same steps launched one by one.

A call of at least 8 steps runs as replays of one captured graph of up to 64 steps; the graph is
captured again when the inlet, schedule, launch shape or starting parity changes. Each check
below reads the state back (eager work and a synchronisation between replays) and compares it
bit for bit with a twin lattice whose graphs are off, after every call of the sequence.
"""
import numpy as np
import pytest

from cases import cylinder_channel_2d, sphere_channel
from orrery_amd import Lbm2DSolver, Lbm3DSolver, Smagorinsky, SolidWall

pytestmark = pytest.mark.gpu

# (steps, inlet scale): a long call (captured at once), the same again (the same graph), an odd
# call (parity 1 afterwards: eager, too short), a long call at odd parity (a new graph), a changed
# inlet twice (eager on the first call, captured on the second), and back to the first inlet.
SEQUENCE = [(300, 1.0), (300, 1.0), (7, 1.0), (260, 1.0), (100, 0.9), (100, 0.9), (129, 1.0), (64, 1.0)]


def run_pair(make, set_inlet, read):
    eager, graphs = make(), make()
    eager.set_graphs(False)
    for k, (steps, scale) in enumerate(SEQUENCE):
        for s in (eager, graphs):
            set_inlet(s, scale)
            s.step(steps)
        a, b = read(eager), read(graphs)
        print(f"call {k}: {steps} steps, inlet x{scale}: identical {all(np.array_equal(x, y) for x, y in zip(a, b))}",
              flush=True)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    assert eager.steps() == graphs.steps() == sum(s for s, _ in SEQUENCE)


@pytest.mark.parametrize("storage,schedule", [("f32", "serial"), ("f32", "fused"), ("f16s", "serial"),
                                              ("f64", "serial")])
def test_d3q19_graph_replays_are_bitwise_eager(storage, schedule):
    c, mask = sphere_channel(70, 21, 19, 7.0, 2.0, 0.05, 60.0)
    c.smagorinsky = Smagorinsky(True, 0.17)
    c.solid_wall = SolidWall.interpolated
    u0 = c.inlet_ux

    def make():
        s = Lbm3DSolver(c, mask, storage)
        s.set_schedule(schedule)
        s.initialize_perturbed(5, 1e-4)
        return s

    def set_inlet(s, scale):
        s.set_inlet_velocity((u0 * scale, 0.0, 0.0))

    def read(s):
        f = s.force()
        return [s.populations(), np.array(f.force), np.array(f.moment)]

    run_pair(make, set_inlet, read)


def test_d3q19_launch_shape_change_recaptures():
    c, mask = sphere_channel(40, 16, 16, 6.0, 2.0, 0.05, 50.0)
    a, b = Lbm3DSolver(c, mask), Lbm3DSolver(c, mask)
    a.set_graphs(False)
    for block, cpt in [(256, 1), (128, 2), (256, 1), (64, 4)]:
        for s in (a, b):
            s.set_launch(block, cpt)
            s.step(256)
        assert np.array_equal(a.populations(), b.populations())


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_d2q9_graph_replays_are_bitwise_eager(storage):
    c, m = cylinder_channel_2d(200, 100, 50.0, 50.5, 10.0, 0.05, 100.0)
    c.smagorinsky = Smagorinsky(True, 0.17)

    def make():
        s = Lbm2DSolver(c, m, storage)
        s.initialize_perturbed(3, 1e-4)
        return s

    def set_inlet(s, scale):
        # The 2D handle has no inlet setter: the scale only changes the 3D cases.
        pass

    def read(s):
        f = s.force()
        return [s.populations(), np.array([f.x, f.y, f.moment])]

    run_pair(make, set_inlet, read)
