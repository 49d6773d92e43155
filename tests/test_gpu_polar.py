"""The batched wing-section sweep: the batch and the one-at-a-time runs solve every case to the
same step with the same coefficients (the batch's lattices evolve bit for bit as they would
alone), and the lift grows with the angle of attack."""
import pytest

from orrery_amd import polar

pytestmark = pytest.mark.gpu


def small_cases():
    return [polar.WingCase(k, 96, 24, 48, 24, aoa, 600.0) for k, aoa in enumerate((-2.0, 2.0, 6.0))]


def test_batched_and_sequential_solves_agree():
    cases = small_cases()
    lattices = [polar.wing_lattice(c) for c in cases]
    batched, tb = polar.run_batched(cases, lattices)
    sequential, ts = polar.run_sequential(cases, lattices)
    by_case = {s.case.index: s for s in sequential}
    for b in batched:
        s = by_case[b.case.index]
        print(f"measured: case {b.case.index} aoa {b.case.aoa_deg}: steps {b.steps} cd {b.cd:.6f} cl {b.cl:.6f} "
              f"converged {b.converged}", flush=True)
        assert (b.steps, b.cd, b.cl, b.converged) == (s.steps, s.cd, s.cl, s.converged)
    print(f"measured: batched {tb:.2f} s, sequential {ts:.2f} s")
    slope = polar.lift_slope(batched)
    assert slope > 0.0
