"""Experiment records run on the device lattice: the JSON request/response round trip, its
result record (reference values, Reynolds and Mach numbers, history), determinism, and the
command-driven runner reaching the same record as a direct run."""
import math

import numpy as np
import pytest

from orrery_amd import experiment as ex
from orrery_amd import runner

pytestmark = pytest.mark.gpu

REQUEST = """{
  "version": 1,
  "domain": "aero",
  "inputs": { "name": "sphere, short lattice run", "geometry": "primitive", "shape": "sphere", "diameter_m": 0.2 },
  "conditions": { "speed_m_s": 10, "angle_of_attack_deg": 5 },
  "solver": { "name": "lbm", "version": "1",
              "settings": { "reference_area": "frontal", "lbm_cells_per_length": 8, "lbm_duration": 3,
                            "steps_per_frame": 40 } },
  "result": {}
}
"""


def test_request_response_round_trip():
    text = runner.serve_request(REQUEST)
    e = ex.parse_aero_experiment(text)
    r = e.result
    print(f"record: cd {r.cd:.5f} cl {r.cl:.5f} iterations {r.iterations} Re {r.reynolds:.4g} warnings {r.warnings}",
          flush=True)
    assert r is not None and r.solver == ex.SolverKind.lbm
    assert r.iterations > 0 and math.isfinite(r.cd) and r.cd > 0.0
    flow = e.conditions.flow()
    assert r.reynolds == flow.reynolds(r.reference_length_m) and r.mach == flow.mach()
    prepared = runner.prepare_geometry(e.geometry)
    assert r.reference_area_m2 == prepared.quantities.frontal_area
    assert r.reference_length_m == prepared.quantities.reference_length
    assert len(r.history_step) == len(r.history_cd) == len(r.history_cl) > 0
    assert ex.save_aero_experiment(e) == text  # the response reads back to the same text
    again = ex.parse_aero_experiment(runner.serve_request(REQUEST))
    assert again == e  # deterministic


def test_runner_commands_reach_the_direct_result():
    e = ex.parse_aero_experiment(REQUEST)
    direct = runner.run_experiment(e).result
    r = runner.ExperimentRunner()
    try:
        r.send(runner.StartRun(e, to_convergence=False))
        p = r.wait({runner.RunState.running, runner.RunState.failed}, timeout=120)
        assert p.status.state == runner.RunState.running, p.status.message
        for _ in range(3):
            r.wake()
        r.send(runner.SetPaused(True))
        p = r.wait({runner.RunState.paused}, timeout=120)
        r.send(runner.StepFrame())
        r.send(runner.SetToConvergence(True))
        r.send(runner.SetPaused(False))
        p = r.wait({runner.RunState.finished, runner.RunState.failed}, timeout=300)
        assert p.status.state == runner.RunState.finished, p.status.message
        assert p.status.run == 1 and p.status.steps_per_second > 0.0
        got = p.record.result
        assert (got.cd, got.cl, got.cm, got.iterations, got.converged) == \
            (direct.cd, direct.cl, direct.cm, direct.iterations, direct.converged)
        assert np.array_equal(got.history_cd, direct.history_cd)
        r.send(runner.StopRun())
        assert r.wait({runner.RunState.idle}, timeout=60).status.state == runner.RunState.idle
    finally:
        r.close()


def test_bundled_naca_section_record_runs_on_the_2d_lattice():
    """The reference's bundled record assets/experiments/aero_naca0012_lbm.json (NACA 0012 at
    4 degrees, Re 400, 40 cells per chord, 40 convective times) through the JSON request path:
    the 2D adapter, per-span reference values, the history, and a positive lift."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "experiments", "aero_naca0012_lbm.json")
    with open(path) as f:
        text = f.read()
    e = ex.parse_aero_experiment(runner.serve_request(text))
    r = e.result
    flow = e.conditions.flow()
    print(f"measured: NACA 0012 section record: cl {r.cl:.4f} cd {r.cd:.4f} cm {r.cm:.4f} St {r.strouhal} "
          f"converged {r.converged} iterations {r.iterations} warnings {r.warnings}", flush=True)
    assert r.solver == ex.SolverKind.lbm and r.iterations == 32000
    assert r.reference_area_m2 == 0.1 and r.reference_length_m == 0.1
    assert r.reynolds == flow.reynolds(0.1) and r.mach == flow.mach()
    assert 0.05 < r.cl < 0.6 and 0.02 < r.cd < 0.5
    assert len(r.history_step) == len(r.history_cd) == 3200


def test_section_record_through_the_runner_commands():
    """A 2D record driven by the command thread reaches the direct run's record."""
    req = REQUEST.replace('"shape": "sphere", "diameter_m": 0.2', '"shape": "naca_section", "naca": "2412", '
                          '"chord_m": 0.1').replace('"lbm_cells_per_length": 8', '"lbm_cells_per_length": 16')
    e = ex.parse_aero_experiment(req)
    direct = runner.run_experiment(e).result
    assert direct.iterations == 960 and direct.cl > 0.0
    r = runner.ExperimentRunner()
    try:
        r.send(runner.StartRun(e, to_convergence=True))
        p = r.wait({runner.RunState.finished, runner.RunState.failed}, timeout=300)
        assert p.status.state == runner.RunState.finished, p.status.message
        got = p.record.result
        assert (got.cd, got.cl, got.cm, got.iterations) == (direct.cd, direct.cl, direct.cm, direct.iterations)
        assert np.array_equal(got.history_cl, direct.history_cl)
    finally:
        r.close()
