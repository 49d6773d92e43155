"""The 256 x 128 x 128 sphere channel (Re 200, LES, 5000 steps) against the oracle's full run,
committed in tests/golden/sphere_channel_cd.json by tests/golden/make_sphere_cd.py: the FP64
lattice reproduces every drag sample to rounding, the binary32 lattice the mean Cd to 1e-3."""
import importlib.util
import json
import os

import numpy as np
import pytest

from orrery_amd import Lbm3DSolver

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "sphere_channel_cd.json")


def maker():
    spec = importlib.util.spec_from_file_location("make_sphere_cd", os.path.join(HERE, "golden", "make_sphere_cd.py"))
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    return module


@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="golden file not generated")
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_drag_history_against_the_oracle(storage):
    with open(GOLDEN) as f:
        golden = json.load(f)
    m = maker()
    config, mask = m.lattice()
    assert int(mask.solid.sum()) == golden["solid_nodes"]
    s = Lbm3DSolver(config, mask, storage)
    s.initialize_perturbed(2024, 1e-4)
    fx = []
    for _ in range(golden["steps"] // golden["sample_every"]):
        s.step(golden["sample_every"])
        fx.append(s.force().force[0])
    fx = np.array(fx)
    ref = np.array(golden["force_x"])
    q = 0.5 * m.U * m.U * np.pi * (0.5 * m.D) ** 2
    cd = float((fx[-100:] / q).mean())
    print(f"{storage}: Cd {cd:.12f} against the oracle's {golden['cd_mean_last_1000_steps']:.12f}; "
          f"largest drag difference {np.abs(fx - ref).max():.3e}", flush=True)
    if storage == "f64":
        assert np.abs(fx - ref).max() <= 1e-12 * np.abs(ref).max()
    else:
        assert abs(cd / golden["cd_mean_last_1000_steps"] - 1.0) < 1e-3
