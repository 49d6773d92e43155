"""Writes sphere_channel_cd.json: the oracle's drag history on the 256 x 128 x 128 sphere channel
(sphere of diameter 32 cells at (64, 64, 64) voxelised on the host, inlet u 0.05, Re 200,
Smagorinsky C 0.17, convective outlet, slip sides, halfway walls, seeded 1e-4 perturbation),
5000 steps with a force sample every 10. Takes about 15 minutes on 8 host threads.

    python tests/golden/make_sphere_cd.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

from oracle import Oracle3D, load  # noqa: E402
from orrery_amd import Lbm3DConfig, Smagorinsky, Wall, WallKind, XBoundary, geometry, tau_of_viscosity  # noqa: E402
from orrery_amd.polar import place_body  # noqa: E402

STEPS = 5000
U, D, RE = 0.05, 32.0, 200.0


def lattice():
    config = Lbm3DConfig(nx=256, ny=128, nz=128, tau=tau_of_viscosity(U * D / RE), x_boundary=XBoundary.inflow_outflow,
                         y_low=Wall(WallKind.slip), y_high=Wall(WallKind.slip), z_low=Wall(WallKind.slip),
                         z_high=Wall(WallKind.slip), inlet_ux=U, smagorinsky=Smagorinsky(True, 0.17))
    mesh = geometry.sphere_mesh(0.5 * D, geometry.RevolutionResolution(64, 0.05)).transformed(
        None, (0.0, 0.0, 0.0), (64.0, 64.0, 64.0))
    return config, place_body(mesh, 256, 128, 128, interpolated=False)


def main():
    load().orc_set_threads(os.cpu_count() or 1)
    config, mask = lattice()
    o = Oracle3D(config, mask)
    o.initialize_perturbed(2024, 1e-4)
    q = 0.5 * U * U * np.pi * (0.5 * D) ** 2
    t0 = time.perf_counter()
    fx = []
    for k in range(STEPS // 10):
        o.step(10)
        fx.append(o.force().force[0])
        if (k + 1) % 50 == 0:
            print(f"step {(k + 1) * 10}: Cd {fx[-1] / q:.6f} ({time.perf_counter() - t0:.0f} s)", flush=True)
    fx = np.array(fx)
    out = {"lattice": "256x128x128 sphere channel, D 32 at (64,64,64), u 0.05, Re 200, LES C 0.17, seed 2024",
           "solid_nodes": int(mask.solid.sum()), "steps": STEPS, "sample_every": 10, "reference_area": "pi (D/2)^2",
           "force_x": fx.tolist(), "cd_mean_last_1000_steps": float((fx[-100:] / q).mean())}
    with open(os.path.join(HERE, "sphere_channel_cd.json"), "w") as f:
        json.dump(out, f)
    print(f"Cd over the last 1000 steps: {out['cd_mean_last_1000_steps']:.12f}")


if __name__ == "__main__":
    main()
