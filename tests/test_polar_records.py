"""The batched sweep's results as experiment records and their comparison table (one table:
every case is the same periodic-span section), with the least-squares lift slope over all
angles and the CSV of src/app/aero/aero_comparison.cpp."""
import math

import numpy as np
import pytest

from orrery_amd import experiment as ex
from orrery_amd import polar


def solves():
    cases = [polar.WingCase(k, 96, 48, 48, 48, aoa, 800.0) for k, aoa in enumerate((6.0, -2.0, 2.0))]
    cl = {6.0: 0.62, -2.0: -0.21, 2.0: 0.2}
    return [polar.Solve(c, 4000, 0.05 + 0.001 * c.aoa_deg ** 2, cl[c.aoa_deg], True, 1.0) for c in cases]


def test_one_table_sorted_by_angle_with_the_lift_slope():
    tables = ex.compare_by_geometry(polar.records(solves()))
    assert len(tables) == 1
    t = tables[0]
    assert t.label == "NACA 0012 section, chord 1 m"
    assert [r.angle_of_attack_deg for r in t.rows] == [-2.0, 2.0, 6.0]
    slope = np.polyfit([-2.0, 2.0, 6.0], [-0.21, 0.2, 0.62], 1)[0]
    assert t.lift_slope_per_rad == pytest.approx(slope * 180.0 / math.pi, rel=1e-13)


def test_csv_rows():
    text = polar.comparison_csv(solves())
    lines = text.split("\r\n")
    assert lines[0].startswith("geometry,name,solver,aoa_deg") and lines[-1] == ""
    cd = 0.05 + 0.001 * 4.0
    assert lines[1] == ('"NACA 0012 section, chord 1 m",case 01 (f32),lbm,-2,0,30,' + ex.exact(cd) + ",-0.21,0,0,0,0,"
                        + ex.exact(-0.21 / cd) + ",1")
    assert len(lines) == 5
