"""Experiment records: the cases of the reference's tests/scene/test_aero_experiment.cpp (round
trips, defaults, the error texts naming the offending key, the optional series), its bundled
experiment files (tests/golden/experiments: assets/experiments and tests/scene/fixtures of the
reference), the number layout of the JSON writer, and the comparison tables and CSV of
src/app/aero/aero_comparison.cpp against hand-computed values."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from orrery_amd import experiment as ex

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(HERE, "golden", "experiments")


def fixture(name):
    with open(os.path.join(FIXTURES, name)) as f:
        return f.read()


def naca_with_result():
    """test_aero_experiment.cpp:22-56"""
    e = ex.AeroExperiment(name="NACA 0012, 4 deg")
    e.geometry.shape = ex.AeroShape.naca_section
    e.geometry.naca = "0012"
    e.geometry.dimensions.chord_m = 0.75
    e.conditions.speed_m_s = 34.5
    e.conditions.angle_of_attack_deg = 4.0
    e.conditions.sideslip_deg = -0.1
    e.solver.solver = ex.SolverKind.lbm
    e.solver.reference_area = ex.ReferenceAreaKind.planform
    e.solver.reference_length_m = 0.75
    e.solver.lbm_cells_per_length = 64.0
    e.solver.lbm_duration = 40.0
    e.solver.steps_per_frame = 250
    e.result = ex.AeroResultRecord(solver=ex.SolverKind.lbm, cd=0.0412345678901234, cl=0.4382910571234567,
                                   cm=-0.0123, cd_pressure=0.03, reynolds=1.77e6, mach=0.1013,
                                   reference_area_m2=0.75, reference_length_m=0.75, iterations=48_000,
                                   converged=True, strouhal=0.1723, warnings=["lbm.reynolds", "lbm.les"],
                                   residuals=[1.0, 0.1, 1e-3, 4.9406564584124654e-324],
                                   history_step=[10.0, 20.0, 30.0], history_cd=[0.5, 0.45, 0.41],
                                   history_cl=[0.1, 0.3, 0.43])
    return e


def test_aero_records_round_trip_exactly():
    """test_aero_experiment.cpp:111-128"""
    e = naca_with_result()
    text = ex.save_aero_experiment(e)
    assert '"domain": "aero"' in text and '"version": 1' in text
    assert ex.parse_aero_experiment(text) == e
    generic = ex.to_experiment(e)
    assert generic.inputs["shape"] == "naca_section" and generic.inputs["chord_m"] == 0.75
    assert "diameter_m" not in generic.inputs
    assert generic.result["warnings"] == "lbm.reynolds; lbm.les"
    assert len(generic.series["history_cd"]) == 3


@pytest.mark.parametrize("shape", list(ex.AeroShape))
def test_every_shape_round_trips_with_its_dimensions(shape):
    """test_aero_experiment.cpp:130-150"""
    e = ex.AeroExperiment(name=shape.value)
    e.geometry.shape = shape
    text = ex.save_aero_experiment(e)
    for key in ex.shape_dimensions(shape):
        assert f'"{key}"' in text
    back = ex.parse_aero_experiment(text)
    assert back.geometry.shape == shape and back.result is None and back.solver.solver is None
    assert back == e


def test_user_models_round_trip():
    """test_aero_experiment.cpp:152-165"""
    e = ex.AeroExperiment(name="glider")
    e.geometry = ex.GeometryRef(source=ex.GeometrySource.url, path="cache/glider.stl", url="https://example.org/g.stl",
                                sha256="ab" * 32, license_note="CC-BY 4.0", unit_scale=0.001, up_axis="y")
    assert e.geometry.user_model
    assert ex.parse_aero_experiment(ex.save_aero_experiment(e)) == e


def test_minimal_fixture_loads_with_defaults():
    """test_aero_experiment.cpp:167-183 on tests/scene/fixtures/aero_sphere_minimal.json"""
    e = ex.parse_aero_experiment(fixture("aero_sphere_minimal.json"))
    assert e.geometry.shape == ex.AeroShape.sphere and e.geometry.dimensions.diameter_m == 0.2
    assert e.conditions.speed_m_s == 10.0 and e.conditions.density_kg_m3 == 1.225
    assert e.solver.solver == ex.SolverKind.newtonian
    assert e.solver.reference_area == ex.ReferenceAreaKind.frontal
    flow = e.conditions.flow()
    assert flow.speed == 10.0 and flow.angle_of_attack == 0.0


@pytest.mark.parametrize("name", ["aero_apollo_capsule.json", "aero_naca0012_lbm.json", "aero_naca0012_panel.json",
                                  "aero_naca2412_wing.json"])
def test_bundled_aero_experiments_load_and_round_trip(name):
    e = ex.parse_aero_experiment(fixture(name))
    assert e.result is None
    assert ex.parse_aero_experiment(ex.save_aero_experiment(e)) == e


def test_bundled_lbm_experiment_values():
    e = ex.parse_aero_experiment(fixture("aero_naca0012_lbm.json"))
    assert e.geometry.shape == ex.AeroShape.naca_section and e.geometry.naca == "0012"
    assert e.geometry.dimensions.chord_m == 0.1
    assert e.conditions.speed_m_s == 0.0584 and e.conditions.viscosity_pa_s == 1.789e-05
    assert e.conditions.angle_of_attack_deg == 4.0
    assert e.solver.solver == ex.SolverKind.lbm and e.solver.reference_area == ex.ReferenceAreaKind.planform
    assert (e.solver.lbm_cells_per_length, e.solver.lbm_duration, e.solver.steps_per_frame) == (40.0, 40.0, 200)


def test_space_records_load_generically_and_are_not_aero():
    """test_aero_experiment.cpp:94-109"""
    text = fixture("black_hole.json")
    e = ex.parse_experiment(text)
    assert e.domain == "space" and not e.series
    assert "series" not in ex.save_experiment(e)
    with pytest.raises(ex.AeroRecordError) as err:
        ex.aero_from_experiment(e)
    assert str(err.value) == "domain: expected 'aero', found 'space'"


def error_of(inputs, conditions, solver, tail=""):
    text = ('{"version": 1, "domain": "aero", "inputs": ' + inputs + ', "conditions": ' + conditions + ', "solver": '
            + solver + tail + "}")
    with pytest.raises(ex.AeroRecordError) as err:
        ex.parse_aero_experiment(text)
    return str(err.value)


def test_errors_name_the_offending_key():
    """test_aero_experiment.cpp:185-219"""
    ok_inputs = '{"geometry": "primitive", "shape": "sphere"}'
    ok_solver = '{"name": "auto"}'
    assert error_of('{"shape": "torus"}', "{}", ok_solver).startswith("inputs.shape: unknown shape 'torus'")
    assert error_of('{"shape": "sphere", "chord_m": 1}', "{}", ok_solver) == \
        "inputs.chord_m: unknown key (allowed: name, geometry, shape, diameter_m)"
    assert error_of('{"shape": "naca_wing", "naca": "00x2"}', "{}", ok_solver) == \
        "inputs.naca: '00x2' is not a NACA 4-digit designation"
    assert error_of('{"geometry": "file"}', "{}", ok_solver) == "inputs.path: a file geometry needs a model path"
    assert error_of(ok_inputs, '{"speed_m_s": -3}', ok_solver) == "conditions.speed_m_s: must be positive and finite"
    assert error_of(ok_inputs, '{"mach": 0.3}', ok_solver).startswith("conditions.mach: unknown key")
    assert error_of(ok_inputs, "{}", '{"name": "rans"}') == \
        "solver.name: unknown solver 'rans' (allowed: auto, panel, newtonian, lbm)"
    assert error_of(ok_inputs, "{}", '{"name": "lbm", "settings": {"steps_per_frame": 2.5}}') == \
        "solver.settings.steps_per_frame: must be an integer in [1, 1000000]"
    assert error_of(ok_inputs, "{}", '{"name": "lbm", "settings": {"reference_area": "side"}}').startswith(
        "solver.settings.reference_area: unknown reference area 'side'")
    assert error_of(ok_inputs, "{}", ok_solver, ', "result": {"cd": 1}') == \
        "result.solver: a result names the solver that produced it"
    assert error_of(ok_inputs, "{}", ok_solver, ', "result": {"solver": "lbm"}, "series": {"history_cd": [1, 2]}') == \
        "series: history_step, history_cd and history_cl must have the same length"


def test_series_round_trip_and_stay_optional():
    """test_aero_experiment.cpp:65-92"""
    e = ex.Experiment(domain="aero", solver=ex.SolverInfo("lbm", "1"),
                      series={"residual": [1.0, 0.5, 1e-300], "cd": [0.25]})
    assert ex.parse_experiment(ex.save_experiment(e)) == e
    assert "series" not in ex.save_experiment(ex.Experiment(domain="aero", solver=ex.SolverInfo("auto")))
    base = '{"version": 1, "domain": "aero", "solver": {"name": "auto"}, "series": '
    with pytest.raises(ex.SceneError) as err:
        ex.parse_experiment(base + '{"residual": [1, "x"]}}')
    assert err.value.code == ex.SceneErrorCode.wrong_type and err.value.path == "series.residual[1]"
    with pytest.raises(ex.SceneError) as err:
        ex.parse_experiment(base + '{"cd": 3}}')
    assert err.value.path == "series.cd"


@pytest.mark.parametrize("text,code,path,message", [
    ('{"version": 2, "domain": "aero", "solver": {"name": "x"}}', "unsupported_version", "version",
     "experiment schema version 2 is not supported (supported: 1)"),
    ('{"version": 1.0, "domain": "aero", "solver": {"name": "x"}}', "wrong_type", "version",
     "expected integer, found number"),
    ('{"domain": "aero"}', "missing_key", "version", "required key 'version' is missing"),
    ('{"version": 1, "domain": "fluid", "solver": {"name": "x"}}', "invalid_value", "domain",
     "unsupported domain 'fluid' (allowed: space, aero, entry)"),
    ('{"version": 1, "domain": "aero", "extra": 1, "solver": {"name": "x"}}', "unknown_key", "extra",
     "unknown key 'extra' (allowed: version, domain, inputs, conditions, solver, result, series)"),
    ('{"version": 1, "domain": "aero", "solver": {"name": "x"}, "inputs": {"a": 1, "a": 2}}', "duplicate_key",
     "inputs.a", "key 'a' appears more than once"),
    ('{"version": 1, "domain": "aero", "solver": {"name": "x"}, "inputs": {"a": true}}', "wrong_type", "inputs.a",
     "expected number or string, found boolean"),
    ('{"version": 1, "domain": "aero", "solver": {"name": "x"}, "inputs": {"a": 1e-400}}', "invalid_value",
     "inputs.a", "number 1e-400 is not zero but underflows to zero"),
    ('{"version": 1, "domain": "aero", "solver": {"name": "x"}, "inputs": {"a": 1e400}}', "invalid_value",
     "inputs.a", "value must be finite"),
    ('{"version": 1, "domain": "aero", "solver": [], "inputs": {}}', "wrong_type", "solver",
     "expected object, found array"),
])
def test_generic_record_errors(text, code, path, message):
    with pytest.raises(ex.SceneError) as err:
        ex.parse_experiment(text)
    assert (err.value.code.name, err.value.path, err.value.message) == (code, path, message)


def test_explicit_zero_literals_and_syntax_errors():
    e = ex.parse_experiment('{"version": 1, "domain": "aero", "solver": {"name": "x"}, "inputs": {"a": -0.0, '
                            '"b": 0e-400}}')
    assert e.inputs["a"] == 0.0 and math.copysign(1.0, e.inputs["a"]) < 0 and e.inputs["b"] == 0.0
    for bad in ('{"version": 1,', '{"version": NaN}', "[1, 2"):
        with pytest.raises(ex.SceneError) as err:
            ex.parse_experiment(bad)
        assert err.value.code == ex.SceneErrorCode.syntax_error


@pytest.mark.parametrize("value,text", [
    (1.789e-05, "1.789e-05"), (0.0003, "0.0003"), (288.15, "288.15"), (6100.0, "6100.0"), (4.0, "4.0"),
    (1e14, "100000000000000.0"), (1e15, "1e+15"), (1e16, "1e+16"), (1.23456e-05, "1.23456e-05"), (-0.0, "-0.0"), (0.0, "0.0"),
    (0.1, "0.1"), (1.0 / 3.0, "0.3333333333333333"), (-2.5e-300, "-2.5e-300"), (1.5e+300, "1.5e+300"),
    (123456789012345.6, "123456789012345.6"), (4.9406564584124654e-324, "5e-324"), (math.nan, "null"),
])
def test_json_number_layout(value, text):
    """Shortest round-trip digits; fixed notation for decimal exponents in [-4, 15), with ".0" on
    integral values; scientific with a signed exponent of two or more digits otherwise."""
    assert ex.format_json_number(value) == text
    if math.isfinite(value):
        assert float(text) == value


def test_saved_layout():
    e = ex.Experiment(domain="aero", inputs={"shape": "sphere", "diameter_m": 0.2}, conditions={},
                      solver=ex.SolverInfo("lbm", "1", {"steps_per_frame": 200.0}), series={"history_cd": [0.5]})
    assert ex.save_experiment(e) == (
        '{\n  "version": 1,\n  "domain": "aero",\n  "inputs": {\n    "diameter_m": 0.2,\n    "shape": "sphere"\n  },\n'
        '  "conditions": {},\n  "solver": {\n    "name": "lbm",\n    "version": "1",\n    "settings": {\n'
        '      "steps_per_frame": 200.0\n    }\n  },\n  "result": {},\n  "series": {\n    "history_cd": [\n'
        '      0.5\n    ]\n  }\n}\n')


@pytest.mark.parametrize("value,text", [(4.0, "4"), (0.1, "0.1"), (1e-05, "1e-05"), (48000.0, "48000"), (1e22, "1e+22"),
                                        (-0.0123, "-0.0123"), (0.0, "0"), (1234567.0, "1234567"), (1.5e-7, "1.5e-07"),
                                        (0.001, "0.001"), (0.0001, "1e-04"), (123e-7, "1.23e-05"),
                                        (100000.0, "1e+05"), (5e-324, "5e-324")])
def test_exact_is_the_plain_shortest_form(value, text):
    assert ex.exact(value) == text


TO_CHARS = r"""
#include <charconv>
#include <cstdio>
#include <cstring>
int main() {
    double v;
    while (std::fread(&v, sizeof v, 1, stdin) == 1) {
        char buf[64];
        auto r = std::to_chars(buf, buf + sizeof buf, v);
        std::fwrite(buf, 1, r.ptr - buf, stdout);
        std::fputc('\n', stdout);
    }
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_exact_matches_the_standard_library(tmp_path):
    """exact() against std::to_chars itself (what aero_comparison.cpp:27-31 calls) on values of
    every magnitude."""
    src = tmp_path / "to_chars.cpp"
    src.write_text(TO_CHARS)
    exe = tmp_path / "to_chars"
    subprocess.run(["g++", "-std=c++17", "-O1", str(src), "-o", str(exe)], check=True)
    rng = np.random.default_rng(3)
    values = np.concatenate([rng.standard_normal(500) * 10.0 ** rng.integers(-12, 13, 500),
                             np.round(rng.standard_normal(300) * 1e4) / 10.0 ** rng.integers(0, 6, 300),
                             10.0 ** rng.integers(-310, 308, 200), rng.integers(-10 ** 6, 10 ** 6, 200).astype(float),
                             [0.0, -0.0, 5e-324, 1.7976931348623157e308, 0.1, 1e-4, 1e5, 123456.0]])
    out = subprocess.run([str(exe)], input=values.astype("<f8").tobytes(), capture_output=True, check=True)
    want = out.stdout.decode().split("\n")[:-1]
    got = [ex.exact(float(v)) for v in values]
    assert got == want


def finished(name, aoa, cd, cl, cm, geometry=None, converged=True):
    e = ex.AeroExperiment(name=name)
    if geometry is not None:
        e.geometry = geometry
    e.conditions.angle_of_attack_deg = aoa
    e.result = ex.AeroResultRecord(solver=ex.SolverKind.lbm, cd=cd, cl=cl, cm=cm, converged=converged)
    return e


def test_comparison_tables_and_lift_slope():
    wing = ex.GeometryRef(shape=ex.AeroShape.naca_wing, naca="0012")
    runs = [finished("b", 4.0, 0.03, 0.4, -0.01, wing), finished("a", 0.0, 0.02, 0.0, 0.0, wing),
            finished("c", 8.0, 0.05, 0.75, -0.02, wing), finished("s", 0.0, 0.47, 0.0, 0.0),
            ex.AeroExperiment(name="pending")]
    tables = ex.compare_by_geometry(runs)
    assert [t.label for t in tables] == ["NACA 0012 wing, chord 1 m, span 4 m", "Sphere, d 1 m"]
    rows = tables[0].rows
    assert [r.name for r in rows] == ["a", "b", "c"]
    assert rows[2].delta_cl == 0.75 and rows[1].delta_cd == pytest.approx(0.01)
    slope = np.polyfit([0.0, 4.0, 8.0], [0.0, 0.4, 0.75], 1)[0]
    assert tables[0].lift_slope_per_deg == pytest.approx(slope, rel=1e-14)
    assert tables[0].lift_slope_per_rad == pytest.approx(slope * 180.0 / math.pi, rel=1e-14)
    assert tables[1].lift_slope_per_deg is None  # one row
    assert rows[1].lift_to_drag == 0.4 / 0.03


def test_comparison_csv():
    g = ex.GeometryRef(source=ex.GeometrySource.file, path="models/=evil,model.stl")
    t = ex.compare_by_geometry([finished('x "1"', 2.0, 0.5, 0.25, -0.125, g, converged=False),
                                finished("-y", 0.0, 0.25, 0.0, 0.0, g)])[0]
    assert ex.comparison_csv(t) == (
        "geometry,name,solver,aoa_deg,sideslip_deg,speed_m_s,cd,cl,cm,delta_cd,delta_cl,delta_cm,l_over_d,converged\r\n"
        "\"'=evil,model.stl\",\"'-y\",lbm,0,0,30,0.25,0,0,0,0,0,0,1\r\n"
        "\"'=evil,model.stl\",\"x \"\"1\"\"\",lbm,2,0,30,0.5,0.25,-0.125,0.25,0.25,-0.125,0.5,0\r\n")
