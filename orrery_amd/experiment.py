"""Experiment records of the aero domain as JSON: the request and response format of the
experiment runner, and the comparison tables (with the lift slope) and CSV export of a set of
finished records.

Three layers, each restating the reference's:
- the generic record (schema version 1): parse_experiment / save_experiment
  (src/scene/experiment.cpp, src/scene/json_support.cpp): strict JSON (duplicate keys, NaN and
  underflowing literals rejected, the document limits of src/scene/include/orrery/scene/limits.h),
  errors as "path: message", output in the layout of an indented ordered JSON dump with
  shortest round-trip numbers;
- the typed aero record: AeroExperiment and its parts, to_experiment / aero_from_experiment /
  parse_aero_experiment / save_aero_experiment (src/scene/aero_experiment.cpp), with the
  reference's validation messages;
- comparisons: geometry_label, compare_by_geometry (deltas against the lowest angle of attack and
  the least-squares lift slope) and comparison_csv (src/app/aero/aero_comparison.cpp:62-187).

JSON syntax errors carry Python's parser message where the reference carries its JSON library's.
"""
from __future__ import annotations

import dataclasses
import enum
import json
import math
from typing import Optional, Union

SCHEMA_VERSION = 1
AERO_SOLVER_VERSION = "1"
MAX_DOCUMENT_BYTES = 16 * 1024 * 1024
MAX_NESTING_DEPTH = 64
MAX_CONTAINER_ELEMENTS = 200_000
MAX_JSON_VALUES = 2_000_000
DOMAINS = ("space", "aero", "entry")
MAX_STEPS_PER_FRAME = 1.0e6


# ---- errors (src/scene/include/orrery/scene/error.h) ----------------------------------------

class SceneErrorCode(enum.Enum):
    io_error = "io error"
    syntax_error = "syntax error"
    duplicate_key = "duplicate key"
    unknown_key = "unknown key"
    missing_key = "missing key"
    wrong_type = "wrong type"
    unsupported_version = "unsupported version"
    invalid_value = "invalid value"
    limit_exceeded = "limit exceeded"


class SceneError(ValueError):
    """A rejected generic record: str() is "path: message", or the message alone."""

    def __init__(self, code: SceneErrorCode, path: str, message: str):
        self.code = code
        self.path = path
        self.message = message
        super().__init__(f"{path}: {message}" if path else message)


class AeroRecordError(ValueError):
    """A rejected aero record; str() names the offending key."""


def child_path(parent: str, key: str) -> str:
    return f"{parent}.{key}" if parent else key


def index_path(parent: str, index: int) -> str:
    return f"{parent}[{index}]"


# ---- strict JSON reading (json_support.cpp) ------------------------------------------------

class _Object:
    """A JSON object as written: its members in document order (duplicates kept for the check)."""

    def __init__(self, pairs):
        self.pairs = pairs

    def get(self, key):
        for k, v in self.pairs:
            if k == key:
                return v
        return None

    def has(self, key) -> bool:
        return any(k == key for k, _ in self.pairs)


class _Underflow(float):
    """A nonzero literal that reads as zero."""
    lexeme = ""


def _parse_float(lexeme: str):
    value = float(lexeme)
    if value == 0.0:
        mantissa = lexeme.lower().split("e")[0]
        if any(ch in "123456789" for ch in mantissa):
            u = _Underflow(value)
            u.lexeme = lexeme
            return u
    return value


def _reject_constant(name):
    raise ValueError(f"{name} is not JSON")


def type_name(value) -> str:
    if value is None:
        return "null"
    if isinstance(value, _Object):
        return "object"
    if isinstance(value, list):
        return "array"
    if isinstance(value, str):
        return "string"
    if isinstance(value, bool):
        return "boolean"
    return "number"


def _as_double(value) -> float:
    """A JSON number as a double (integer literals beyond the double range read as infinite)."""
    try:
        return float(value)
    except OverflowError:
        return math.inf if value > 0 else -math.inf


def _is_number(value) -> bool:
    return isinstance(value, (int, float)) and not isinstance(value, bool)


def _strict_check(root) -> None:
    """Underflow, duplicate keys (the first in document order) and the document limits, with the
    path of the offending value."""
    duplicate = None
    count = 0
    stack = [(root, "", 0)]
    while stack:
        value, path, depth = stack.pop()
        count += 1
        if count > MAX_JSON_VALUES:
            raise SceneError(SceneErrorCode.limit_exceeded, path,
                             f"document has more than {MAX_JSON_VALUES} values")
        if isinstance(value, _Underflow):
            shown = value.lexeme if len(value.lexeme) <= 40 else value.lexeme[:40] + "..."
            raise SceneError(SceneErrorCode.invalid_value, path, f"number {shown} is not zero but underflows to zero")
        if isinstance(value, (_Object, list)):
            if depth >= MAX_NESTING_DEPTH:
                raise SceneError(SceneErrorCode.limit_exceeded, path,
                                 f"nesting is deeper than {MAX_NESTING_DEPTH} levels")
            if isinstance(value, _Object):
                if len(value.pairs) > MAX_CONTAINER_ELEMENTS:
                    raise SceneError(SceneErrorCode.limit_exceeded, path,
                                     f"object has more than {MAX_CONTAINER_ELEMENTS} members")
                seen = set()
                children = []
                for k, v in value.pairs:
                    p = child_path(path, k)
                    if k in seen and duplicate is None:
                        duplicate = SceneError(SceneErrorCode.duplicate_key, p, f"key '{k}' appears more than once")
                    seen.add(k)
                    children.append((v, p, depth + 1))
            else:
                if len(value) > MAX_CONTAINER_ELEMENTS:
                    raise SceneError(SceneErrorCode.limit_exceeded, path,
                                     f"array has more than {MAX_CONTAINER_ELEMENTS} elements")
                children = [(v, index_path(path, i), depth + 1) for i, v in enumerate(value)]
            stack.extend(reversed(children))
    if duplicate is not None:
        raise duplicate


def _parse_document(text: str):
    if len(text.encode("utf-8")) > MAX_DOCUMENT_BYTES:
        raise SceneError(SceneErrorCode.limit_exceeded, "", f"document is larger than {MAX_DOCUMENT_BYTES} bytes")
    try:
        root = json.loads(text, object_pairs_hook=_Object, parse_float=_parse_float, parse_constant=_reject_constant)
    except (ValueError, RecursionError) as e:
        raise SceneError(SceneErrorCode.syntax_error, "", f"syntax error: {e}") from None
    _strict_check(root)
    return root


class _Reader:
    """ObjectReader (json_support.h:70-99)."""

    def __init__(self, value, path: str):
        if not isinstance(value, _Object):
            raise SceneError(SceneErrorCode.wrong_type, path, f"expected object, found {type_name(value)}")
        self.value = value
        self.path = path

    def path_of(self, key: str) -> str:
        return child_path(self.path, key)

    def has(self, key: str) -> bool:
        return self.value.has(key)

    def allow_only(self, allowed) -> None:
        for k, _ in self.value.pairs:
            if k not in allowed:
                raise SceneError(SceneErrorCode.unknown_key, self.path_of(k),
                                 f"unknown key '{k}' (allowed: {', '.join(allowed)})")

    def required(self, key: str):
        if not self.has(key):
            raise SceneError(SceneErrorCode.missing_key, self.path_of(key), f"required key '{key}' is missing")
        return self.value.get(key)

    def required_string(self, key: str) -> str:
        v = self.required(key)
        if not isinstance(v, str):
            raise SceneError(SceneErrorCode.wrong_type, self.path_of(key), f"expected string, found {type_name(v)}")
        return v

    def required_integer(self, key: str) -> int:
        v = self.required(key)
        if not (isinstance(v, int) and not isinstance(v, bool)):
            raise SceneError(SceneErrorCode.wrong_type, self.path_of(key), f"expected integer, found {type_name(v)}")
        return min(v, 0x7FFFFFFFFFFFFFFF)

    def optional_string(self, key: str) -> Optional[str]:
        return self.required_string(key) if self.has(key) else None


# ---- the generic record (experiment.h, experiment.cpp) -------------------------------------

ParameterValue = Union[float, str]


@dataclasses.dataclass
class SolverInfo:
    name: str = ""
    version: str = ""
    settings: dict = dataclasses.field(default_factory=dict)


@dataclasses.dataclass
class Experiment:
    domain: str = "space"
    inputs: dict = dataclasses.field(default_factory=dict)
    conditions: dict = dataclasses.field(default_factory=dict)
    solver: SolverInfo = dataclasses.field(default_factory=SolverInfo)
    result: dict = dataclasses.field(default_factory=dict)
    series: dict = dataclasses.field(default_factory=dict)


def _read_parameters(parent: _Reader, key: str) -> dict:
    out = {}
    if not parent.has(key):
        return out
    obj = _Reader(parent.required(key), parent.path_of(key))
    for k, v in obj.value.pairs:
        path = obj.path_of(k)
        if isinstance(v, str):
            out[k] = v
        elif _is_number(v):
            number = _as_double(v)
            if not math.isfinite(number):
                raise SceneError(SceneErrorCode.invalid_value, path, "value must be finite")
            out[k] = number
        else:
            raise SceneError(SceneErrorCode.wrong_type, path, f"expected number or string, found {type_name(v)}")
    return out


def _read_series(parent: _Reader, key: str) -> dict:
    out = {}
    if not parent.has(key):
        return out
    obj = _Reader(parent.required(key), parent.path_of(key))
    for k, values in obj.value.pairs:
        path = obj.path_of(k)
        if not isinstance(values, list):
            raise SceneError(SceneErrorCode.wrong_type, path, f"expected an array of numbers, found {type_name(values)}")
        series = []
        for i, v in enumerate(values):
            element = index_path(path, i)
            if not _is_number(v):
                raise SceneError(SceneErrorCode.wrong_type, element, f"expected number, found {type_name(v)}")
            number = _as_double(v)
            if not math.isfinite(number):
                raise SceneError(SceneErrorCode.invalid_value, element, "value must be finite")
            series.append(number)
        out[k] = series
    return out


def parse_experiment(text: str) -> Experiment:
    """parse_experiment (experiment.cpp:88-126); raises SceneError."""
    root = _Reader(_parse_document(text), "")
    version = root.required_integer("version")
    if version != SCHEMA_VERSION:
        raise SceneError(SceneErrorCode.unsupported_version, root.path_of("version"),
                         f"experiment schema version {version} is not supported (supported: {SCHEMA_VERSION})")
    root.allow_only(("version", "domain", "inputs", "conditions", "solver", "result", "series"))
    e = Experiment()
    domain = root.required_string("domain")
    if domain not in DOMAINS:
        raise SceneError(SceneErrorCode.invalid_value, root.path_of("domain"),
                         f"unsupported domain '{domain}' (allowed: space, aero, entry)")
    e.domain = domain
    e.inputs = _read_parameters(root, "inputs")
    e.conditions = _read_parameters(root, "conditions")
    solver = _Reader(root.required("solver"), root.path_of("solver"))
    solver.allow_only(("name", "version", "settings"))
    e.solver.name = solver.required_string("name")
    e.solver.version = solver.optional_string("version") or ""
    e.solver.settings = _read_parameters(solver, "settings")
    e.result = _read_parameters(root, "result")
    e.series = _read_series(root, "series")
    return e


# ---- writing: an ordered JSON dump with indent 2 ------------------------------------------

def _shortest_digits(value: float) -> tuple[str, int]:
    """Shortest round-trip significant digits and the decimal exponent of the first digit."""
    r = repr(abs(value))
    if "e" in r:
        m, _, ex = r.partition("e")
        digits = m.replace(".", "")
        return digits.rstrip("0") or "0", int(ex)
    whole, _, frac = r.partition(".")
    frac = frac.rstrip("0")
    if whole != "0":
        digits = (whole + frac).rstrip("0")
        return digits, len(whole) - 1
    stripped = frac.lstrip("0")
    return stripped.rstrip("0"), -(len(frac) - len(stripped)) - 1


def _exponent(e: int) -> str:
    sign = "-" if e < 0 else "+"
    e = abs(e)
    return f"{sign}{e:02d}" if e < 100 else f"{sign}{e}"


def format_json_number(value: float) -> str:
    """A double as the ordered JSON dump writes it: shortest round-trip digits, fixed notation
    for decimal exponents in [-4, 15) (integral values keep a ".0"), scientific otherwise;
    non-finite values become null."""
    if not math.isfinite(value):
        return "null"
    if value == 0.0:
        return "-0.0" if math.copysign(1.0, value) < 0 else "0.0"
    digits, e10 = _shortest_digits(value)
    sign = "-" if value < 0 else ""
    k = len(digits)
    n = e10 + 1  # position of the decimal point after the first n digits
    if k <= n <= 15:
        return sign + digits + "0" * (n - k) + ".0"
    if 0 < n <= 15:
        return sign + digits[:n] + "." + digits[n:]
    if -4 < n <= 0:
        return sign + "0." + "0" * (-n) + digits
    body = digits if k == 1 else digits[0] + "." + digits[1:]
    return sign + body + "e" + _exponent(n - 1)


def _json_string(text: str) -> str:
    out = ['"']
    for ch in text:
        o = ord(ch)
        if ch == '"':
            out.append('\\"')
        elif ch == "\\":
            out.append("\\\\")
        elif ch == "\b":
            out.append("\\b")
        elif ch == "\f":
            out.append("\\f")
        elif ch == "\n":
            out.append("\\n")
        elif ch == "\r":
            out.append("\\r")
        elif ch == "\t":
            out.append("\\t")
        elif o < 0x20:
            out.append(f"\\u{o:04x}")
        else:
            out.append(ch)
    out.append('"')
    return "".join(out)


def _dump(value, indent: int) -> str:
    pad = " " * (indent + 2)
    if isinstance(value, dict):
        if not value:
            return "{}"
        items = [f"{pad}{_json_string(k)}: {_dump(v, indent + 2)}" for k, v in value.items()]
        return "{\n" + ",\n".join(items) + "\n" + " " * indent + "}"
    if isinstance(value, list):
        if not value:
            return "[]"
        return "[\n" + ",\n".join(pad + _dump(v, indent + 2) for v in value) + "\n" + " " * indent + "]"
    if isinstance(value, str):
        return _json_string(value)
    if isinstance(value, bool):
        return "true" if value else "false"
    if isinstance(value, int):
        return str(value)
    return format_json_number(float(value))


def _parameters(p: dict) -> dict:
    return {k: (v if isinstance(v, str) else float(v)) for k, v in sorted(p.items())}


def save_experiment(e: Experiment) -> str:
    """save_experiment (experiment.cpp:158-179): keys in schema order, parameters sorted."""
    root = {"version": SCHEMA_VERSION, "domain": e.domain, "inputs": _parameters(e.inputs),
            "conditions": _parameters(e.conditions),
            "solver": {"name": e.solver.name, "version": e.solver.version,
                       "settings": _parameters(e.solver.settings)},
            "result": _parameters(e.result)}
    if e.series:
        root["series"] = {k: [float(x) for x in v] for k, v in sorted(e.series.items())}
    return _dump(root, 0) + "\n"


# ---- the typed aero record (aero_experiment.h, aero_experiment.cpp) ------------------------

class AeroShape(enum.Enum):
    sphere = "sphere"
    cube = "cube"
    cylinder = "cylinder"
    cone = "cone"
    ellipsoid = "ellipsoid"
    flat_plate = "flat_plate"
    wedge = "wedge"
    capsule = "capsule"
    naca_wing = "naca_wing"
    naca_section = "naca_section"


SHAPE_DISPLAY = {AeroShape.sphere: "Sphere", AeroShape.cube: "Cube", AeroShape.cylinder: "Cylinder (cross-flow)",
                 AeroShape.cone: "Cone", AeroShape.ellipsoid: "Ellipsoid", AeroShape.flat_plate: "Flat plate",
                 AeroShape.wedge: "Wedge", AeroShape.capsule: "Entry capsule", AeroShape.naca_wing: "NACA wing",
                 AeroShape.naca_section: "NACA section (2D)"}

DIMENSION_KEYS = ("diameter_m", "length_m", "width_m", "height_m", "side_m", "chord_m", "span_m", "thickness_m",
                  "half_angle_deg")


def shape_dimensions(shape: AeroShape) -> tuple[str, ...]:
    """aero_experiment.cpp:557-579"""
    return {AeroShape.sphere: ("diameter_m",), AeroShape.capsule: ("diameter_m",), AeroShape.cube: ("side_m",),
            AeroShape.cylinder: ("diameter_m", "length_m"), AeroShape.cone: ("diameter_m", "length_m"),
            AeroShape.ellipsoid: ("length_m", "width_m", "height_m"),
            AeroShape.flat_plate: ("chord_m", "span_m", "thickness_m"),
            AeroShape.wedge: ("length_m", "half_angle_deg", "span_m"), AeroShape.naca_wing: ("chord_m", "span_m"),
            AeroShape.naca_section: ("chord_m",)}[shape]


class GeometrySource(enum.Enum):
    primitive = "primitive"
    file = "file"
    url = "url"


class SolverKind(enum.Enum):
    """aero::SolverKind with its solver_name (src/aero/regime.cpp:57-67)."""
    panel = "panel"
    newtonian = "newtonian"
    lbm = "lbm"


class ReferenceAreaKind(enum.Enum):
    frontal = "frontal"
    planform = "planform"
    wetted = "wetted"


@dataclasses.dataclass
class ShapeDimensions:
    diameter_m: float = 1.0
    length_m: float = 2.0
    width_m: float = 1.0
    height_m: float = 1.0
    side_m: float = 1.0
    chord_m: float = 1.0
    span_m: float = 4.0
    thickness_m: float = 0.02
    half_angle_deg: float = 10.0


@dataclasses.dataclass
class GeometryRef:
    source: GeometrySource = GeometrySource.primitive
    shape: AeroShape = AeroShape.sphere
    dimensions: ShapeDimensions = dataclasses.field(default_factory=ShapeDimensions)
    naca: str = "0012"
    capsule: str = "apollo"
    path: str = ""
    url: str = ""
    sha256: str = ""
    license_note: str = ""
    unit_scale: float = 1.0
    up_axis: str = "z"

    @property
    def user_model(self) -> bool:
        return self.source != GeometrySource.primitive


@dataclasses.dataclass
class AeroConditions:
    speed_m_s: float = 30.0
    density_kg_m3: float = 1.225
    viscosity_pa_s: float = 1.789e-5
    temperature_k: float = 288.15
    angle_of_attack_deg: float = 0.0
    sideslip_deg: float = 0.0
    gamma: float = 1.4
    gas_constant_j_kg_k: float = 287.05

    def flow(self):
        """The solver's view (angles in radians)."""
        from .aero import FlowConditions
        degree = math.pi / 180.0
        return FlowConditions(speed=self.speed_m_s, density=self.density_kg_m3, viscosity=self.viscosity_pa_s,
                              temperature=self.temperature_k, angle_of_attack=self.angle_of_attack_deg * degree,
                              sideslip=self.sideslip_deg * degree, gamma=self.gamma,
                              gas_constant=self.gas_constant_j_kg_k)


@dataclasses.dataclass
class AeroSolverSettings:
    solver: Optional[SolverKind] = None  # None: "auto"
    reference_area: ReferenceAreaKind = ReferenceAreaKind.frontal
    reference_length_m: Optional[float] = None
    lbm_cells_per_length: float = 0.0
    lbm_duration: float = 0.0
    steps_per_frame: int = 100


@dataclasses.dataclass
class AeroResultRecord:
    solver: SolverKind = SolverKind.panel
    cd: float = 0.0
    cl: float = 0.0
    cy: float = 0.0
    cm: float = 0.0
    cd_pressure: float = 0.0
    cd_induced: float = 0.0
    cd_friction: float = 0.0
    reynolds: float = 0.0
    mach: float = 0.0
    reference_area_m2: float = 0.0
    reference_length_m: float = 0.0
    iterations: int = 0
    converged: bool = False
    strouhal: Optional[float] = None
    warnings: list = dataclasses.field(default_factory=list)
    residuals: list = dataclasses.field(default_factory=list)
    history_step: list = dataclasses.field(default_factory=list)
    history_cd: list = dataclasses.field(default_factory=list)
    history_cl: list = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class AeroExperiment:
    name: str = ""
    geometry: GeometryRef = dataclasses.field(default_factory=GeometryRef)
    conditions: AeroConditions = dataclasses.field(default_factory=AeroConditions)
    solver: AeroSolverSettings = dataclasses.field(default_factory=AeroSolverSettings)
    result: Optional[AeroResultRecord] = None


def naca_designation_ok(digits: str) -> bool:
    """Naca4::from_designation (src/geom/naca.cpp:85-105)."""
    if len(digits) != 4 or any(c < "0" or c > "9" for c in digits):
        return False
    m, p, t = int(digits[0]), int(digits[1]), int(digits[2:])
    return t != 0 and (m == 0) == (p == 0)


def _join(names) -> str:
    return ", ".join(names) if names else "none"


class _Params:
    """The Reader of aero_experiment.cpp:93-186 over one parameter map."""

    def __init__(self, parameters: dict, where: str):
        self.p = parameters
        self.where = where

    def path(self, key: str) -> str:
        return f"{self.where}.{key}"

    def unknown_key(self, allowed) -> Optional[str]:
        for k in sorted(self.p):
            if k not in allowed:
                return f"{self.path(k)}: unknown key (allowed: {_join(allowed)})"
        return None

    def has(self, key: str) -> bool:
        return key in self.p

    def number(self, key: str) -> Optional[float]:
        if key not in self.p:
            return None
        v = self.p[key]
        if isinstance(v, str):
            raise AeroRecordError(f"{self.path(key)}: expected a number")
        return float(v)

    def text(self, key: str) -> Optional[str]:
        if key not in self.p:
            return None
        v = self.p[key]
        if not isinstance(v, str):
            raise AeroRecordError(f"{self.path(key)}: expected a string")
        return v

    def checked(self, key: str, fallback: float, valid, rule: str) -> float:
        given = self.number(key)
        if given is None:
            return fallback
        if not valid(given):
            raise AeroRecordError(f"{self.path(key)}: {rule}")
        return given

    def positive(self, key: str, fallback: float) -> float:
        return self.checked(key, fallback, lambda v: v > 0.0 and math.isfinite(v), "must be positive and finite")

    def text_or(self, key: str, fallback: str) -> str:
        v = self.text(key)
        return fallback if v is None else v


def _read_geometry(inputs: dict) -> GeometryRef:
    """aero_experiment.cpp:188-307"""
    r = _Params(inputs, "inputs")
    ref = GeometryRef()
    source = r.text_or("geometry", "primitive")
    if source not in ("primitive", "file", "url"):
        raise AeroRecordError(f"{r.path('geometry')}: unknown geometry source '{source}' (allowed: primitive, file, url)")
    ref.source = GeometrySource(source)
    allowed = ["name", "geometry"]
    if ref.source == GeometrySource.primitive:
        name = r.text_or("shape", "sphere")
        try:
            ref.shape = AeroShape(name)
        except ValueError:
            raise AeroRecordError(f"{r.path('shape')}: unknown shape '{name}' "
                                  f"(allowed: {_join([s.value for s in AeroShape])})") from None
        allowed.append("shape")
        for key in shape_dimensions(ref.shape):
            allowed.append(key)
            setattr(ref.dimensions, key, r.positive(key, getattr(ref.dimensions, key)))
        if ref.dimensions.half_angle_deg >= 90.0:
            raise AeroRecordError(f"{r.path('half_angle_deg')}: must be below 90 degrees")
        if ref.shape in (AeroShape.naca_wing, AeroShape.naca_section):
            allowed.append("naca")
            naca = r.text_or("naca", ref.naca)
            if not naca_designation_ok(naca):
                raise AeroRecordError(f"{r.path('naca')}: '{naca}' is not a NACA 4-digit designation")
            ref.naca = naca
        if ref.shape == AeroShape.capsule:
            allowed.append("capsule")
            capsule = r.text_or("capsule", ref.capsule)
            if capsule not in ("apollo", "orion"):
                raise AeroRecordError(f"{r.path('capsule')}: unknown capsule '{capsule}' (allowed: apollo, orion)")
            ref.capsule = capsule
    else:
        allowed += ["path", "unit_scale", "up_axis", "license"]
        if ref.source == GeometrySource.url:
            allowed += ["url", "sha256"]
        path = r.text_or("path", "")
        url = r.text_or("url", "")
        sha = r.text_or("sha256", "")
        lic = r.text_or("license", "")
        up = r.text_or("up_axis", "z")
        scale = r.positive("unit_scale", 1.0)
        ref.path, ref.url, ref.sha256, ref.license_note, ref.unit_scale = path, url, sha, lic, scale
        if up not in ("z", "y"):
            raise AeroRecordError(f"{r.path('up_axis')}: unknown up axis '{up}' (allowed: z, y)")
        ref.up_axis = up
        if not ref.path and not ref.url:
            raise AeroRecordError(f"{r.path('path')}: a {source} geometry needs a model path")
        if ref.source == GeometrySource.url and not ref.url:
            raise AeroRecordError(f"{r.path('url')}: a url geometry needs its source URL")
    problem = r.unknown_key(allowed)
    if problem:
        raise AeroRecordError(problem)
    return ref


def _read_conditions(parameters: dict) -> AeroConditions:
    """aero_experiment.cpp:309-345"""
    r = _Params(parameters, "conditions")
    keys = ("speed_m_s", "density_kg_m3", "viscosity_pa_s", "temperature_k", "angle_of_attack_deg", "sideslip_deg",
            "gamma", "gas_constant_j_kg_k")
    problem = r.unknown_key(keys)
    if problem:
        raise AeroRecordError(problem)
    c = AeroConditions()
    finite = math.isfinite
    return AeroConditions(
        speed_m_s=r.positive("speed_m_s", c.speed_m_s),
        density_kg_m3=r.positive("density_kg_m3", c.density_kg_m3),
        viscosity_pa_s=r.positive("viscosity_pa_s", c.viscosity_pa_s),
        temperature_k=r.positive("temperature_k", c.temperature_k),
        angle_of_attack_deg=r.checked("angle_of_attack_deg", c.angle_of_attack_deg, finite, "must be finite"),
        sideslip_deg=r.checked("sideslip_deg", c.sideslip_deg, finite, "must be finite"),
        gamma=r.checked("gamma", c.gamma, lambda v: v > 1.0 and math.isfinite(v), "must be above 1 and finite"),
        gas_constant_j_kg_k=r.positive("gas_constant_j_kg_k", c.gas_constant_j_kg_k))


def _read_solver(info: SolverInfo) -> AeroSolverSettings:
    """aero_experiment.cpp:347-398"""
    s = AeroSolverSettings()
    if info.name != "auto":
        try:
            s.solver = SolverKind(info.name)
        except ValueError:
            raise AeroRecordError(f"solver.name: unknown solver '{info.name}' "
                                  "(allowed: auto, panel, newtonian, lbm)") from None
    if info.version and info.version != AERO_SOLVER_VERSION:
        raise AeroRecordError(f"solver.version: '{info.version}' is not supported "
                              f"(supported: {AERO_SOLVER_VERSION})")
    r = _Params(info.settings, "solver.settings")
    problem = r.unknown_key(("reference_area", "reference_length_m", "lbm_cells_per_length", "lbm_duration",
                             "steps_per_frame"))
    if problem:
        raise AeroRecordError(problem)
    area = r.text_or("reference_area", "frontal")
    try:
        s.reference_area = ReferenceAreaKind(area)
    except ValueError:
        raise AeroRecordError(f"{r.path('reference_area')}: unknown reference area '{area}' "
                              "(allowed: frontal, planform, wetted)") from None
    if r.has("reference_length_m"):
        s.reference_length_m = r.positive("reference_length_m", 1.0)
    cells = r.positive("lbm_cells_per_length", 1.0)
    duration = r.positive("lbm_duration", 1.0)
    steps = r.checked("steps_per_frame", float(s.steps_per_frame),
                      lambda v: 1.0 <= v <= MAX_STEPS_PER_FRAME and math.floor(v) == v,
                      "must be an integer in [1, 1000000]")
    s.lbm_cells_per_length = cells if r.has("lbm_cells_per_length") else 0.0
    s.lbm_duration = duration if r.has("lbm_duration") else 0.0
    s.steps_per_frame = int(steps)
    return s


def _split_codes(text: str) -> list[str]:
    out = []
    for part in text.split(";"):
        code = part.strip(" ")
        if code:
            out.append(code)
    return out


def _read_result(e: Experiment) -> Optional[AeroResultRecord]:
    """aero_experiment.cpp:421-531"""
    if not e.result and not e.series:
        return None
    r = _Params(e.result, "result")
    keys = ("solver", "cd", "cl", "cy", "cm", "cd_pressure", "cd_induced", "cd_friction", "reynolds", "mach",
            "reference_area_m2", "reference_length_m", "iterations", "converged", "warnings", "strouhal")
    problem = r.unknown_key(keys)
    if problem:
        raise AeroRecordError(problem)
    out = AeroResultRecord()
    solver = r.text("solver")
    if solver is None:
        raise AeroRecordError(f"{r.path('solver')}: a result names the solver that produced it")
    try:
        out.solver = SolverKind(solver)
    except ValueError:
        raise AeroRecordError(f"{r.path('solver')}: unknown solver '{solver}' (allowed: panel, newtonian, lbm)") from None
    for key in ("cd", "cl", "cy", "cm", "cd_pressure", "cd_induced", "cd_friction", "reynolds", "mach",
                "reference_area_m2", "reference_length_m"):
        setattr(out, key, r.checked(key, 0.0, lambda v: True, ""))
    iterations = r.checked("iterations", 0.0, lambda v: 0.0 <= v < 9.0e15 and math.floor(v) == v,
                           "must be a non-negative integer")
    converged = r.checked("converged", 0.0, lambda v: v in (0.0, 1.0), "must be 0 or 1")
    warnings = r.text_or("warnings", "")
    out.iterations = int(iterations)
    out.converged = converged == 1.0
    out.warnings = _split_codes(warnings)
    if r.has("strouhal"):
        out.strouhal = r.positive("strouhal", 1.0)
    allowed = ("residual", "history_step", "history_cd", "history_cl")
    for k in sorted(e.series):
        if k not in allowed:
            raise AeroRecordError(f"series.{k}: unknown key (allowed: {_join(allowed)})")
    out.residuals = list(e.series.get("residual", []))
    out.history_step = list(e.series.get("history_step", []))
    out.history_cd = list(e.series.get("history_cd", []))
    out.history_cl = list(e.series.get("history_cl", []))
    if len(out.history_cd) != len(out.history_step) or len(out.history_cl) != len(out.history_step):
        raise AeroRecordError("series: history_step, history_cd and history_cl must have the same length")
    return out


def to_experiment(a: AeroExperiment) -> Experiment:
    """aero_experiment.cpp:621-712"""
    out = Experiment(domain="aero")
    g = a.geometry
    if a.name:
        out.inputs["name"] = a.name
    if g.source == GeometrySource.primitive:
        out.inputs["geometry"] = "primitive"
        out.inputs["shape"] = g.shape.value
        for key in shape_dimensions(g.shape):
            out.inputs[key] = float(getattr(g.dimensions, key))
        if g.shape in (AeroShape.naca_wing, AeroShape.naca_section):
            out.inputs["naca"] = g.naca
        if g.shape == AeroShape.capsule:
            out.inputs["capsule"] = g.capsule
    else:
        out.inputs["geometry"] = g.source.value
        out.inputs["path"] = g.path
        out.inputs["unit_scale"] = float(g.unit_scale)
        out.inputs["up_axis"] = g.up_axis
        if g.license_note:
            out.inputs["license"] = g.license_note
        if g.source == GeometrySource.url:
            out.inputs["url"] = g.url
            if g.sha256:
                out.inputs["sha256"] = g.sha256
    c = a.conditions
    out.conditions = {"speed_m_s": c.speed_m_s, "density_kg_m3": c.density_kg_m3,
                      "viscosity_pa_s": c.viscosity_pa_s, "temperature_k": c.temperature_k,
                      "angle_of_attack_deg": c.angle_of_attack_deg, "sideslip_deg": c.sideslip_deg,
                      "gamma": c.gamma, "gas_constant_j_kg_k": c.gas_constant_j_kg_k}
    s = a.solver
    out.solver = SolverInfo(s.solver.value if s.solver else "auto", AERO_SOLVER_VERSION,
                            {"reference_area": s.reference_area.value})
    if s.reference_length_m is not None:
        out.solver.settings["reference_length_m"] = float(s.reference_length_m)
    if s.lbm_cells_per_length > 0.0:
        out.solver.settings["lbm_cells_per_length"] = float(s.lbm_cells_per_length)
    if s.lbm_duration > 0.0:
        out.solver.settings["lbm_duration"] = float(s.lbm_duration)
    out.solver.settings["steps_per_frame"] = float(s.steps_per_frame)
    if a.result is not None:
        r = a.result
        out.result = {"solver": r.solver.value, "cd": r.cd, "cl": r.cl, "cy": r.cy, "cm": r.cm,
                      "cd_pressure": r.cd_pressure, "cd_induced": r.cd_induced, "cd_friction": r.cd_friction,
                      "reynolds": r.reynolds, "mach": r.mach, "reference_area_m2": r.reference_area_m2,
                      "reference_length_m": r.reference_length_m, "iterations": float(r.iterations),
                      "converged": 1.0 if r.converged else 0.0}
        if r.strouhal is not None:
            out.result["strouhal"] = r.strouhal
        if r.warnings:
            out.result["warnings"] = "; ".join(r.warnings)
        for key, values in (("residual", r.residuals), ("history_step", r.history_step),
                            ("history_cd", r.history_cd), ("history_cl", r.history_cl)):
            if values:
                out.series[key] = [float(v) for v in values]
    return out


def aero_from_experiment(e: Experiment) -> AeroExperiment:
    """aero_experiment.cpp:714-749; raises AeroRecordError."""
    if e.domain != "aero":
        raise AeroRecordError(f"domain: expected 'aero', found '{e.domain}'")
    out = AeroExperiment()
    out.name = _Params(e.inputs, "inputs").text_or("name", "")
    out.geometry = _read_geometry(e.inputs)
    out.conditions = _read_conditions(e.conditions)
    out.solver = _read_solver(e.solver)
    out.result = _read_result(e)
    return out


def parse_aero_experiment(text: str) -> AeroExperiment:
    """parse_aero_experiment: a generic error becomes its "path: message" text."""
    try:
        e = parse_experiment(text)
    except SceneError as err:
        raise AeroRecordError(str(err)) from None
    return aero_from_experiment(e)


def save_aero_experiment(a: AeroExperiment) -> str:
    return save_experiment(to_experiment(a))


# ---- comparisons (src/app/aero/aero_comparison.cpp:24-170) ---------------------------------

def exact(value: float) -> str:
    """The shortest round-trip decimal text of a double, in the plain style of std::to_chars:
    fixed or scientific notation, whichever is shorter (fixed on a tie)."""
    if value == 0.0:
        return "-0" if math.copysign(1.0, value) < 0 else "0"
    if not math.isfinite(value):
        return "nan" if math.isnan(value) else ("-inf" if value < 0 else "inf")
    digits, e10 = _shortest_digits(value)
    sign = "-" if value < 0 else ""
    sci = (digits[0] + ("." + digits[1:] if len(digits) > 1 else "") + "e" + _exponent(e10))
    if e10 >= 0:
        fixed = digits + "0" * (e10 + 1 - len(digits)) if len(digits) <= e10 + 1 else \
            digits[:e10 + 1] + "." + digits[e10 + 1:]
    else:
        fixed = "0." + "0" * (-e10 - 1) + digits
    return sign + (fixed if len(fixed) <= len(sci) else sci)


def _trimmed(value: float) -> str:
    return f"{value:.3g}"


def _formula_like(text: str) -> bool:
    return bool(text) and text[0] in "=+-@\t\r"


def csv_field(text: str) -> str:
    """RFC 4180 quoting; a cell that would start a formula is quoted behind an apostrophe."""
    formula = _formula_like(text)
    if not formula and not any(c in text for c in ',"\r\n'):
        return text
    return ('"\'' if formula else '"') + text.replace('"', '""') + '"'


def geometry_label(ref: GeometryRef) -> str:
    """aero_comparison.cpp:62-95"""
    if ref.source != GeometrySource.primitive:
        name = ref.path.replace("\\", "/").rsplit("/", 1)[-1] if ref.path else ""
        if name:
            return name
        return ref.url if ref.url else "User model"
    d = ref.dimensions

    def m(v):
        return _trimmed(v) + " m"

    s = ref.shape
    if s == AeroShape.sphere:
        return "Sphere, d " + m(d.diameter_m)
    if s == AeroShape.cube:
        return "Cube, side " + m(d.side_m)
    if s == AeroShape.cylinder:
        return "Cylinder, d " + m(d.diameter_m) + ", span " + m(d.length_m)
    if s == AeroShape.cone:
        return "Cone, d " + m(d.diameter_m) + ", length " + m(d.length_m)
    if s == AeroShape.ellipsoid:
        return "Ellipsoid " + m(d.length_m) + " x " + m(d.width_m) + " x " + m(d.height_m)
    if s == AeroShape.flat_plate:
        return "Flat plate " + m(d.chord_m) + " x " + m(d.span_m)
    if s == AeroShape.wedge:
        return "Wedge " + _trimmed(d.half_angle_deg) + " deg, " + m(d.length_m)
    if s == AeroShape.capsule:
        return ("Orion-type capsule, d " if ref.capsule == "orion" else "Apollo-type capsule, d ") + m(d.diameter_m)
    if s == AeroShape.naca_wing:
        return "NACA " + ref.naca + " wing, chord " + m(d.chord_m) + ", span " + m(d.span_m)
    return "NACA " + ref.naca + " section, chord " + m(d.chord_m)


@dataclasses.dataclass
class ComparisonRow:
    name: str
    angle_of_attack_deg: float
    sideslip_deg: float
    speed_m_s: float
    solver: SolverKind
    cd: float
    cl: float
    cm: float
    lift_to_drag: float
    converged: bool
    delta_cd: float = 0.0
    delta_cl: float = 0.0
    delta_cm: float = 0.0


@dataclasses.dataclass
class ComparisonTable:
    geometry: GeometryRef
    label: str
    rows: list = dataclasses.field(default_factory=list)
    lift_slope_per_deg: Optional[float] = None
    lift_slope_per_rad: Optional[float] = None


def compare_by_geometry(experiments) -> list[ComparisonTable]:
    """compare_by_geometry (aero_comparison.cpp:97-164): finished records grouped by geometry,
    rows by angle of attack (then name), deltas against the first row, and the least-squares
    lift slope over all rows."""
    tables: list[ComparisonTable] = []
    for e in experiments:
        if e.result is None:
            continue
        table = next((t for t in tables if t.geometry == e.geometry), None)
        if table is None:
            table = ComparisonTable(e.geometry, geometry_label(e.geometry))
            tables.append(table)
        r = e.result
        table.rows.append(ComparisonRow(e.name, e.conditions.angle_of_attack_deg, e.conditions.sideslip_deg,
                                        e.conditions.speed_m_s, r.solver, r.cd, r.cl, r.cm,
                                        r.cl / r.cd if r.cd != 0.0 else 0.0, r.converged))
    for table in tables:
        table.rows.sort(key=lambda row: (row.angle_of_attack_deg, row.name))
        base = table.rows[0]
        mean_a = 0.0
        mean_cl = 0.0
        for row in table.rows:
            row.delta_cd = row.cd - base.cd
            row.delta_cl = row.cl - base.cl
            row.delta_cm = row.cm - base.cm
            mean_a += row.angle_of_attack_deg
            mean_cl += row.cl
        n = float(len(table.rows))
        mean_a /= n
        mean_cl /= n
        sxx = 0.0
        sxy = 0.0
        for row in table.rows:
            sxx += (row.angle_of_attack_deg - mean_a) * (row.angle_of_attack_deg - mean_a)
            sxy += (row.angle_of_attack_deg - mean_a) * (row.cl - mean_cl)
        if len(table.rows) >= 2 and sxx > 0.0:
            table.lift_slope_per_deg = sxy / sxx
            table.lift_slope_per_rad = sxy / sxx / (math.pi / 180.0)
    return tables


def comparison_csv(table: ComparisonTable) -> str:
    """comparison_csv (aero_comparison.cpp:166-181): CRLF rows, shortest round-trip numbers."""
    out = ["geometry,name,solver,aoa_deg,sideslip_deg,speed_m_s,cd,cl,cm,delta_cd,delta_cl,delta_cm,l_over_d,"
           "converged\r\n"]
    for row in table.rows:
        out.append(",".join([csv_field(table.label), csv_field(row.name), row.solver.value,
                             exact(row.angle_of_attack_deg), exact(row.sideslip_deg), exact(row.speed_m_s),
                             exact(row.cd), exact(row.cl), exact(row.cm), exact(row.delta_cd), exact(row.delta_cl),
                             exact(row.delta_cm), exact(row.lift_to_drag), "1" if row.converged else "0"]) + "\r\n")
    return "".join(out)


__all__ = ["SceneError", "SceneErrorCode", "AeroRecordError", "Experiment", "SolverInfo", "parse_experiment",
           "save_experiment", "format_json_number", "AeroShape", "GeometrySource", "SolverKind", "ReferenceAreaKind",
           "ShapeDimensions", "GeometryRef", "AeroConditions", "AeroSolverSettings", "AeroResultRecord",
           "AeroExperiment", "shape_dimensions", "to_experiment", "aero_from_experiment", "parse_aero_experiment",
           "save_aero_experiment", "exact", "csv_field", "geometry_label", "ComparisonRow", "ComparisonTable",
           "compare_by_geometry", "comparison_csv", "naca_designation_ok"]
