"""ctypes binding of oracle/_ref/liborrery_ref.so: the reference's setup code compiled from its own
sources (oracle/ref/Makefile). Present only where the reference tree was available to build it;
tests that need it skip otherwise and fall back to the committed fixtures in tests/golden/."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from orrery_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBRARY = os.path.join(ROOT, "oracle", "_ref", "liborrery_ref.so")
P = ctypes.POINTER
_f64p = P(ctypes.c_double)
_u8p = P(ctypes.c_uint8)
_u32p = P(ctypes.c_uint32)
_lib = None


def available() -> bool:
    return os.path.exists(LIBRARY)


def load():
    global _lib
    if _lib is not None:
        return _lib
    lib = ctypes.CDLL(LIBRARY)
    u64 = ctypes.c_uint64
    sig = {
        "ref3d_validate": (ctypes.c_int, [P(abi.CConfig3D), _u8p, _f64p, ctypes.c_char_p, ctypes.c_size_t]),
        "ref3d_boundary_links": (ctypes.c_size_t, [P(abi.CConfig3D), _u8p, _f64p, _u32p, _u32p, _f64p, ctypes.c_size_t]),
        "ref3d_classify": (None, [P(abi.CConfig3D), _u8p, _u8p]),
        "ref3d_resolve": (ctypes.c_int, [P(abi.CConfig3D), _u8p, u64, u64, u64, ctypes.c_int, P(u64), P(ctypes.c_int),
                                         _f64p]),
        "ref3d_link_fraction": (ctypes.c_double, [P(abi.CConfig3D), _u8p, _f64p, u64, u64, u64, ctypes.c_int]),
        "ref3d_sponge": (ctypes.c_double, [P(abi.CConfig3D), u64, u64, u64]),
        "ref3d_inlet_west": (None, [_f64p, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "ref3d_equilibrium": (ctypes.c_double, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double]),
        "ref3d_initial_populations": (None, [P(abi.CConfig3D), _u8p, _f64p, _f64p]),
        "ref3d_gpu_memory_bytes": (u64, [P(abi.CConfig3D), u64, u64]),
        "ref3d_sphere_mask": (None, [u64, u64, u64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, _u8p, _f64p]),
        "ref3d_box_mask": (None, [u64, u64, u64, _f64p, _f64p, _u8p, _f64p]),
        "ref2d_validate": (ctypes.c_int, [P(abi.CConfig2D), _u8p, ctypes.c_char_p, ctypes.c_size_t]),
        "ref2d_boundary_links": (ctypes.c_size_t, [P(abi.CConfig2D), _u8p, _u32p, _u32p, ctypes.c_size_t]),
        "ref2d_resolve": (ctypes.c_int, [P(abi.CConfig2D), _u8p, u64, u64, ctypes.c_int, P(u64), P(ctypes.c_int),
                                         _f64p]),
        "ref2d_zou_he_west": (None, [_f64p, ctypes.c_double, ctypes.c_double]),
        "ref2d_circle_mask": (None, [u64, u64, ctypes.c_double, ctypes.c_double, ctypes.c_double, _u8p]),
        "ref_lattice_tau": (ctypes.c_double, [ctypes.c_double] * 6),
        "ref2d_polygon_mask": (None, [u64, u64, _f64p, ctypes.c_size_t, _u8p]),
        "ref_naca_section": (None, [ctypes.c_double] * 3 + [ctypes.c_int, _f64p]),
        "ref_naca_section_area": (ctypes.c_double, [ctypes.c_double] * 3),
        "ref_mesh_revolution": (ctypes.c_size_t, [ctypes.c_int] + [ctypes.c_double] * 5
                                + [ctypes.c_int, ctypes.c_double, _f64p, ctypes.c_size_t, _u32p, ctypes.c_size_t,
                                   P(ctypes.c_size_t)]),
        "ref_mesh_wedge": (ctypes.c_size_t, [ctypes.c_double] * 3 + [_f64p, ctypes.c_size_t, _u32p, ctypes.c_size_t,
                                                                     P(ctypes.c_size_t)]),
        "ref_mesh_naca_wing": (ctypes.c_size_t, [ctypes.c_double] * 5 + [ctypes.c_int, ctypes.c_int, _f64p,
                                                                         ctypes.c_size_t, _u32p, ctypes.c_size_t,
                                                                         P(ctypes.c_size_t)]),
        "ref_voxelize": (None, [_f64p, ctypes.c_size_t, _u32p, ctypes.c_size_t, _f64p, ctypes.c_double, ctypes.c_size_t,
                                ctypes.c_size_t, ctypes.c_size_t, ctypes.c_double, _u8p, _f64p]),
        "ref_reference_quantities": (None, [_f64p, ctypes.c_size_t, _u32p, ctypes.c_size_t, ctypes.c_size_t, _f64p]),
        "ref_signed_distance_band": (None, [_f64p, ctypes.c_size_t, _u32p, ctypes.c_size_t, _f64p, ctypes.c_double,
                                            ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, _u8p, ctypes.c_double,
                                            _f64p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a, t=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(P(t))


def sphere_mask(nx, ny, nz, centre, radius):
    n = nx * ny * nz
    solid = np.empty(n, dtype=np.uint8)
    dist = np.empty(n)
    load().ref3d_sphere_mask(nx, ny, nz, *[float(c) for c in centre], float(radius), ptr(solid, ctypes.c_uint8),
                             ptr(dist))
    return solid, dist


def box_mask(nx, ny, nz, lo, hi):
    n = nx * ny * nz
    solid = np.empty(n, dtype=np.uint8)
    dist = np.empty(n)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    load().ref3d_box_mask(nx, ny, nz, ptr(lo), ptr(hi), ptr(solid, ctypes.c_uint8), ptr(dist))
    return solid, dist


def circle_mask(nx, ny, cx, cy, r):
    solid = np.empty(nx * ny, dtype=np.uint8)
    load().ref2d_circle_mask(nx, ny, cx, cy, r, ptr(solid, ctypes.c_uint8))
    return solid


def polygon_mask(nx, ny, points):
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 2))
    solid = np.empty(nx * ny, dtype=np.uint8)
    load().ref2d_polygon_mask(nx, ny, ptr(pts), len(pts), ptr(solid, ctypes.c_uint8))
    return solid


def naca_section(m, p, t, n):
    out = np.empty((2 * n, 2))
    load().ref_naca_section(m, p, t, n, ptr(out))
    return out


def naca_section_area(m, p, t):
    return load().ref_naca_section_area(m, p, t)


def boundary_links3(config, mask):
    lib = load()
    c = config.to_c()
    solid = np.ascontiguousarray(mask.solid, dtype=np.uint8)
    dist = None if mask.distance is None else np.ascontiguousarray(mask.distance)
    n = lib.ref3d_boundary_links(ctypes.byref(c), ptr(solid, ctypes.c_uint8), ptr(dist), None, None, None, 0)
    cell = np.empty(n, dtype=np.uint32)
    d = np.empty(n, dtype=np.uint32)
    q = np.empty(n)
    lib.ref3d_boundary_links(ctypes.byref(c), ptr(solid, ctypes.c_uint8), ptr(dist), ptr(cell, ctypes.c_uint32),
                             ptr(d, ctypes.c_uint32), ptr(q), n)
    return cell, d, q


def classify3(config, mask):
    c = config.to_c()
    out = np.empty(config.cells, dtype=np.uint8)
    load().ref3d_classify(ctypes.byref(c), ptr(np.ascontiguousarray(mask.solid, dtype=np.uint8), ctypes.c_uint8),
                          ptr(out, ctypes.c_uint8))
    return out


def validate3(config, mask):
    c = config.to_c()
    buf = ctypes.create_string_buffer(512)
    dist = None if mask.distance is None else np.ascontiguousarray(mask.distance)
    rc = load().ref3d_validate(ctypes.byref(c), ptr(np.ascontiguousarray(mask.solid, dtype=np.uint8), ctypes.c_uint8),
                               ptr(dist), buf, 512)
    return buf.value.decode() if rc else None


def validate2(config, mask):
    c = config.to_c()
    buf = ctypes.create_string_buffer(512)
    rc = load().ref2d_validate(ctypes.byref(c), ptr(np.ascontiguousarray(mask.solid, dtype=np.uint8), ctypes.c_uint8),
                               buf, 512)
    return buf.value.decode() if rc else None


def boundary_links2(config, mask):
    lib = load()
    c = config.to_c()
    solid = np.ascontiguousarray(mask.solid, dtype=np.uint8)
    n = lib.ref2d_boundary_links(ctypes.byref(c), ptr(solid, ctypes.c_uint8), None, None, 0)
    cell = np.empty(n, dtype=np.uint32)
    d = np.empty(n, dtype=np.uint32)
    lib.ref2d_boundary_links(ctypes.byref(c), ptr(solid, ctypes.c_uint8), ptr(cell, ctypes.c_uint32),
                             ptr(d, ctypes.c_uint32), n)
    return cell, d


def _mesh(call):
    """Runs a ref_mesh_* binding twice: once for the sizes, once into arrays of that size."""
    nt = ctypes.c_size_t()
    nv = call(None, 0, None, 0, ctypes.byref(nt))
    v = np.empty((nv, 3))
    t = np.empty((nt.value, 3), dtype=np.uint32)
    call(ptr(v), nv, ptr(t, ctypes.c_uint32), nt.value, ctypes.byref(nt))
    return v, t


def revolution_mesh(kind, a, b=0.0, c=0.0, d=0.0, e=0.0, around=64, arc_step=0.05):
    lib = load()
    return _mesh(lambda *r: lib.ref_mesh_revolution(kind, a, b, c, d, e, around, arc_step, *r))


def wedge_mesh(length, half_angle, span):
    lib = load()
    return _mesh(lambda *r: lib.ref_mesh_wedge(length, half_angle, span, *r))


def naca_wing(m, p, t, chord, span, n, div):
    lib = load()
    return _mesh(lambda *r: lib.ref_mesh_naca_wing(m, p, t, chord, span, n, div, *r))


def voxelize(v, t, origin, spacing, nx, ny, nz, threshold=0.5):
    v = np.ascontiguousarray(v, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.uint32)
    o = np.asarray(origin, dtype=np.float64)
    solid = np.empty(nx * ny * nz, dtype=np.uint8)
    w = np.empty(nx * ny * nz)
    load().ref_voxelize(ptr(v), len(v), ptr(t, ctypes.c_uint32), len(t), ptr(o), spacing, nx, ny, nz, threshold,
                        ptr(solid, ctypes.c_uint8), ptr(w))
    return solid, w


def signed_distance_band(v, t, origin, spacing, nx, ny, nz, solid, band_cells=3.0):
    v = np.ascontiguousarray(v, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.uint32)
    o = np.asarray(origin, dtype=np.float64)
    d = np.empty(nx * ny * nz)
    load().ref_signed_distance_band(ptr(v), len(v), ptr(t, ctypes.c_uint32), len(t), ptr(o), spacing, nx, ny, nz,
                                    ptr(np.ascontiguousarray(solid, dtype=np.uint8), ctypes.c_uint8), band_cells,
                                    ptr(d))
    return d


REFERENCE_QUANTITY_NAMES = ("wetted_area", "frontal_area", "planform_area", "frontal_area_bound",
                            "planform_area_bound", "reference_length", "span", "volume", "centre_x", "centre_y",
                            "centre_z")


def reference_quantities(v, t, resolution):
    """reference_quantities of the reference (flow +x, lift +z) on a mesh, as a dict."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.uint32)
    out = np.empty(len(REFERENCE_QUANTITY_NAMES))
    load().ref_reference_quantities(ptr(v), len(v), ptr(t, ctypes.c_uint32), len(t), resolution, ptr(out))
    return dict(zip(REFERENCE_QUANTITY_NAMES, out.tolist()))
