"""ctypes mirror of include/orrery_lbm.h: the configuration structs and the library loader.

The structs follow Lbm3DConfig / Lbm2DConfig of the reference
(src/aero/lbm3d/include/orrery/aero/lbm3d/lbm3d.h:83-103,
src/aero/lbm/include/orrery/aero/lbm/lbm2d.h:43-78).
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# ORRERY_LBM_LIBRARY names another build of the same library (the host-sanitizer build of the
# CPU tests); by default the in-tree build.
LIBRARY = os.environ.get("ORRERY_LBM_LIBRARY") or os.path.join(HERE, "liborrery_lbm.so")

LBM_OK, LBM_EINVAL, LBM_ENOMEM, LBM_EDEVICE, LBM_ESTATE = 0, 1, 2, 3, 4
STORAGE = {"f32": 0, "f16s": 1, "bf16s": 2, "f64": 3}
SCHEDULE = {"fused": 0, "serial": 1, "streams": 2}
STORAGE_BYTES = {"f32": 4, "f16s": 2, "bf16s": 2, "f64": 8}


class CWall(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("velocity_x", ctypes.c_double)]


class CConfig3D(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint64), ("ny", ctypes.c_uint64), ("nz", ctypes.c_uint64),
        ("tau", ctypes.c_double),
        ("x_boundary", ctypes.c_int32), ("outlet", ctypes.c_int32),
        ("y_low", CWall), ("y_high", CWall), ("z_low", CWall), ("z_high", CWall),
        ("inlet_ux", ctypes.c_double), ("inlet_uy", ctypes.c_double), ("inlet_uz", ctypes.c_double),
        ("body_force_x", ctypes.c_double), ("body_force_y", ctypes.c_double),
        ("body_force_z", ctypes.c_double),
        ("smagorinsky_enabled", ctypes.c_int32), ("smagorinsky_constant", ctypes.c_double),
        ("sponge_strength", ctypes.c_double),
        ("sponge_inlet", ctypes.c_uint64), ("sponge_outlet", ctypes.c_uint64),
        ("sponge_sides", ctypes.c_uint64),
        ("solid_wall", ctypes.c_int32),
    ]


class CConfig2D(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint64), ("ny", ctypes.c_uint64),
        ("tau", ctypes.c_double),
        ("x_boundary", ctypes.c_int32), ("outlet", ctypes.c_int32),
        ("bottom", CWall), ("top", CWall),
        ("inlet_ux", ctypes.c_double), ("inlet_uy", ctypes.c_double),
        ("body_force_x", ctypes.c_double), ("body_force_y", ctypes.c_double),
        ("smagorinsky_enabled", ctypes.c_int32), ("smagorinsky_constant", ctypes.c_double),
    ]


class COscillation(ctypes.Structure):
    _fields_ = [("period", ctypes.c_double), ("frequency", ctypes.c_double), ("cycles", ctypes.c_size_t),
                ("amplitude", ctypes.c_double), ("period_spread", ctypes.c_double)]


class CLaunchChoice(ctypes.Structure):
    _fields_ = [("schedule", ctypes.c_int32), ("block", ctypes.c_int32), ("cells_per_thread", ctypes.c_int32),
                ("groups_a", ctypes.c_int32), ("groups_b", ctypes.c_int32), ("ms_per_step", ctypes.c_double)]


D = ctypes.c_double
U64 = ctypes.c_uint64
I32 = ctypes.c_int32


class CBatchTiming(ctypes.Structure):
    """lbm_batch_timing"""
    _fields_ = [("step_ms", D), ("bulk_ms", D * 4), ("boundary_ms", D * 4), ("force_ms", D),
                ("bulk_nodes", U64 * 4), ("boundary_nodes", U64 * 4)]


class CAeroOptions(ctypes.Structure):
    """lbm_aero3d_options (LbmSolver3DOptions, lbm_aero_solver3d.h:45-81)."""
    _fields_ = [("cells_per_length", D), ("lattice_velocity", D), ("upstream", D), ("downstream", D), ("width", D),
                ("sides", I32), ("wall", I32), ("les", I32), ("smagorinsky_constant", D), ("duration", D),
                ("average_window", D), ("start_ramp", D), ("sponge_strength", D), ("sponge_inlet", D),
                ("sponge_outlet", D), ("sponge_sides", D), ("steps_per_call", U64), ("sample_every", U64),
                ("max_cells", U64), ("storage", I32)]


class CFlowConditions(ctypes.Structure):
    _fields_ = [(n, D) for n in ("speed", "density", "viscosity", "temperature", "angle_of_attack", "sideslip",
                                 "gamma", "gas_constant")]


class CReferenceValues(ctypes.Structure):
    _fields_ = [("area", D), ("length", D), ("moment_centre", D * 3)]


class CAeroResult(ctypes.Structure):
    _fields_ = [("cd", D), ("cl", D), ("cy", D), ("cm", D), ("force", D * 3), ("moment", D * 3),
                ("iterations", U64), ("converged", I32), ("residual", D), ("samples", ctypes.c_size_t),
                ("wall_samples", ctypes.c_size_t), ("warnings", ctypes.c_size_t)]


class CAeroInfo(ctypes.Structure):
    _fields_ = [("config", CConfig3D), ("origin", D * 3), ("centre", D * 3), ("first", U64 * 3), ("count", U64 * 3),
                ("axes", D * 9), ("dx", D), ("dt", D), ("reynolds", D), ("lattice_velocity", D),
                ("cells_per_length", D), ("total_steps", U64), ("window_samples", ctypes.c_size_t),
                ("ramp_steps", D), ("blockage", D), ("surface_area", D), ("solid_nodes", ctypes.c_size_t)]


class CAero2DOptions(ctypes.Structure):
    """lbm_aero2d_options (LbmSolverOptions, lbm_aero_solver.h:38-52)."""
    _fields_ = [("cells_per_length", D), ("lattice_velocity", D), ("upstream", D), ("downstream", D), ("height", D),
                ("les", I32), ("smagorinsky_constant", D), ("duration", D), ("average_window", D),
                ("steps_per_call", U64), ("sample_every", U64), ("storage", I32)]


class CAero2DInfo(ctypes.Structure):
    _fields_ = [("config", CConfig2D), ("origin", D * 2), ("centre", D * 2), ("dx", D), ("dt", D), ("reynolds", D),
                ("lattice_velocity", D), ("cells_per_length", D), ("total_steps", U64),
                ("window_samples", ctypes.c_size_t), ("blockage", D), ("solid_nodes", ctypes.c_size_t)]


P = ctypes.POINTER
c_u8p = P(ctypes.c_uint8)
c_u32p = P(ctypes.c_uint32)
c_f64p = P(ctypes.c_double)
c_sizep = P(ctypes.c_size_t)

# (name, restype, argtypes) of every export declared in include/orrery_lbm.h.
EXPORTS = [
    ("lbm_last_error", ctypes.c_char_p, []),
    ("lbm_version", ctypes.c_char_p, []),
    ("lbm_device_memory", ctypes.c_int, [c_sizep, c_sizep]),
    ("lbm_set_memory_limit", ctypes.c_int, [ctypes.c_size_t]),
    ("lbm_device_name", ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    ("lbm3d_create", ctypes.c_int, [P(CConfig3D), c_u8p, c_f64p, ctypes.c_int32, P(ctypes.c_void_p)]),
    ("lbm3d_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("lbm3d_initialize", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm3d_initialize_perturbed", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double]),
    ("lbm3d_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64]),
    ("lbm3d_steps", ctypes.c_uint64, [ctypes.c_void_p]),
    ("lbm3d_set_inlet", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm3d_get_inlet", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm3d_force", ctypes.c_int, [ctypes.c_void_p, c_f64p, c_f64p]),
    ("lbm3d_mass", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm3d_macroscopic", ctypes.c_int, [ctypes.c_void_p, c_f64p, c_f64p, c_f64p, c_f64p]),
    ("lbm3d_sample", ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_uint32), ctypes.c_size_t, c_f64p, c_f64p, c_f64p,
                                    c_f64p]),
    ("lbm3d_sample_fixed", ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_uint32), ctypes.c_size_t, c_f64p, c_f64p,
                                          c_f64p, c_f64p]),
    ("lbm3d_max_speed", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm3d_populations", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm3d_links", ctypes.c_int, [ctypes.c_void_p, c_u32p, c_u32p, c_f64p, ctypes.c_size_t, c_sizep]),
    ("lbm3d_device_bytes", ctypes.c_int, [ctypes.c_void_p, c_sizep]),
    ("lbm3d_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("lbm3d_set_schedule", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("lbm3d_set_graphs", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("lbm3d_get_launch", ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int32)]),
    ("lbm3d_autotune", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, P(CLaunchChoice),
                                      ctypes.c_size_t, c_sizep, P(CLaunchChoice)]),
    ("lbm3d_set_launch", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    ("lbm3d_set_cells_per_thread", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    ("lbm3d_get_cells_per_thread", ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_int32), P(ctypes.c_int32)]),
    ("lbm3d_time_kernels", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, c_f64p]),
    ("lbm3d_set_occupancy", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    ("lbm3d_get_occupancy", ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_int32), P(ctypes.c_int32)]),
    ("lbm3d_step_bytes", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm3d_counts", ctypes.c_int, [ctypes.c_void_p, c_sizep, c_sizep, c_sizep, c_sizep]),
    ("lbm3d_stream_counts", ctypes.c_int, [ctypes.c_void_p, c_sizep, c_sizep, c_sizep]),
    ("lbm3d_time_steps", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, c_f64p, c_f64p, c_f64p]),
    ("lbm2d_create", ctypes.c_int, [P(CConfig2D), c_u8p, ctypes.c_int32, P(ctypes.c_void_p)]),
    ("lbm2d_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("lbm2d_initialize", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm2d_initialize_perturbed", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double]),
    ("lbm2d_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64]),
    ("lbm2d_steps", ctypes.c_uint64, [ctypes.c_void_p]),
    ("lbm2d_set_graphs", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("lbm2d_set_schedule", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("lbm2d_force", ctypes.c_int, [ctypes.c_void_p, c_f64p, c_f64p, c_f64p]),
    ("lbm2d_mass", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm2d_macroscopic", ctypes.c_int, [ctypes.c_void_p, c_f64p, c_f64p, c_f64p]),
    ("lbm2d_populations", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm2d_links", ctypes.c_int, [ctypes.c_void_p, c_u32p, c_u32p, ctypes.c_size_t, c_sizep]),
    ("lbm2d_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("lbm3d_validate", ctypes.c_int, [P(CConfig3D), c_u8p, c_f64p]),
    ("lbm2d_validate", ctypes.c_int, [P(CConfig2D), c_u8p]),
    ("lbm3d_classify", ctypes.c_int, [P(CConfig3D), c_u8p, c_u8p]),
    ("lbm3d_boundary_links", ctypes.c_int, [P(CConfig3D), c_u8p, c_f64p, c_u32p, c_u32p, c_f64p, ctypes.c_size_t,
                                            c_sizep]),
    ("lbm2d_boundary_links", ctypes.c_int, [P(CConfig2D), c_u8p, c_u32p, c_u32p, ctypes.c_size_t, c_sizep]),
    ("lbm_perturbation", ctypes.c_double, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32]),
    ("lbm_batch_create", ctypes.c_int, [ctypes.c_size_t, ctypes.c_uint32, P(ctypes.c_void_p)]),
    ("lbm_batch_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("lbm_batch_required_bytes", ctypes.c_int, [P(CConfig3D), c_u8p, c_f64p, ctypes.c_int32, c_sizep]),
    ("lbm_batch_add", ctypes.c_int, [ctypes.c_void_p, P(CConfig3D), c_u8p, c_f64p, ctypes.c_int32, ctypes.c_uint64,
                                     ctypes.c_double, P(ctypes.c_int32)]),
    ("lbm_batch_remove", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("lbm_batch_run", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64]),
    ("lbm_batch_run_timed", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(CBatchTiming)]),
    ("lbm_batch_count", ctypes.c_int, [ctypes.c_void_p, c_sizep]),
    ("lbm_batch_lattice", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_void_p)]),
    ("lbm_batch_history", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_f64p, c_f64p, ctypes.c_size_t,
                                         c_sizep]),
    ("lbm_batch_bytes", ctypes.c_int, [ctypes.c_void_p, c_sizep, c_sizep]),
    ("lbm_batch_set_launch", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    ("lbm_batch_set_occupancy", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    ("lbm_batch_autotune", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, P(CLaunchChoice)]),
    ("lbm_batch_set_inlet", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_f64p]),
    ("lbm_force_coefficients", None, [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, c_f64p,
                                      c_f64p]),
    ("lbm_mean_since", ctypes.c_double, [c_f64p, c_f64p, ctypes.c_size_t, ctypes.c_double]),
    ("lbm_dominant_oscillation", ctypes.c_int, [c_f64p, c_f64p, ctypes.c_size_t, ctypes.c_double,
                                                P(COscillation)]),
    ("lbm_steady_since", ctypes.c_int, [c_f64p, c_f64p, c_f64p, ctypes.c_size_t, ctypes.c_double,
                                        ctypes.c_double]),
    ("lbm_settled_oscillation", ctypes.c_int, [c_f64p, c_f64p, c_f64p, ctypes.c_size_t, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_double, P(COscillation)]),
    ("lbm_forces_settled", ctypes.c_int, [c_f64p, c_f64p, c_f64p, c_f64p, ctypes.c_size_t, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    ("lbm_geom_last_error", ctypes.c_char_p, []),
    ("lbm_mesh_create", ctypes.c_int, [c_f64p, ctypes.c_size_t, P(ctypes.c_uint32), ctypes.c_size_t,
                                       P(ctypes.c_void_p)]),
    ("lbm_mesh_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("lbm_mesh_size", ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_size_t), P(ctypes.c_size_t)]),
    ("lbm_mesh_data", ctypes.c_int, [ctypes.c_void_p, c_f64p, P(ctypes.c_uint32)]),
    ("lbm_mesh_revolution", ctypes.c_int, [ctypes.c_int32] + [ctypes.c_double] * 5
     + [ctypes.c_int32, ctypes.c_double, P(ctypes.c_void_p)]),
    ("lbm_mesh_wedge", ctypes.c_int, [ctypes.c_double] * 3 + [P(ctypes.c_void_p)]),
    ("lbm_mesh_box", ctypes.c_int, [ctypes.c_double] * 3 + [P(ctypes.c_void_p)]),
    ("lbm_mesh_naca_wing", ctypes.c_int, [ctypes.c_double] * 5 + [ctypes.c_int32, ctypes.c_int32,
                                                                  P(ctypes.c_void_p)]),
    ("lbm_mesh_transform", ctypes.c_int, [ctypes.c_void_p, c_f64p, c_f64p, c_f64p]),
    ("lbm_mesh_bounds", ctypes.c_int, [ctypes.c_void_p, c_f64p, c_f64p]),
    ("lbm_voxelize", ctypes.c_int, [ctypes.c_void_p, c_f64p, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64,
                                    ctypes.c_uint64, ctypes.c_double, P(ctypes.c_uint8), c_f64p]),
    ("lbm_signed_distance_band", ctypes.c_int, [ctypes.c_void_p, c_f64p, ctypes.c_double, ctypes.c_uint64,
                                                ctypes.c_uint64, ctypes.c_uint64, P(ctypes.c_uint8),
                                                ctypes.c_double, c_f64p, P(ctypes.c_uint8)]),
    ("lbm_winding_numbers", ctypes.c_int, [ctypes.c_void_p, c_f64p, ctypes.c_size_t, c_f64p]),
    ("lbm_voxelize_device", ctypes.c_int, [ctypes.c_void_p, c_f64p, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64,
                                           ctypes.c_uint64, ctypes.c_double, P(ctypes.c_uint8), c_f64p, c_sizep]),
    ("lbm_signed_distance_band_device", ctypes.c_int, [ctypes.c_void_p, c_f64p, ctypes.c_double, ctypes.c_uint64,
                                                       ctypes.c_uint64, ctypes.c_uint64, P(ctypes.c_uint8),
                                                       ctypes.c_double, c_f64p, P(ctypes.c_uint8)]),
    ("lbm_aero3d_last_error", ctypes.c_char_p, []),
    ("lbm_aero3d_default_options", None, [P(CAeroOptions)]),
    ("lbm_default_flow_conditions", None, [P(CFlowConditions)]),
    ("lbm_aero3d_create", ctypes.c_int, [P(CAeroOptions), P(ctypes.c_void_p)]),
    ("lbm_aero3d_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("lbm_aero3d_plan", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, P(CReferenceValues), P(CFlowConditions)]),
    ("lbm_aero3d_setup", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, P(CReferenceValues), P(CFlowConditions)]),
    ("lbm_aero3d_step", ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_int32)]),
    ("lbm_aero3d_get_result", ctypes.c_int, [ctypes.c_void_p, P(CAeroResult)]),
    ("lbm_aero3d_history", ctypes.c_int, [ctypes.c_void_p, c_f64p, c_f64p, c_f64p, c_f64p, ctypes.c_size_t,
                                          P(ctypes.c_size_t)]),
    ("lbm_aero3d_residuals", ctypes.c_int, [ctypes.c_void_p, c_f64p, ctypes.c_size_t, P(ctypes.c_size_t)]),
    ("lbm_aero3d_instantaneous", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm_aero3d_surface", ctypes.c_int, [ctypes.c_void_p, c_f64p, c_f64p, c_f64p, c_f64p, ctypes.c_size_t,
                                          P(ctypes.c_size_t)]),
    ("lbm_aero3d_warning_count", ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_size_t)]),
    ("lbm_aero3d_warning", ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t,
                                          ctypes.c_char_p, ctypes.c_size_t]),
    ("lbm_aero3d_get_info", ctypes.c_int, [ctypes.c_void_p, P(CAeroInfo)]),
    ("lbm_aero3d_mask", ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_uint8), c_f64p]),
    ("lbm_aero3d_lattice", ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_void_p)]),
    ("lbm_aero3d_velocity", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm_aero2d_last_error", ctypes.c_char_p, []),
    ("lbm_aero2d_default_options", None, [P(CAero2DOptions)]),
    ("lbm_aero2d_create", ctypes.c_int, [P(CAero2DOptions), P(ctypes.c_void_p)]),
    ("lbm_aero2d_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("lbm_aero2d_plan", ctypes.c_int, [ctypes.c_void_p, c_f64p, ctypes.c_size_t, P(CReferenceValues),
                                       P(CFlowConditions)]),
    ("lbm_aero2d_setup", ctypes.c_int, [ctypes.c_void_p, c_f64p, ctypes.c_size_t, P(CReferenceValues),
                                        P(CFlowConditions)]),
    ("lbm_aero2d_step", ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_int32)]),
    ("lbm_aero2d_get_result", ctypes.c_int, [ctypes.c_void_p, P(CAeroResult)]),
    ("lbm_aero2d_history", ctypes.c_int, [ctypes.c_void_p, c_f64p, c_f64p, c_f64p, ctypes.c_size_t, c_sizep]),
    ("lbm_aero2d_residuals", ctypes.c_int, [ctypes.c_void_p, c_f64p, ctypes.c_size_t, c_sizep]),
    ("lbm_aero2d_instantaneous", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm_aero2d_strouhal", ctypes.c_int, [ctypes.c_void_p, c_f64p, P(ctypes.c_int32)]),
    ("lbm_aero2d_surface", ctypes.c_int, [ctypes.c_void_p, c_f64p, c_f64p, c_f64p, c_f64p, ctypes.c_size_t,
                                          c_sizep]),
    ("lbm_aero2d_warning_count", ctypes.c_int, [ctypes.c_void_p, c_sizep]),
    ("lbm_aero2d_warning", ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t,
                                          ctypes.c_char_p, ctypes.c_size_t]),
    ("lbm_aero2d_get_info", ctypes.c_int, [ctypes.c_void_p, P(CAero2DInfo)]),
    ("lbm_aero2d_mask", ctypes.c_int, [ctypes.c_void_p, c_u8p]),
    ("lbm_aero2d_lattice", ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_void_p)]),
    ("lbm_aero2d_velocity", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm2d_sample", ctypes.c_int, [ctypes.c_void_p, c_u32p, ctypes.c_size_t, c_f64p, c_f64p, c_f64p]),
    ("lbm2d_max_speed", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("lbm_naca_section", ctypes.c_int, [D, D, D, I32, c_f64p]),
    ("lbm_polygon_mask", ctypes.c_int, [U64, U64, c_f64p, ctypes.c_size_t, c_u8p]),
]

_lib = None


def load() -> ctypes.CDLL:
    """Loads liborrery_lbm.so (built in-tree by __graft_entry__.build()); raises when missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY):
        raise RuntimeError(
            f"orrery_amd: the HIP library {LIBRARY} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIBRARY)
    for name, restype, argtypes in EXPORTS:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


def ptr(array, ctype):
    """Pointer to the data of a contiguous numpy array (None passes NULL)."""
    if array is None:
        return None
    return array.ctypes.data_as(P(ctype))
