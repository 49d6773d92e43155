"""MI355X-native lattice Boltzmann solver for Orrery's aero library (D3Q19 / D2Q9 on gfx950).

The compute path is liborrery_lbm.so (HIP kernels + C ABI, include/orrery_lbm.h); this package is
the Python host mirror of the reference's solver API. Importing it does not load the library;
the first solver constructed does, and fails loudly when it is not built.
"""
from .config import (Lbm2DConfig, Lbm3DConfig, Outlet, Smagorinsky, SolidWall, Sponge, Wall, WallKind,
                     XBoundary, tau_of_viscosity, viscosity_of_tau)
from .masks import (SolidMask2D, SolidMask3D, box_mask, circle_mask, fluid_mask, fluid_mask_3d,
                    mask_from_sdf, mask_from_sdf_3d, polygon_mask, sphere_mask)
from .solver import (BoundaryLinks, Force2D, Force3D, Lbm2DSolver, Lbm3DSolver, LbmError, MacroFields2D, boundary_links, classify_nodes, validate,
                     MacroFields3D, device_memory, make_hip_solver, set_memory_limit)

__all__ = [
    "Lbm2DConfig", "Lbm3DConfig", "Outlet", "Smagorinsky", "SolidWall", "Sponge", "Wall", "WallKind",
    "XBoundary", "tau_of_viscosity", "viscosity_of_tau", "SolidMask2D", "SolidMask3D", "box_mask",
    "circle_mask", "fluid_mask", "fluid_mask_3d", "mask_from_sdf", "mask_from_sdf_3d", "polygon_mask", "sphere_mask",
    "BoundaryLinks", "Force2D", "Force3D", "Lbm2DSolver", "Lbm3DSolver", "LbmError", "MacroFields2D",
    "MacroFields3D", "device_memory", "make_hip_solver", "set_memory_limit", "boundary_links", "classify_nodes",
    "validate",
]
