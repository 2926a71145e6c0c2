"""mpm_amd -- MI355X (gfx950) MLS-MPM substep engine.

The compute path is libmpm_hip.so (hand-written HIP kernels behind the C-ABI
in include/mpm_hip.h); this package only binds it (ctypes), builds it and
generates seeded workloads. There is no CPU fallback.
"""
from ._lib import (BOUNDARY_CELLS, BOUNDARY_FRACTION, COMPAT_FIXED, COMPAT_REFERENCE, INTERP_CUBIC,
                   INTERP_QUADRATIC, MATERIAL_FIXED_COROTATED, MATERIAL_JELLY, MATERIAL_SNOW, MATERIAL_WC_LIQUID,
                   LibraryNotBuilt, MPMError, lib, mpm_material, mpm_params)
from .engine import Batch, Engine, Params, device_svd, make_material, material_as_numpy

__all__ = [
    "Batch", "Engine", "Params", "make_material", "material_as_numpy", "device_svd", "lib", "MPMError",
    "LibraryNotBuilt", "mpm_material", "mpm_params", "MATERIAL_SNOW", "MATERIAL_FIXED_COROTATED",
    "MATERIAL_JELLY", "MATERIAL_WC_LIQUID", "BOUNDARY_FRACTION", "BOUNDARY_CELLS",
    "COMPAT_REFERENCE", "COMPAT_FIXED", "INTERP_QUADRATIC", "INTERP_CUBIC",
]
