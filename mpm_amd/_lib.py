"""ctypes binding of libmpm_hip.so (the C-ABI declared in include/mpm_hip.h).

The library is loaded from the package directory only (it is built in-tree by
mpm_amd.build). Loading it needs no GPU; calls that touch the device return
MPM_ERR_NOGPU when none is visible. There is no CPU fallback: if the library
is missing, every entry point raises.

PyTorch-ROCm is imported before the library is opened: torch ships its own
HIP runtime (libamdhip64.so.7), and with it already mapped the engine's
dependency on that soname resolves to the same copy. One HIP runtime per
process is what lets the engine and torch share device pointers and streams
(the adjoint path hands torch tensors to the C-ABI).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import torch  # noqa: F401  (maps torch's HIP runtime before libmpm_hip.so)

LIB_PATH = Path(__file__).resolve().parent / "libmpm_hip.so"

MPM_OK = 0
MPM_ERR_INVALID = 1
MPM_ERR_HIP = 2
MPM_ERR_STATE = 3
MPM_ERR_CAPACITY = 4
MPM_ERR_IO = 5
MPM_ERR_NOGPU = 6

MATERIAL_SNOW = 0
MATERIAL_FIXED_COROTATED = 1
MATERIAL_JELLY = 2
MATERIAL_WC_LIQUID = 3

BOUNDARY_FRACTION = 0
BOUNDARY_CELLS = 1

COMPAT_REFERENCE = 0
COMPAT_FIXED = 1

INTERP_QUADRATIC = 0
INTERP_CUBIC = 1


class mpm_params(C.Structure):
    _fields_ = [
        ("dt", C.c_float),
        ("N", C.c_uint32),
        ("gravity", C.c_float),
        ("boundary_mode", C.c_int32),
        ("boundary", C.c_float),
        ("compat", C.c_int32),
    ]


class mpm_material(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("particle_volume", C.c_float),
        ("particle_mass", C.c_float),
        ("mu0", C.c_float),
        ("lambda0", C.c_float),
        ("hardening", C.c_float),
        ("clamp_lo", C.c_float),
        ("clamp_hi", C.c_float),
        ("bulk_modulus", C.c_float),
    ]


class MPMError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[mpm status {code}] {message}")
        self.code = code


class LibraryNotBuilt(RuntimeError):
    pass


_P = C.c_void_p
_f = C.POINTER(C.c_float)
_u8 = C.POINTER(C.c_uint8)
_u32 = C.POINTER(C.c_uint32)
_i64p = C.POINTER(C.c_int64)

# name -> (restype, argtypes); every symbol declared in include/mpm_hip.h
SIGNATURES = {
    "mpm_material_init": (C.c_int, [C.POINTER(mpm_material), C.c_int32, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_double]),
    "mpm_create": (C.c_int, [C.POINTER(mpm_params), C.POINTER(_P)]),
    "mpm_destroy": (C.c_int, [_P]),
    "mpm_last_error": (C.c_char_p, [_P]),
    "mpm_abi_version": (C.c_int, []),
    "mpm_set_materials": (C.c_int, [_P, C.POINTER(mpm_material), C.c_int32]),
    "mpm_add_particles_soa": (C.c_int, [_P, C.c_int64, _f, _f, _f, _f, _f, _u8]),
    "mpm_add_particles_aos104": (C.c_int, [_P, C.c_void_p, C.c_int64]),
    "mpm_clear_particles": (C.c_int, [_P]),
    "mpm_upload": (C.c_int, [_P, _P]),
    "mpm_append_particles": (C.c_int, [_P, _P]),
    "mpm_mesh_build": (C.c_int, [_P, C.c_int32, C.c_int32, _P, _P]),
    "mpm_mesh_triangles": (C.c_int, [_P, _P, _P]),
    "mpm_mesh_distance_field": (C.c_int, [_P, C.c_int32, _P]),
    "mpm_mesh_write_obj": (C.c_int, [_P, C.c_char_p]),
    "mpm_count": (C.c_int, [_P, _i64p]),
    "mpm_mesh_indexed": (C.c_int, [_P, _i64p, _i64p, _P, _P]),
    "mpm_mesh_case_table": (C.c_int, [_P, _P]),
    "mpm_render_frame": (C.c_int, [_P, C.c_int32, C.c_int32, _P, _P]),
    "mpm_mesh_smooth": (C.c_int, [C.c_int64, _P, C.c_int64, _P]),
    "mpm_mesh_decimate": (C.c_int, [C.c_int64, _P, C.c_int64, _P, C.c_int64, _P, _i64p, _P, _i64p]),
    "mpm_obj_write": (C.c_int, [C.c_char_p, C.c_int64, _P, C.c_int64, _P]),
    "mpm_bgeo_write": (C.c_int, [C.c_char_p, C.c_int64, _P, _P]),
    "mpm_remove_particles": (C.c_int, [_P, C.c_int64, C.c_int64, _P, _P]),
    "mpm_advance": (C.c_int, [_P, C.c_int32, _P]),
    "mpm_time": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "mpm_download_soa": (C.c_int, [_P, _f, _f, _f, _f, _f, _u8, _P]),
    "mpm_download_aos104": (C.c_int, [_P, C.c_void_p, _P]),
    "mpm_download_grid": (C.c_int, [_P, _f, _P]),
    "mpm_save_state": (C.c_int, [_P, C.c_char_p]),
    "mpm_load_state": (C.c_int, [_P, C.c_char_p]),
    "mpm_set_option": (C.c_int, [_P, C.c_char_p, C.c_int64]),
    "mpm_get_option": (C.c_int, [_P, C.c_char_p, _i64p]),
    "mpm_timing_read": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double), _i64p]),
    "mpm_timing_reset": (C.c_int, [_P]),
    "mpm_autotune": (C.c_int, [_P, C.c_int32, C.c_char_p, C.c_int32, _P]),
    "mpm_debug_p2g": (C.c_int, [_P, _f, _P]),
    "mpm_debug_sort": (C.c_int, [_P, _u32, _u32, _P]),
    "mpm_debug_svd": (C.c_int, [C.c_int64, _f, _f, _f, _f]),
    "mpm_device_info": (C.c_int, [C.c_char_p, C.c_int32]),
    "mpm_debug_selftest": (C.c_int, [C.c_char_p, C.c_int32]),
    "mpm_batch_create": (C.c_int, [C.POINTER(_P)]),
    "mpm_batch_destroy": (C.c_int, [_P]),
    "mpm_batch_last_error": (C.c_char_p, [_P]),
    "mpm_batch_add": (C.c_int, [_P, _P]),
    "mpm_batch_count": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "mpm_batch_advance": (C.c_int, [_P, C.c_int32, _P]),
    "mpm_batch_set_option": (C.c_int, [_P, C.c_char_p, C.c_int64]),
    "mpm_batch_get_option": (C.c_int, [_P, C.c_char_p, _i64p]),
    "mpm_batch_autotune": (C.c_int, [_P, C.c_int32, C.c_char_p, C.c_int32, _P]),
    "mpm_scene_load": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.POINTER(_P)]),
    "mpm_scene_destroy": (C.c_int, [_P]),
    "mpm_scene_last_error": (C.c_char_p, []),
    "mpm_scene_info": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _i64p]),
    "mpm_scene_materials": (C.c_int, [_P, C.POINTER(mpm_material)]),
    "mpm_scene_object": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), _i64p, _i64p,
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mpm_scene_particles_soa": (C.c_int, [_P, _f, _f, _u8]),
    "mpm_sampler_create": (C.c_int, [C.c_uint32, C.POINTER(_P)]),
    "mpm_sampler_destroy": (C.c_int, [_P]),
    "mpm_sampler_fill_mesh": (C.c_int, [_P, C.c_char_p, C.c_double, _f, C.c_float, _i64p]),
    "mpm_sampler_points": (C.c_int, [_P, _f]),
    "mpm_debug_glibc_rand": (C.c_int, [C.c_uint32, C.c_int32, C.POINTER(C.c_int32)]),
    "mpm_diff_sizes": (C.c_int, [C.POINTER(mpm_params), _i64p, _i64p]),
    # device pointers travel as void*
    "mpm_diff_substep": (C.c_int, [C.POINTER(mpm_params), C.c_float, C.c_float, C.c_int64,
                                   _P, _P, _P, _P, _P, _P, _P]),
    "mpm_diff_substep_noalloc": (C.c_int, [C.POINTER(mpm_params), C.c_float, C.c_float, C.c_int64,
                                           _P, _P, _P, _P, _P, _P]),
    "mpm_diff_forward_steps": (C.c_int, [C.POINTER(mpm_params), C.c_float, C.c_float, C.c_int64,
                                         C.c_int32, _P, _P, C.c_int64, _P, _P, _P, _P, C.c_int64, _P]),
    "mpm_diff_backward_steps": (C.c_int, [C.POINTER(mpm_params), C.c_float, C.c_float, C.c_int64,
                                          C.c_int32, _P, C.c_int64, _P, _P, _P, C.c_int64, _P, _P, _P, _P,
                                          _P, _P, _P]),
    "mpm_diff_substep_backward": (C.c_int, [C.POINTER(mpm_params), C.c_float, C.c_float,
                                            C.c_int64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
}


def selftest() -> str:
    """Status of each step of mpm_debug_selftest."""
    buf = C.create_string_buffer(4096)
    lib().mpm_debug_selftest(buf, len(buf))
    return buf.value.decode()


def device_info() -> str:
    """Runtime/device description string (see mpm_device_info)."""
    buf = C.create_string_buffer(1024)
    lib().mpm_device_info(buf, len(buf))
    return buf.value.decode()

_lib = None


def lib() -> C.CDLL:
    """Load libmpm_hip.so (once). Raises LibraryNotBuilt if it is absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise LibraryNotBuilt(
                f"{LIB_PATH} is not built; run `python -m mpm_amd.build` (or __graft_entry__.build())")
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(rc: int, handle=None, what: str = "") -> None:
    if rc != MPM_OK:
        msg = lib().mpm_last_error(handle)
        raise MPMError(rc, (msg.decode() if msg else "") or what)
