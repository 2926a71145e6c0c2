"""The C-ABI boundary on the CPU: libmpm_hip.so loads, exports every entry
point include/mpm_hip.h declares, and its host-side logic (material
construction, argument validation, error reporting) behaves -- without
launching any kernel."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import mpm_amd
from mpm_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "mpm_hip.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mpm_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("mpm_create", "mpm_destroy", "mpm_set_materials", "mpm_add_particles_soa",
                 "mpm_add_particles_aos104", "mpm_upload", "mpm_append_particles", "mpm_mesh_build",
                 "mpm_mesh_triangles", "mpm_mesh_distance_field", "mpm_mesh_write_obj", "mpm_advance", "mpm_download_soa",
                 "mpm_download_aos104", "mpm_save_state", "mpm_load_state", "mpm_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = mpm_amd.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_abi_version():
    assert mpm_amd.lib().mpm_abi_version() == 1


def test_struct_layouts_match_header():
    assert C.sizeof(mpm_amd.mpm_params) == 24
    assert C.sizeof(mpm_amd.mpm_material) == 36


@pytest.mark.parametrize("args", [
    (mpm_amd.MATERIAL_SNOW, 1.0 / 500000, 700.0, 1.4e5, 0.2, 10.0, 0.975, 1.0075, 0.0),
    (mpm_amd.MATERIAL_SNOW, 1.0 / 10000, 200.0, 1.4e5, 0.2, 0.0, 0.0, 1e30, 0.0),
    (mpm_amd.MATERIAL_JELLY, 3e-7, 1000.0, 1e5, 0.3, 10.0, 0.0, 0.0, 0.0),
    (mpm_amd.MATERIAL_WC_LIQUID, 2.16e-7, 1000.0, 0.0, 0.0, 0.0, 0.0, 0.0, 4e4),
])
def test_material_init_matches_float32_constructor(args):
    a = mpm_amd.make_material(*args)
    b = mpm_amd.material_as_numpy(*args)
    for name, _ in mpm_amd.mpm_material._fields_:
        assert getattr(a, name) == getattr(b, name), name


def test_material_init_rejects_unknown_kind():
    m = mpm_amd.mpm_material()
    rc = mpm_amd.lib().mpm_material_init(C.byref(m), 9, 1.0, 1.0, 1.0, 0.2, 0, 0, 0, 0)
    assert rc == _lib.MPM_ERR_INVALID
    assert b"kind" in mpm_amd.lib().mpm_last_error(None)


@pytest.mark.parametrize("bad", [
    dict(N=2), dict(N=2048), dict(dt=0.0), dict(dt=float("nan")), dict(boundary_mode=7),
    dict(compat=5), dict(boundary=-1.0), dict(boundary_mode=1, boundary=2.5),
])
def test_create_validates_parameters(bad):
    p = mpm_amd.Params(**{**dict(dt=1e-4, N=16), **bad}).to_c()
    h = C.c_void_p()
    rc = mpm_amd.lib().mpm_create(C.byref(p), C.byref(h))
    assert rc == _lib.MPM_ERR_INVALID
    assert not h.value
    assert mpm_amd.lib().mpm_last_error(None)


def test_null_handle_is_an_error():
    lib = mpm_amd.lib()
    assert lib.mpm_advance(None, 1, None) == _lib.MPM_ERR_INVALID
    assert b"null handle" in lib.mpm_last_error(None)
    assert lib.mpm_destroy(None) == 0


def test_create_without_gpu_reports_nogpu():
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    p = mpm_amd.Params(dt=1e-4, N=16).to_c()
    h = C.c_void_p()
    rc = mpm_amd.lib().mpm_create(C.byref(p), C.byref(h))
    assert rc == _lib.MPM_ERR_NOGPU
    assert mpm_amd.lib().mpm_last_error(None)


def test_debug_svd_rejects_bad_arguments():
    assert mpm_amd.lib().mpm_debug_svd(-1, None, None, None, None) == _lib.MPM_ERR_INVALID
    assert mpm_amd.lib().mpm_debug_svd(0, None, None, None, None) == 0


def test_engine_shares_torchs_hip_runtime():
    # one HIP runtime per process: the engine binds the copy torch ships
    import torch
    mpm_amd.lib()
    maps = Path("/proc/self/maps").read_text()
    libs = {line.split()[-1] for line in maps.splitlines() if "libamdhip64" in line}
    torch_dir = str(Path(torch.__file__).resolve().parent)
    assert libs and all(p.startswith(torch_dir) for p in libs), libs


def test_diff_sizes():
    p = mpm_amd.Params(dt=1e-4, N=32, compat=_lib.COMPAT_FIXED).to_c()
    g, t = C.c_int64(), C.c_int64()
    assert mpm_amd.lib().mpm_diff_sizes(C.byref(p), C.byref(g), C.byref(t)) == 0
    assert (g.value, t.value) == (4 * 32 ** 3, 8 * 32 ** 3)


@pytest.mark.parametrize("bad", [
    dict(compat=_lib.COMPAT_REFERENCE), dict(N=2), dict(mass=0.0), dict(volume=float("nan")),
    dict(n=-1), dict(null_grid=True),
])
def test_diff_substep_validates_arguments(bad):
    # rejected before any device work: every pointer is null or a dummy host address
    p = mpm_amd.Params(dt=1e-4, N=bad.get("N", 16),
                       compat=bad.get("compat", _lib.COMPAT_FIXED)).to_c()
    dummy = C.c_void_p(8)
    grid = None if bad.get("null_grid") else dummy
    rc = mpm_amd.lib().mpm_diff_substep(C.byref(p), bad.get("mass", 1.0), bad.get("volume", 1.0),
                                        bad.get("n", 0), None, None, None, None, grid, dummy, None)
    assert rc == _lib.MPM_ERR_INVALID
    assert mpm_amd.lib().mpm_last_error(None)
    rc = mpm_amd.lib().mpm_diff_substep_backward(
        C.byref(p), bad.get("mass", 1.0), bad.get("volume", 1.0), bad.get("n", 0), None, None, None,
        dummy, None, None, None, None, grid, None)
    assert rc == _lib.MPM_ERR_INVALID


def test_diff_substep_rejects_aliased_output():
    p = mpm_amd.Params(dt=1e-4, N=16, compat=_lib.COMPAT_FIXED).to_c()
    d = C.c_void_p(8)
    rc = mpm_amd.lib().mpm_diff_substep(C.byref(p), 1.0, 1.0, 4, d, d, d, d, d, d, None)
    assert rc == _lib.MPM_ERR_INVALID
    assert b"alias" in mpm_amd.lib().mpm_last_error(None)


def test_diff_substep_rejects_misaligned_grid_and_tape():
    p = mpm_amd.Params(dt=1e-4, N=16, compat=_lib.COMPAT_FIXED).to_c()
    a, b, odd = C.c_void_p(256), C.c_void_p(512), C.c_void_p(1032)
    lib = mpm_amd.lib()
    assert lib.mpm_diff_substep(C.byref(p), 1.0, 1.0, 4, a, b, a, a, odd, a, None) == _lib.MPM_ERR_INVALID
    assert b"16-byte aligned" in lib.mpm_last_error(None)
    assert lib.mpm_diff_substep_noalloc(C.byref(p), 1.0, 1.0, 4, a, b, a, a, odd, None) == _lib.MPM_ERR_INVALID
    assert b"16-byte aligned" in lib.mpm_last_error(None)
