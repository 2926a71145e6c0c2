"""CPU oracle of the MLS-MPM substep -- TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/_build/libmpm_oracle.so (built from mpm_oracle.hpp by
oracle/Makefile). Imported by tests/, __graft_entry__.smoke() and the CPU
baseline leg of bench.py, as the checker; the engine never uses it.

Every function here takes and returns the engine's SoA layout: x (n,3),
v (n,3), C (n,3,3), F (n,3,3) row-major, Jp (n,), material (n,) uint8.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libmpm_oracle.so"


class mpm_params(C.Structure):
    _fields_ = [("dt", C.c_float), ("N", C.c_uint32), ("gravity", C.c_float),
                ("boundary_mode", C.c_int32), ("boundary", C.c_float), ("compat", C.c_int32)]


class mpm_material(C.Structure):
    _fields_ = [("kind", C.c_int32), ("particle_volume", C.c_float), ("particle_mass", C.c_float),
                ("mu0", C.c_float), ("lambda0", C.c_float), ("hardening", C.c_float),
                ("clamp_lo", C.c_float), ("clamp_hi", C.c_float), ("bulk_modulus", C.c_float)]


_lib = None


def build() -> Path:
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        _lib.oracle_rsqrt_f32.restype = C.c_float
        _lib.oracle_rsqrt_f32.argtypes = [C.c_float]
        _lib.oracle_time_advance_f32.restype = C.c_double
        _lib.oracle_openmp_max_threads.restype = C.c_int
    return _lib


def _ptr(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def _conv_params(p) -> mpm_params:
    return mpm_params(p.dt, p.N, p.gravity, p.boundary_mode, p.boundary, p.compat)


def _conv_mats(mats):
    arr = (mpm_material * len(mats))()
    for i, m in enumerate(mats):
        for name, _ in mpm_material._fields_:
            setattr(arr[i], name, getattr(m, name))
    return arr


def _state_arrays(state, dtype):
    n = state["x"].shape[0]
    out = {}
    for k, shape in (("x", (n, 3)), ("v", (n, 3)), ("C", (n, 3, 3)), ("F", (n, 3, 3)), ("Jp", (n,))):
        out[k] = np.ascontiguousarray(state[k], dtype=dtype).reshape(shape).copy()
    out["material"] = np.ascontiguousarray(state["material"], dtype=np.uint8).copy()
    return out


INTERP_QUADRATIC = 0
INTERP_CUBIC = 1


def advance(params, materials, state, steps: int, dtype=np.float32, threads: int = 1,
            return_grid: bool = False, interp: int = INTERP_QUADRATIC):
    """Run `steps` reference substeps on the CPU; returns the new state dict
    (and the last grid, shape (N,N,N,4), after its grid update). `interp`
    selects the interpolation kernel (quadratic: the reference's; cubic:
    generic per-particle D^-1)."""
    s = _state_arrays(state, dtype)
    n = s["x"].shape[0]
    ct = C.c_float if dtype == np.float32 else C.c_double
    fn = lib().oracle_advance_f32 if dtype == np.float32 else lib().oracle_advance_f64
    N = params.N
    grid = np.zeros((N, N, N, 4), dtype) if return_grid else None
    cp = _conv_params(params)
    cm = _conv_mats(materials)
    rc = fn(C.byref(cp), cm, len(materials), C.c_int64(n), _ptr(s["x"], ct), _ptr(s["v"], ct),
            _ptr(s["C"], ct), _ptr(s["F"], ct), _ptr(s["Jp"], ct), _ptr(s["material"], C.c_uint8),
            int(steps), int(threads), _ptr(grid, ct), int(interp))
    if rc != 0:
        raise RuntimeError("oracle advance rejected its arguments")
    return (s, grid) if return_grid else s


def advance_lame(params, material, state, mu, lam, steps: int, dtype=np.float32, threads: int = 1):
    """`steps` substeps with every particle of `material` but its mu0, lambda0
    replaced per particle by mu[i], lam[i] (the differentiable step's
    forward: fixed-corotated with compat FIXED). Returns the new state dict.
    threads > 1 scatters with OpenMP atomics (summation order varies)."""
    s = _state_arrays(state, dtype)
    n = s["x"].shape[0]
    ct = C.c_float if dtype == np.float32 else C.c_double
    fn = lib().oracle_advance_lame_f32 if dtype == np.float32 else lib().oracle_advance_lame_f64
    lame = np.ascontiguousarray(np.stack([np.asarray(mu, dtype).reshape(n),
                                          np.asarray(lam, dtype).reshape(n)], -1))
    cp = _conv_params(params)
    cm = _conv_mats([material])
    rc = fn(C.byref(cp), cm, C.c_int64(n), _ptr(s["x"], ct), _ptr(s["v"], ct), _ptr(s["C"], ct),
            _ptr(s["F"], ct), _ptr(s["Jp"], ct), _ptr(lame, ct), int(steps), int(threads))
    if rc != 0:
        raise RuntimeError("oracle advance_lame rejected its arguments")
    return s


def p2g(params, materials, state, dtype=np.float32, interp: int = INTERP_QUADRATIC):
    s = _state_arrays(state, dtype)
    n = s["x"].shape[0]
    ct = C.c_float if dtype == np.float32 else C.c_double
    fn = lib().oracle_p2g_f32 if dtype == np.float32 else lib().oracle_p2g_f64
    N = params.N
    grid = np.zeros((N, N, N, 4), dtype)
    cp = _conv_params(params)
    rc = fn(C.byref(cp), _conv_mats(materials), len(materials), C.c_int64(n), _ptr(s["x"], ct),
            _ptr(s["v"], ct), _ptr(s["C"], ct), _ptr(s["F"], ct), _ptr(s["Jp"], ct),
            _ptr(s["material"], C.c_uint8), _ptr(grid, ct), int(interp))
    if rc != 0:
        raise RuntimeError("oracle p2g rejected its arguments")
    return grid


def kernel(params, x, interp: int = INTERP_QUADRATIC, generic: bool = False):
    """The selected kernel at positions x (n, 3), in FP64: base (n, 3) int32,
    weights (n, 3, 4) (entries past the kernel's size are 0) and D^-1 (n, 3, 3)
    (the generic inverse for the cubic kernel, or for either with generic)."""
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    n = x.shape[0]
    base = np.empty((n, 3), np.int32)
    w = np.empty((n, 3, 4), np.float64)
    D = np.empty((n, 3, 3), np.float64)
    cp = _conv_params(params)
    rc = lib().oracle_kernel_f64(C.byref(cp), int(interp), int(bool(generic)), C.c_int64(n),
                                 _ptr(x, C.c_double), _ptr(base, C.c_int32), _ptr(w, C.c_double),
                                 _ptr(D, C.c_double))
    if rc != 0:
        raise RuntimeError("oracle kernel rejected its arguments")
    return base, w, D


def grid_update(params, grid, dtype=np.float32):
    g = np.ascontiguousarray(grid, dtype=dtype).copy()
    ct = C.c_float if dtype == np.float32 else C.c_double
    fn = lib().oracle_grid_update_f32 if dtype == np.float32 else lib().oracle_grid_update_f64
    cp = _conv_params(params)
    if fn(C.byref(cp), _ptr(g, ct)) != 0:
        raise RuntimeError("oracle grid update rejected its arguments")
    return g


def svd(A, dtype=np.float32):
    A = np.ascontiguousarray(A, dtype=dtype).reshape(-1, 3, 3)
    n = A.shape[0]
    U = np.empty_like(A)
    V = np.empty_like(A)
    S = np.empty((n, 3), dtype)
    ct = C.c_float if dtype == np.float32 else C.c_double
    fn = lib().oracle_svd_f32 if dtype == np.float32 else lib().oracle_svd_f64
    fn(C.c_int64(n), _ptr(A, ct), _ptr(U, ct), _ptr(S, ct), _ptr(V, ct))
    return U, S, V


def stress(material, F, Jp, compat=0, dtype=np.float32):
    F = np.ascontiguousarray(F, dtype=dtype).reshape(-1, 3, 3)
    Jp = np.ascontiguousarray(Jp, dtype=dtype).reshape(-1)
    out = np.empty_like(F)
    ct = C.c_float if dtype == np.float32 else C.c_double
    fn = lib().oracle_stress_f32 if dtype == np.float32 else lib().oracle_stress_f64
    m = _conv_mats([material])
    fn(m, int(compat), C.c_int64(F.shape[0]), _ptr(F, ct), _ptr(Jp, ct), _ptr(out, ct))
    return out


def plasticity(material, F, Jp, dtype=np.float32):
    F = np.ascontiguousarray(F, dtype=dtype).reshape(-1, 3, 3).copy()
    Jp = np.ascontiguousarray(Jp, dtype=dtype).reshape(-1).copy()
    ct = C.c_float if dtype == np.float32 else C.c_double
    fn = lib().oracle_plasticity_f32 if dtype == np.float32 else lib().oracle_plasticity_f64
    fn(_conv_mats([material]), C.c_int64(F.shape[0]), _ptr(F, ct), _ptr(Jp, ct))
    return F, Jp


def rsqrt_f32(x: float) -> float:
    return lib().oracle_rsqrt_f32(float(x))


def time_advance_f32(params, materials, state, steps: int, threads: int) -> float:
    """Seconds for `steps` FP32 substeps with `threads` OpenMP threads."""
    s = _state_arrays(state, np.float32)
    n = s["x"].shape[0]
    cp = _conv_params(params)
    ct = C.c_float
    return lib().oracle_time_advance_f32(C.byref(cp), _conv_mats(materials), len(materials),
                                         C.c_int64(n), _ptr(s["x"], ct), _ptr(s["v"], ct),
                                         _ptr(s["C"], ct), _ptr(s["F"], ct), _ptr(s["Jp"], ct),
                                         _ptr(s["material"], C.c_uint8), int(steps), int(threads))


def openmp_threads() -> int:
    return lib().oracle_openmp_max_threads()


def check_threads() -> int:
    """Threads for oracle runs inside tests: the OpenMP budget of the host
    (OMP_NUM_THREADS, else the CPUs this process may use), at most 16."""
    import os
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return min(16, int(env))
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except (AttributeError, OSError):
        return max(1, min(16, os.cpu_count() or 1))
