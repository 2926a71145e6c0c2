"""Pins of the oracle's interpolation-kernel plugin (CPU): the reference's
quadratic B-spline with its constant D^-1 (InterpolationKernel.cuh:55-74) and
a cubic B-spline whose D^-1 goes through the generic per-particle inverse
(InterpolationKernelBase::D_inv, InterpolationKernel.cuh:21-50).

The reference ships no kernel besides the quadratic one, so the cubic
kernel's parity is pinned by mathematics, not by the reference:
  * B-spline values at known points (N(0) = 2/3, N(1/2) = 23/48, N(1) = 1/6,
    N(3/2) = 1/48, N(2) = 0);
  * partition of unity, zero first moment and second moment dx^2/4 I
    (quadratic) or dx^2/3 I (cubic) at interior positions;
  * the generic D^-1 equals the closed form 3/dx^2 I for the cubic kernel,
    and for the quadratic kernel the reference's own constant 4/dx^2 I
    (InterpolationKernel.cuh:71-73);
  * P2G conserves mass, momentum and angular momentum (APIC term m dx^2/3
    axial(C));
  * an affine velocity field is reproduced exactly by G2P;
  * a textbook MLS-MPM substep with the cubic B-spline and the analytic D^-1
    (numpy, dense over the grid) against the FP64 oracle.
"""
import math

import numpy as np
import pytest

import oracle
from mpm_amd import COMPAT_REFERENCE, MATERIAL_SNOW, Params, material_as_numpy
from test_oracle_step import ELASTIC, SNOW, axial, interior_state

QUAD, CUBIC = oracle.INTERP_QUADRATIC, oracle.INTERP_CUBIC


def bspline3(t):
    t = np.abs(t)
    return np.where(t < 1, 0.5 * t ** 3 - t ** 2 + 2.0 / 3.0,
                    np.where(t < 2, (2 - t) ** 3 / 6.0, 0.0))


def test_cubic_bspline_values(oracle_lib):
    N = 16
    p = Params(N=N)
    dx = 1.0 / N
    # x/dx = 5.5: base = int(4.5) = 4, nodes 4..7 at distances 1.5, 0.5, 0.5, 1.5
    # x/dx = 5.0: base 4, distances 1, 0, 1, 2
    base, w, _ = oracle.kernel(p, np.array([[5.5 * dx, 5.0 * dx, 5.25 * dx]]), CUBIC)
    assert list(base[0]) == [4, 4, 4]
    assert np.allclose(w[0, 0], [1 / 48, 23 / 48, 23 / 48, 1 / 48], rtol=0, atol=1e-15)
    assert np.allclose(w[0, 1], [1 / 6, 2 / 3, 1 / 6, 0.0], rtol=0, atol=1e-15)
    t = np.abs(5.25 - np.array([4, 5, 6, 7]))
    assert np.allclose(w[0, 2], bspline3(t), rtol=0, atol=1e-15)


@pytest.mark.parametrize("interp,second", [(QUAD, 0.25), (CUBIC, 1.0 / 3.0)])
def test_partition_of_unity_and_moments(oracle_lib, interp, second):
    N = 32
    p = Params(N=N)
    dx = 1.0 / N
    rng = np.random.default_rng(3)
    x = rng.uniform(0.2, 0.8, (2000, 3))
    base, w, _ = oracle.kernel(p, x, interp)
    ns = 3 if interp == QUAD else 4
    assert np.all(w[:, :, ns:] == 0)
    d = (base[:, :, None] + np.arange(4)[None, None, :]) * dx - x[:, :, None]  # node - particle
    assert np.abs(w.sum(-1) - 1).max() < 1e-14
    assert np.abs((w * d).sum(-1)).max() < 1e-14
    assert np.abs((w * d * d).sum(-1) / dx ** 2 - second).max() < 1e-12


@pytest.mark.parametrize("interp,closed", [(QUAD, 4.0), (CUBIC, 3.0)])
def test_generic_dinv_matches_closed_form(oracle_lib, interp, closed):
    N = 64
    p = Params(N=N)
    dx = 1.0 / N
    rng = np.random.default_rng(4)
    x = rng.uniform(0.1, 0.9, (1000, 3))
    _, _, D = oracle.kernel(p, x, interp, generic=True)
    ref = closed / dx ** 2 * np.eye(3)
    assert np.abs(D - ref).max() / (closed / dx ** 2) < 1e-10
    if interp == QUAD:  # the constant the reference uses for its quadratic kernel
        _, _, Dc = oracle.kernel(p, x, interp, generic=False)
        assert np.abs(Dc - ref).max() / (4.0 / dx ** 2) < 1e-15


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 2e-4)])
def test_cubic_p2g_conserves_mass_and_momenta(oracle_lib, dtype, tol):
    N = 32
    st = interior_state(400, 5, N)
    p = Params(dt=1e-4, N=N, gravity=0.0, boundary=0.0)
    g = oracle.p2g(p, [SNOW], st, dtype, interp=CUBIC).astype(np.float64)
    m = float(np.float32(SNOW.particle_mass)) if dtype == np.float32 else SNOW.particle_mass
    n = st["x"].shape[0]
    assert abs(g[..., 3].sum() - n * m) / (n * m) < tol
    mom = g[..., :3].reshape(-1, 3).sum(0)
    ref = m * st["v"].sum(0)
    assert np.abs(mom - ref).max() / np.abs(ref).max() < tol
    dx = 1.0 / N
    idx = np.stack(np.meshgrid(*(np.arange(N),) * 3, indexing="ij"), -1).reshape(-1, 3) * dx
    L = np.cross(idx, g[..., :3].reshape(-1, 3)).sum(0)
    Lref = (np.cross(st["x"], m * st["v"]) + m * dx * dx / 3.0 * axial(st["C"])).sum(0)
    assert np.abs(L - Lref).max() / np.abs(Lref).max() < tol


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 2e-5)])
def test_cubic_affine_velocity_field_is_reproduced(oracle_lib, dtype, tol):
    N = 32
    rng = np.random.default_rng(6)
    n = 3000
    x = rng.uniform(0.25, 0.75, size=(n, 3))
    b = np.array([0.3, -0.2, 0.1])
    A = rng.uniform(-2, 2, size=(3, 3))
    st = {"x": x, "v": b + x @ A.T, "C": np.broadcast_to(A, (n, 3, 3)).copy(),
          "F": np.broadcast_to(np.eye(3), (n, 3, 3)).copy(), "Jp": np.ones(n),
          "material": np.zeros(n, np.uint8)}
    p = Params(dt=1e-5, N=N, gravity=0.0, boundary=0.0)
    out, grid = oracle.advance(p, [ELASTIC], st, 1, dtype, return_grid=True, interp=CUBIC)
    has = grid[..., 3] > 0
    ii = np.argwhere(has) / N
    gv = grid[has][:, :3].astype(np.float64)
    assert np.abs(gv - (b + ii @ A.T)).max() < tol * 10
    assert np.abs(out["v"] - (b + x @ A.T)).max() < tol * 10
    assert np.abs(out["C"] - A).max() / np.abs(A).max() < tol * 10


def textbook_cubic_step(st, mat, N, dt, g):
    """One MLS-MPM substep (Hu et al. 2018) with the cubic B-spline and its
    analytic D^-1 = 3/dx^2 I, dense over the grid, float64."""
    dx = 1.0 / N
    Dinv = 3.0 / (dx * dx)
    nodes = np.stack(np.meshgrid(*(np.arange(N),) * 3, indexing="ij"), -1).reshape(-1, 3) * dx
    x, v, C, F, Jp = (st[k].astype(np.float64) for k in ("x", "v", "C", "F", "Jp"))
    n = x.shape[0]
    mass = np.zeros(N ** 3)
    mom = np.zeros((N ** 3, 3))
    W = np.zeros((n, N ** 3))
    for p in range(n):
        d = nodes - x[p]
        W[p] = bspline3(d[:, 0] / dx) * bspline3(d[:, 1] / dx) * bspline3(d[:, 2] / dx)
        U, s, Vt = np.linalg.svd(F[p])
        if np.linalg.det(U) * np.linalg.det(Vt) < 0:
            U[:, 2] *= -1
        R = U @ Vt
        e = math.exp(mat.hardening * (1.0 - Jp[p]))
        tau = 2 * mat.mu0 * e * (F[p] - R) @ F[p].T + mat.lambda0 * e * (Jp[p] - 1) * Jp[p] * np.eye(3)
        aff = -dt * mat.particle_volume * Dinv * tau + mat.particle_mass * C[p]
        mass += W[p] * mat.particle_mass
        mom += W[p][:, None] * (mat.particle_mass * v[p] + d @ aff.T)
    vel = np.zeros_like(mom)
    has = mass > 0
    vel[has] = mom[has] / mass[has][:, None]
    vel[has, 1] += dt * g
    out = {k: st[k].astype(np.float64).copy() for k in st}
    for p in range(n):
        d = nodes - x[p]
        vp = (W[p][:, None] * vel).sum(0)
        Cp = Dinv * (W[p][:, None, None] * vel[:, :, None] * d[:, None, :]).sum(0)
        Fp = (np.eye(3) + dt * Cp) @ F[p]
        U, s, Vt = np.linalg.svd(Fp)
        if np.linalg.det(U) * np.linalg.det(Vt) < 0:
            U[:, 2] *= -1
            s[2] *= -1
        sc = np.clip(s, mat.clamp_lo, mat.clamp_hi)
        Fn = U @ np.diag(sc) @ Vt
        out["Jp"][p] = np.clip(Jp[p] * np.linalg.det(Fp) / np.linalg.det(Fn), float(np.float32(0.6)), 20.0)
        out["F"][p] = Fn
        out["v"][p] = vp
        out["C"][p] = Cp
        out["x"][p] = x[p] + dt * vp
    return out


@pytest.mark.parametrize("mat", [SNOW, ELASTIC], ids=["snow", "elastic"])
def test_textbook_cubic_substep(oracle_lib, mat):
    N = 10
    rng = np.random.default_rng(9)
    n = 12
    st = {"x": rng.uniform(0.35, 0.65, (n, 3)), "v": rng.uniform(-1, 1, (n, 3)),
          "C": rng.uniform(-2, 2, (n, 3, 3)),
          "F": np.eye(3) + 0.02 * rng.uniform(-1, 1, (n, 3, 3)), "Jp": rng.uniform(0.9, 1.1, n),
          "material": np.zeros(n, np.uint8)}
    dt = float(np.float32(1e-4))
    g = float(np.float32(-9.8))
    p = Params(dt=1e-4, N=N, gravity=-9.8, boundary=0.0, compat=COMPAT_REFERENCE)
    mat64 = material_as_numpy(mat.kind, mat.particle_volume, mat.particle_mass / mat.particle_volume)
    for f in ("mu0", "lambda0", "hardening", "clamp_lo", "clamp_hi", "particle_mass", "particle_volume"):
        setattr(mat64, f, getattr(mat, f))
    ref = textbook_cubic_step(st, mat64, N, dt, g)
    out = oracle.advance(p, [mat64], st, 1, np.float64, interp=CUBIC)
    for k in ("x", "v", "C", "F", "Jp"):
        err = np.abs(out[k] - ref[k]).max() / max(np.abs(ref[k]).max(), 1e-30)
        assert err < 1e-9, (k, err)


def test_cubic_differs_from_quadratic(oracle_lib):
    """The kernel choice reaches the step (a plugin that is silently ignored
    would pass every test above on the quadratic path)."""
    N = 16
    st = interior_state(50, 7, N)
    p = Params(dt=1e-4, N=N, gravity=0.0, boundary=0.0)
    a = oracle.advance(p, [ELASTIC], st, 1, np.float64, interp=QUAD)
    b = oracle.advance(p, [ELASTIC], st, 1, np.float64, interp=CUBIC)
    assert np.abs(a["v"] - b["v"]).max() > 1e-3 * np.abs(a["v"]).max()
