"""Pins of the oracle's substep (CPU).

Independent checks, each chosen so that a plausible slip (a dropped term, a
wrong index or sign, a transposed operand, a wrong D^-1) fails one of them:
  * P2G conserves mass, linear momentum and (with the APIC term) angular
    momentum: sum_i x_i x p_i = sum_p [x_p x m v_p + m dx^2/4 axial(C_p)];
  * an affine velocity field (v = b + A x, C = A) survives a substep exactly
    (partition of unity, zero first moment, second moment dx^2/4 I);
  * free fall: v = dt g, C = 0 after one substep;
  * wall rule counts for N = 16 and N = 128 (fraction 0.05) and cell layers;
  * closed forms of the Kirchhoff stress and of the plasticity return map on
    F = Q diag(s) P^T;
  * the reference's index quirks (frozen particles beyond N^3, aliased upper
    rows, truncating base cell, out-of-domain particles);
  * a textbook MLS-MPM substep (Hu et al. 2018) written in numpy from the
    piecewise quadratic B-spline, against the FP64 oracle.
"""
import math

import numpy as np
import pytest

import oracle
from mpm_amd import (BOUNDARY_CELLS, BOUNDARY_FRACTION, COMPAT_FIXED, COMPAT_REFERENCE,
                     MATERIAL_FIXED_COROTATED, MATERIAL_JELLY, MATERIAL_SNOW, MATERIAL_WC_LIQUID,
                     Params, material_as_numpy)


def random_rotations(n, rng):
    Q = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    d = np.linalg.det(Q)
    Q[:, :, 0] *= d[:, None]
    return Q


def interior_state(n, seed, N=32, F_noise=0.1, C_scale=5.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.3, 0.7, size=(n, 3))
    v = rng.uniform(-1, 1, size=(n, 3))
    C = rng.uniform(-C_scale, C_scale, size=(n, 3, 3))
    F = np.eye(3) + F_noise * rng.uniform(-1, 1, size=(n, 3, 3))
    return {"x": x, "v": v, "C": C, "F": F, "Jp": np.ones(n), "material": np.zeros(n, np.uint8)}


SNOW = material_as_numpy(MATERIAL_SNOW, 2e-6, 400.0, E=1.4e5, nu=0.2, hardening=10.0,
                         clamp_lo=0.975, clamp_hi=1.0075)
ELASTIC = material_as_numpy(MATERIAL_SNOW, 2e-6, 1000.0, E=1e4, nu=0.3, hardening=0.0,
                            clamp_lo=0.0, clamp_hi=1e30)
LIQUID = material_as_numpy(MATERIAL_WC_LIQUID, 2e-6, 1000.0, E=0, nu=0, hardening=0,
                           clamp_lo=0, clamp_hi=0, bulk_modulus=4e4)


def axial(C):
    # (C21 - C12, C02 - C20, C10 - C01) for each matrix
    return np.stack([C[:, 2, 1] - C[:, 1, 2], C[:, 0, 2] - C[:, 2, 0], C[:, 1, 0] - C[:, 0, 1]], -1)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 2e-4)])
def test_p2g_conserves_mass_and_momenta(oracle_lib, dtype, tol):
    N = 32
    st = interior_state(400, 1, N)
    p = Params(dt=1e-4, N=N, gravity=0.0, boundary=0.0)
    g = oracle.p2g(p, [SNOW], st, dtype).astype(np.float64)
    m = float(np.float32(SNOW.particle_mass)) if dtype == np.float32 else SNOW.particle_mass
    n = st["x"].shape[0]
    assert abs(g[..., 3].sum() - n * m) / (n * m) < tol
    mom = g[..., :3].reshape(-1, 3).sum(0)
    ref = m * st["v"].sum(0)
    assert np.abs(mom - ref).max() / np.abs(ref).max() < tol
    dx = 1.0 / N
    idx = np.stack(np.meshgrid(*(np.arange(N),) * 3, indexing="ij"), -1).reshape(-1, 3) * dx
    L = np.cross(idx, g[..., :3].reshape(-1, 3)).sum(0)
    Lref = (np.cross(st["x"], m * st["v"]) + m * dx * dx / 4.0 * axial(st["C"])).sum(0)
    assert np.abs(L - Lref).max() / np.abs(Lref).max() < tol


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 2e-5)])
def test_affine_velocity_field_is_reproduced(oracle_lib, dtype, tol):
    N = 32
    rng = np.random.default_rng(2)
    n = 3000
    x = rng.uniform(0.25, 0.75, size=(n, 3))
    b = np.array([0.3, -0.2, 0.1])
    A = rng.uniform(-2, 2, size=(3, 3))
    st = {"x": x, "v": b + x @ A.T, "C": np.broadcast_to(A, (n, 3, 3)).copy(),
          "F": np.broadcast_to(np.eye(3), (n, 3, 3)).copy(), "Jp": np.ones(n),
          "material": np.zeros(n, np.uint8)}
    p = Params(dt=1e-5, N=N, gravity=0.0, boundary=0.0)
    out, grid = oracle.advance(p, [ELASTIC], st, 1, dtype, return_grid=True)
    # grid: every node with mass carries v = b + A x_node
    has = grid[..., 3] > 0
    ii = np.argwhere(has) / N
    gv = grid[has][:, :3].astype(np.float64)
    assert np.abs(gv - (b + ii @ A.T)).max() < tol * 10
    # particles: v and C of the field at the (old) particle position
    assert np.abs(out["v"] - (b + x @ A.T)).max() < tol * 10
    assert np.abs(out["C"] - A).max() / np.abs(A).max() < tol * 10


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-6)])
def test_free_fall(oracle_lib, dtype, tol):
    n = 500
    rng = np.random.default_rng(3)
    st = {"x": rng.uniform(0.3, 0.7, (n, 3)), "v": np.zeros((n, 3)), "C": np.zeros((n, 3, 3)),
          "F": np.broadcast_to(np.eye(3), (n, 3, 3)).copy(), "Jp": np.ones(n),
          "material": np.zeros(n, np.uint8)}
    p = Params(dt=1e-4, N=32, gravity=-9.8, boundary=0.0)
    out = oracle.advance(p, [ELASTIC], st, 1, dtype)
    gdt = np.float64(np.float32(1e-4)) * np.float64(np.float32(-9.8))
    assert np.abs(out["v"][:, 1] - gdt).max() < tol * 10
    assert np.abs(out["v"][:, [0, 2]]).max() < tol
    assert np.abs(out["C"]).max() < tol * 32 * 32
    assert np.abs(out["x"] - (st["x"] + 1e-4 * out["v"])).max() < tol


def wall_layers(N, mode, b):
    """Count sticky layers at the low/high end of the x axis and floor layers
    in y, from the grid update of a uniform field."""
    g = np.zeros((N, N, N, 4))
    g[..., 0] = 1.0
    g[..., 1] = -1.0
    g[..., 2] = 1.0
    g[..., 3] = 1.0
    p = Params(dt=1e-4, N=N, gravity=0.0, boundary_mode=mode, boundary=b)
    out = oracle.grid_update(p, g, np.float64)
    mid = N // 2
    xs = out[:, mid, mid, 0]
    low = int(np.argmax(xs != 0))
    high = int(np.argmax(xs[::-1] != 0))
    floor = int(np.argmax(out[mid, :, mid, 1] != 0))  # v_y clamped to >= 0 -> 0 here
    return low, high, floor


def test_wall_layers(oracle_lib):
    # reference fraction 0.05 on index/N: N=128 -> 7 low, 6 high; N=16 -> 1 low, 0 high
    assert wall_layers(128, BOUNDARY_FRACTION, 0.05) == (7, 6, 7)
    assert wall_layers(16, BOUNDARY_FRACTION, 0.05) == (1, 0, 1)
    assert wall_layers(64, BOUNDARY_FRACTION, 0.05) == (4, 3, 4)
    assert wall_layers(32, BOUNDARY_CELLS, 3) == (3, 3, 3)


def test_grid_update_closed_form(oracle_lib):
    N = 16
    rng = np.random.default_rng(4)
    g = np.zeros((N, N, N, 4))
    g[..., 3] = rng.uniform(0.5, 2.0, (N, N, N))
    g[..., :3] = rng.uniform(-1, 1, (N, N, N, 3))
    g[3, 3, 3, 3] = 0.0
    p = Params(dt=1e-3, N=N, gravity=-9.8, boundary_mode=BOUNDARY_CELLS, boundary=2)
    out = oracle.grid_update(p, g, np.float64)
    gdt = np.float64(np.float32(1e-3)) * np.float64(np.float32(-9.8))
    i = (slice(2, N - 2), slice(2, N - 2), slice(2, N - 2))
    keep = g[i][..., 3] > 0
    exp = g[i][..., :3] / np.where(keep, g[i][..., 3], 1.0)[..., None]
    exp[..., 1] += gdt
    assert np.allclose(out[i][..., :3][keep], exp[keep], rtol=1e-14, atol=1e-15)
    assert np.all(out[i][..., 3][keep] == 1.0)
    assert np.all(out[3, 3, 3] == g[3, 3, 3])  # massless node untouched


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 2e-3)])
def test_stress_closed_form(oracle_lib, dtype, tol):
    rng = np.random.default_rng(5)
    n = 200
    Q = random_rotations(n, rng)
    d = rng.uniform(0.8, 1.25, (n, 3))
    F = Q * d[:, None, :]  # Q diag(d)
    Jp = rng.uniform(0.8, 1.2, n)
    for mat in (SNOW, ELASTIC):
        tau = oracle.stress(mat, F, Jp, COMPAT_REFERENCE, dtype).astype(np.float64)
        e = np.exp(mat.hardening * (1.0 - Jp))
        mu, lam = mat.mu0 * e, mat.lambda0 * e
        dev = Q @ (((d - 1.0) * d)[:, :, None] * np.swapaxes(Q, 1, 2))
        exp = 2.0 * mu[:, None, None] * dev + (lam * (Jp - 1.0) * Jp)[:, None, None] * np.eye(3)
        scale = np.abs(exp).max()
        assert np.abs(tau - exp).max() / scale < tol
    # fixed-corotated with compat FIXED: volumetric term on J = det F
    fc = material_as_numpy(MATERIAL_FIXED_COROTATED, 2e-6, 1000.0, E=1e4, nu=0.3)
    tau = oracle.stress(fc, F, Jp, COMPAT_FIXED, dtype).astype(np.float64)
    J = d.prod(1)
    dev = Q @ (((d - 1.0) * d)[:, :, None] * np.swapaxes(Q, 1, 2))
    exp = 2.0 * fc.mu0 * dev + (fc.lambda0 * (J - 1.0) * J)[:, None, None] * np.eye(3)
    assert np.abs(tau - exp).max() / np.abs(exp).max() < tol
    # rotations carry no stress with Jp = 1
    tau0 = oracle.stress(SNOW, Q, np.ones(n), COMPAT_REFERENCE, dtype).astype(np.float64)
    assert np.abs(tau0).max() < (1e-9 if dtype == np.float64 else 0.5)
    # liquid: tau = K (J - 1) J I with J held in Jp
    tl = oracle.stress(LIQUID, F, Jp, COMPAT_REFERENCE, dtype).astype(np.float64)
    assert np.allclose(tl, (LIQUID.bulk_modulus * (Jp - 1) * Jp)[:, None, None] * np.eye(3),
                       rtol=1e-6, atol=1e-6 * LIQUID.bulk_modulus)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 2e-3)])
def test_plasticity_closed_form(oracle_lib, dtype, tol):
    rng = np.random.default_rng(6)
    n = 300
    Q, P = random_rotations(n, rng), random_rotations(n, rng)
    s = np.sort(rng.uniform(0.9, 1.05, (n, 3)), axis=1)[:, ::-1]
    F = Q @ (s[:, :, None] * np.swapaxes(P, 1, 2))
    Jp = rng.uniform(0.7, 1.3, n)
    Fo, Jo = oracle.plasticity(SNOW, F, Jp, dtype)
    sc = np.clip(s, SNOW.clamp_lo, SNOW.clamp_hi)
    Fexp = Q @ (sc[:, :, None] * np.swapaxes(P, 1, 2))
    # the reference clamps with real(0.6), real(20.0) (float constants)
    Jexp = np.clip(Jp * s.prod(1) / sc.prod(1), float(np.float32(0.6)), 20.0)
    assert np.abs(Fo - Fexp).max() < tol
    assert np.abs(Jo - Jexp).max() / Jexp.max() < tol
    # liquid: F <- J^(1/3) I, Jp <- det F
    Fl, Jl = oracle.plasticity(LIQUID, F, Jp, dtype)
    J = np.linalg.det(F)
    assert np.allclose(Jl, J, rtol=1e-5)
    assert np.allclose(Fl, np.cbrt(J)[:, None, None] * np.eye(3), rtol=1e-5, atol=1e-6)
    # jelly: F untouched, Jp clamped
    jelly = material_as_numpy(MATERIAL_JELLY, 2e-6, 1000.0, E=1e5, nu=0.3, hardening=10.0)
    Fj, Jj = oracle.plasticity(jelly, F, np.array([0.5, 1.0, 30.0] * 100), dtype)
    assert np.array_equal(Fj, F.astype(dtype))
    assert np.allclose(Jj, np.array([0.6, 1.0, 20.0] * 100), rtol=1e-6)


def one_particle(x, v=(0, 0, 0)):
    return {"x": np.array([x], np.float64), "v": np.array([v], np.float64),
            "C": np.zeros((1, 3, 3)), "F": np.eye(3)[None].copy(), "Jp": np.ones(1),
            "material": np.zeros(1, np.uint8)}


def test_quirk_frozen_beyond_n_cubed(oracle_lib):
    # mpm.cu:114: G2P skips particles whose index exceeds N^3 (compat REFERENCE)
    N = 4
    n = 100
    rng = np.random.default_rng(7)
    st = {"x": rng.uniform(0.4, 0.6, (n, 3)), "v": np.zeros((n, 3)), "C": np.zeros((n, 3, 3)),
          "F": np.broadcast_to(np.eye(3), (n, 3, 3)).copy(), "Jp": np.ones(n),
          "material": np.zeros(n, np.uint8)}
    p = Params(dt=1e-3, N=N, gravity=-9.8, boundary=0.0)
    out = oracle.advance(p, [ELASTIC], st, 1, np.float64)
    assert np.array_equal(out["v"][N ** 3 + 1:], st["v"][N ** 3 + 1:])
    assert np.all(out["v"][: N ** 3 + 1, 1] < 0)
    p.compat = COMPAT_FIXED
    out = oracle.advance(p, [ELASTIC], st, 1, np.float64)
    assert np.all(out["v"][:, 1] < 0)


def test_quirk_upper_rows_alias(oracle_lib):
    # mpm.cu:42-44: P2G bounds j and k by N + base, so a stencil reaching j = N
    # lands in row i+1, j = 0 of the linear index (compat REFERENCE)
    N = 8
    dx = 1.0 / N
    x = (3.5 * dx, (N - 0.25) * dx, 3.5 * dx)  # base_y = trunc(N - 0.75) = N - 1
    p = Params(dt=1e-4, N=N, gravity=0.0, boundary=0.0)
    g = oracle.p2g(p, [ELASTIC], one_particle(x), np.float64)
    total = ELASTIC.particle_mass
    assert abs(g[..., 3].sum() - total) / total < 1e-12  # nothing dropped, some aliased
    assert g[4, 0, :, 3].sum() > 0 and g[4, 1, :, 3].sum() > 0  # rows j = N, N+1 of i = 3
    p.compat = COMPAT_FIXED
    g2 = oracle.p2g(p, [ELASTIC], one_particle(x), np.float64)
    assert g2[..., 3].sum() < total  # clipped at the upper face
    assert g2[4, 0, :, 3].sum() == 0


def test_quirk_truncating_base_cell(oracle_lib):
    # InterpolationKernel.cuh:59 casts x/dx - 0.5 to int (toward zero): a
    # particle at 0.2 dx keeps base 0 and weights 0.5*1.3^2, 0.75-0.8^2, 0.5*0.3^2
    N = 8
    dx = 1.0 / N
    x = (0.2 * dx, 4.5 * dx, 4.5 * dx)
    p = Params(dt=1e-4, N=N, gravity=0.0, boundary=0.0)
    g = oracle.p2g(p, [ELASTIC], one_particle(x), np.float64)
    wx = g[..., 3].sum(axis=(1, 2)) / ELASTIC.particle_mass
    assert np.allclose(wx[:3], [0.5 * 1.3 ** 2, 0.75 - 0.8 ** 2, 0.5 * 0.3 ** 2], rtol=1e-12)
    assert np.all(wx[3:] == 0)


def test_quirk_out_of_domain_frozen(oracle_lib):
    N = 8
    st = one_particle((1.6, 0.5, 0.5), (1.0, 2.0, 3.0))
    p = Params(dt=1e-4, N=N, gravity=-9.8, boundary=0.0)
    out = oracle.advance(p, [ELASTIC], st, 3, np.float64)
    for k in ("x", "v", "C", "F", "Jp"):
        assert np.array_equal(out[k], st[k])


# --- textbook MLS-MPM (Hu et al. 2018, Alg. 1) in numpy, float64 -------------
def bspline(t):
    t = np.abs(t)
    return np.where(t < 0.5, 0.75 - t * t, np.where(t < 1.5, 0.5 * (1.5 - t) ** 2, 0.0))


def textbook_step(st, mat, N, dt, g):
    dx = 1.0 / N
    nodes = np.stack(np.meshgrid(*(np.arange(N),) * 3, indexing="ij"), -1).reshape(-1, 3) * dx
    x, v, C, F, Jp = (st[k].astype(np.float64) for k in ("x", "v", "C", "F", "Jp"))
    n = x.shape[0]
    mass = np.zeros(N ** 3)
    mom = np.zeros((N ** 3, 3))
    W = np.zeros((n, N ** 3))
    for p in range(n):
        d = nodes - x[p]
        W[p] = bspline(d[:, 0] / dx) * bspline(d[:, 1] / dx) * bspline(d[:, 2] / dx)
        U, s, Vt = np.linalg.svd(F[p])
        if np.linalg.det(U) * np.linalg.det(Vt) < 0:
            U[:, 2] *= -1
            s[2] *= -1
        R = U @ Vt
        e = math.exp(mat.hardening * (1.0 - Jp[p]))
        tau = 2 * mat.mu0 * e * (F[p] - R) @ F[p].T + mat.lambda0 * e * (Jp[p] - 1) * Jp[p] * np.eye(3)
        aff = -dt * mat.particle_volume * 4.0 / (dx * dx) * tau + mat.particle_mass * C[p]
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
        Cp = 4.0 / (dx * dx) * (W[p][:, None, None] * vel[:, :, None] * d[:, None, :]).sum(0)
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
def test_textbook_mls_mpm_substep(oracle_lib, mat):
    N = 8
    rng = np.random.default_rng(8)
    n = 12
    st = {"x": rng.uniform(0.35, 0.65, (n, 3)), "v": rng.uniform(-1, 1, (n, 3)),
          "C": rng.uniform(-2, 2, (n, 3, 3)),
          "F": np.eye(3) + 0.02 * rng.uniform(-1, 1, (n, 3, 3)), "Jp": rng.uniform(0.9, 1.1, n),
          "material": np.zeros(n, np.uint8)}
    dt = float(np.float32(1e-4))
    g = float(np.float32(-9.8))
    # compat REFERENCE: the volumetric term uses Jp, as the textbook step above
    p = Params(dt=1e-4, N=N, gravity=-9.8, boundary=0.0, compat=COMPAT_REFERENCE)
    mat64 = material_as_numpy(mat.kind, mat.particle_volume, mat.particle_mass / mat.particle_volume)
    for f in ("mu0", "lambda0", "hardening", "clamp_lo", "clamp_hi", "particle_mass", "particle_volume"):
        setattr(mat64, f, getattr(mat, f))
    ref = textbook_step(st, mat64, N, dt, g)
    out = oracle.advance(p, [mat64], st, 1, np.float64)
    for k in ("x", "v", "C", "F", "Jp"):
        err = np.abs(out[k] - ref[k]).max() / max(np.abs(ref[k]).max(), 1e-30)
        assert err < 1e-9, (k, err)
