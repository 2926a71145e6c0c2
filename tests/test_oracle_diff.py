"""Pins of the oracle's forward of the differentiable substep (CPU):
fixed-corotated, J = det F, compat FIXED, per-particle Lame parameters
(oracle.advance_lame).

  * with every particle's (mu, lambda) equal to the material's, it is the
    oracle's own substep with that material, bit for bit (FP32 and FP64);
  * with distinct per-particle values it matches a textbook MLS-MPM substep
    (Hu et al. 2018, Alg. 1) written in numpy with tau = 2 mu (F - R) F^T +
    lambda (J - 1) J I, so a parameter applied to the wrong particle, a
    swapped mu/lambda or J taken from anything but det F fails.
"""
import numpy as np
import pytest

import oracle
from mpm_amd import COMPAT_FIXED, MATERIAL_FIXED_COROTATED, Params, material_as_numpy, scenes

MAT = material_as_numpy(MATERIAL_FIXED_COROTATED, 2e-6, 1000.0, E=1e4, nu=0.3)


def state(n, seed):
    rng = np.random.default_rng(seed)
    return {"x": rng.uniform(0.3, 0.7, (n, 3)), "v": rng.uniform(-1, 1, (n, 3)),
            "C": rng.uniform(-3, 3, (n, 3, 3)), "F": np.eye(3) + 0.1 * rng.uniform(-1, 1, (n, 3, 3)),
            "Jp": np.ones(n), "material": np.zeros(n, np.uint8)}


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_uniform_lame_is_the_material_step(oracle_lib, dtype):
    st = state(300, 1)
    p = Params(dt=1e-4, N=16, gravity=-9.8, compat=COMPAT_FIXED)
    mu = np.full(300, MAT.mu0)
    lam = np.full(300, MAT.lambda0)
    a = oracle.advance_lame(p, MAT, st, mu, lam, 5, dtype=dtype)
    b = oracle.advance(p, [MAT], st, 5, dtype)
    for k in ("x", "v", "C", "F", "Jp"):
        assert np.array_equal(a[k], b[k]), k


def bspline(t):
    t = np.abs(t)
    return np.where(t < 0.5, 0.75 - t * t, np.where(t < 1.5, 0.5 * (1.5 - t) ** 2, 0.0))


def textbook_fc_step(st, mu, lam, mass, vol, N, dt, g):
    dx = 1.0 / N
    nodes = np.stack(np.meshgrid(*(np.arange(N),) * 3, indexing="ij"), -1).reshape(-1, 3) * dx
    x, v, C, F = (st[k].astype(np.float64) for k in ("x", "v", "C", "F"))
    n = x.shape[0]
    M = np.zeros(N ** 3)
    P = np.zeros((N ** 3, 3))
    W = np.zeros((n, N ** 3))
    for p in range(n):
        d = nodes - x[p]
        W[p] = bspline(d[:, 0] / dx) * bspline(d[:, 1] / dx) * bspline(d[:, 2] / dx)
        U, s, Vt = np.linalg.svd(F[p])
        R = U @ Vt
        if np.linalg.det(R) < 0:
            U[:, 2] *= -1
            R = U @ Vt
        J = np.linalg.det(F[p])
        tau = 2 * mu[p] * (F[p] - R) @ F[p].T + lam[p] * (J - 1) * J * np.eye(3)
        aff = -dt * vol * 4.0 / (dx * dx) * tau + mass * C[p]
        M += W[p] * mass
        P += W[p][:, None] * (mass * v[p] + d @ aff.T)
    vel = np.zeros_like(P)
    has = M > 0
    vel[has] = P[has] / M[has][:, None]
    vel[has, 1] += dt * g
    out = {}
    out["v"] = (W[:, :, None] * vel[None]).sum(1)
    out["C"] = np.stack([4.0 / (dx * dx) * (W[p][:, None, None] * vel[:, :, None] *
                                            (nodes - x[p])[:, None, :]).sum(0) for p in range(n)])
    out["F"] = np.stack([(np.eye(3) + dt * out["C"][p]) @ F[p] for p in range(n)])
    out["x"] = x + dt * out["v"]
    return out


def test_per_particle_lame_matches_textbook(oracle_lib):
    n = 12
    st = state(n, 2)
    rng = np.random.default_rng(3)
    E = np.exp(np.log(1e4) + rng.normal(size=n))
    mu = E / 2.6
    lam = E * 0.3 / (1.3 * 0.4)
    dt = float(np.float32(1e-4))
    g = float(np.float32(-9.8))
    p = Params(dt=1e-4, N=8, gravity=-9.8, boundary=0.0, compat=COMPAT_FIXED)
    ref = textbook_fc_step(st, mu, lam, MAT.particle_mass, MAT.particle_volume, 8, dt, g)
    out = oracle.advance_lame(p, MAT, st, mu, lam, 1, dtype=np.float64)
    for k in ("x", "v", "C", "F"):
        err = np.abs(out[k] - ref[k]).max() / max(np.abs(ref[k]).max(), 1e-30)
        assert err < 1e-9, (k, err)
    # the per-particle values matter: the material's uniform values give another step
    other = oracle.advance_lame(p, MAT, st, np.full(n, mu.mean()), np.full(n, lam.mean()), 1,
                                dtype=np.float64)
    assert np.abs(other["C"] - ref["C"]).max() > 1e-6 * np.abs(ref["C"]).max()


def test_diff_cube_workload():
    c = scenes.diff_cube()
    assert c.x.shape == (32768, 3) and c.logE.shape == (32768,)
    assert c.x.min() > 0.35 - 0.3 / 128 - 1e-6 and c.x.max() < 0.65 + 0.3 / 128 + 1e-6
    assert abs(float(c.logE.mean()) - np.log(1e4)) < 0.01
    assert c.params.compat == COMPAT_FIXED and c.params.N == 32 and c.substeps == 200


# ---------------------------------------------------------------- FP64 reference gradients
def dropped_cube(side):
    # a jittered cube just above the floor walls, so the stiffness matters
    return scenes.diff_cube(side=side, lo=(0.45, 0.08, 0.45), spacing=0.009375)


def oracle_loss(c, x0, v0, logE, steps):
    E = np.exp(np.asarray(logE, np.float64))
    mu, lam = E / (2.0 * (1.0 + c.nu)), E * c.nu / ((1.0 + c.nu) * (1.0 - 2.0 * c.nu))
    n = x0.shape[0]
    st = {"x": np.asarray(x0, np.float64), "v": np.tile(np.asarray(v0, np.float64), (n, 1)),
          "C": np.zeros((n, 3, 3)), "F": np.tile(np.eye(3), (n, 1, 1)), "Jp": np.ones(n),
          "material": np.zeros(n, np.uint8)}
    m = material_as_numpy(MATERIAL_FIXED_COROTATED, c.particle_volume, 1000.0, E=1e4, nu=c.nu)
    m.particle_mass = c.particle_mass
    m.particle_volume = c.particle_volume
    out = oracle.advance_lame(c.params, m, st, mu, lam, steps, dtype=np.float64)
    return out, float(((out["x"].mean(0) - np.asarray(c.target)) ** 2).sum())


def test_torch_fp64_forward_matches_oracle(oracle_lib):
    """The autograd restatement computes the C++ FP64 oracle's step (a cube
    that reaches the floor walls within the 30 substeps), to FP64 roundoff."""
    torch = pytest.importorskip("torch")
    from oracle import diff_torch

    c = dropped_cube(6)
    v0 = np.array([0.3, -1.5, 0.2])
    steps = 30
    n = c.x.shape[0]
    E = np.exp(c.logE.astype(np.float64))
    mu, lam = E / 2.6, E * 0.3 / (1.3 * 0.4)
    T = lambda a: torch.tensor(np.asarray(a, np.float64))
    xT, vT, CT, FT = diff_torch.rollout(T(c.x), T(np.tile(v0, (n, 1))), torch.zeros((n, 3, 3), dtype=torch.float64),
                                        torch.eye(3, dtype=torch.float64).expand(n, 3, 3).clone(), T(mu), T(lam),
                                        c.particle_mass, c.particle_volume, c.params, steps)
    ref, _ = oracle_loss(c, c.x, v0, c.logE, steps)
    got = {"x": xT, "v": vT, "C": CT, "F": FT}
    for k in ("x", "v", "C", "F"):
        a = got[k].detach().numpy()
        err = np.abs(a - ref[k]).max() / np.abs(ref[k]).max()
        # FP64 roundoff (C carries D^-1 = 4 / dx^2); a wrong term shows at 1e-3 or more
        assert err < 1e-9, (k, err)


def test_torch_fp64_gradients_match_central_differences(oracle_lib):
    """dL/dv0, dL/dlogE and dL/dx0 of the autograd restatement against FP64
    central differences of the C++ oracle (eps 1e-3): the v0 components and
    the directional derivatives along random log E and x0 directions within
    1e-5 of the gradient's norm."""
    pytest.importorskip("torch")
    from oracle import diff_torch

    c = dropped_cube(5)
    v0 = np.array([0.3, -3.0, 0.2])  # reaches the floor walls, so the stiffness matters
    steps = 80
    loss, g_v0, g_logE, g_x0 = diff_torch.com_loss_gradients(c.x, c.logE, v0, c.nu, c.particle_mass,
                                                            c.particle_volume, c.params, steps, c.target,
                                                            every=15)
    _, l_ref = oracle_loss(c, c.x, v0, c.logE, steps)
    assert abs(loss - l_ref) < 1e-12 * l_ref
    eps = 1e-3
    logE = c.logE.astype(np.float64)
    x0 = c.x.astype(np.float64)
    L = lambda x, v, le: oracle_loss(c, x, v, le, steps)[1]
    fd_v0 = np.array([(L(x0, v0 + eps * e, logE) - L(x0, v0 - eps * e, logE)) / (2 * eps) for e in np.eye(3)])
    assert np.abs(g_v0 - fd_v0).max() < 1e-5 * np.linalg.norm(fd_v0), (g_v0, fd_v0)
    rng = np.random.default_rng(2)
    for _ in range(2):
        d = rng.standard_normal(logE.shape[0])
        d /= np.linalg.norm(d)
        fd = (L(x0, v0, logE + eps * d) - L(x0, v0, logE - eps * d)) / (2 * eps)
        assert abs(g_logE @ d - fd) < 1e-5 * np.linalg.norm(g_logE), (g_logE @ d, fd)
    d = rng.standard_normal(x0.shape)
    d /= np.linalg.norm(d)
    fd = (L(x0 + eps * d, v0, logE) - L(x0 - eps * d, v0, logE)) / (2 * eps)
    assert abs((g_x0 * d).sum() - fd) < 1e-5 * np.linalg.norm(g_x0), ((g_x0 * d).sum(), fd)
    # the stiffness matters here (the cube lands on the floor)
    assert np.linalg.norm(g_logE) > 1e-4 * np.linalg.norm(g_v0)
