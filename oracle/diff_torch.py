"""FP64 reference gradients of the differentiable substep -- TEST
INFRASTRUCTURE ONLY (the checker of the GPU adjoint; nothing in mpm_amd/
imports it).

The oracle's differentiable step (oracle.advance_lame: the reference's
substep, src/mpm.cu:14-178, with the MMFixedCorotated stress of
MaterialModel.cuh:56-61 taking J = det F, per-particle Lame parameters,
compat FIXED, no plasticity) restated with float64 torch operations on the
CPU, so that torch.autograd gives the gradient of a loss with respect to the
initial state and to every particle's parameter, element by element.

The polar rotation R of F is Higham's Newton iteration R <- (R + R^-T) / 2
from R = F, which converges quadratically to the orthogonal polar factor
for det F > 0 and is differentiable through its unrolled iterations (the
SVD's backward divides by differences of singular values, which vanish at
F = I). Segments of substeps are wrapped in torch.utils.checkpoint to bound
memory.

Pinned by tests/test_oracle_diff.py: the forward against the C++ FP64 oracle
(to FP64 roundoff), the gradients with respect to v0, log E and x0 against
FP64 central differences of the C++ oracle.
"""
from __future__ import annotations

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

F64 = torch.float64


class Consts:
    """The FP64 oracle's step constants (oracle/mpm_oracle.hpp make_consts):
    dx = 1/N, dx_inv = 1/dx, D^-1 = 4 dx_inv^2; dt and gravity are the
    float32 parameters widened; walls decided in float32 as node/N against
    the boundary fraction (or as node layers)."""

    def __init__(self, params):
        self.N = int(params.N)
        self.dt = float(np.float32(params.dt))
        self.dx = 1.0 / float(self.N)
        self.dx_inv = 1.0 / self.dx
        self.dinv = (4.0 * self.dx_inv) * self.dx_inv
        self.g = float(np.float32(params.gravity))
        idx = np.arange(self.N)
        if params.boundary_mode == 1:  # node layers
            b = int(params.boundary)
            lo = idx < b
            hi = idx >= self.N - b
        else:
            b = np.float32(params.boundary)
            f = idx.astype(np.float32) / np.float32(self.N)
            lo = f < b
            hi = f > np.float32(1.0) - b
        X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
        sticky = lo[X] | hi[X] | hi[Y] | lo[Z] | hi[Z]
        self.sticky = torch.from_numpy(sticky.reshape(-1))
        self.floor = torch.from_numpy(lo[Y].reshape(-1))
        o = np.stack(np.meshgrid(*(np.arange(3),) * 3, indexing="ij"), -1).reshape(27, 3)
        self.offs = torch.from_numpy(o)


def polar_rotation(F: torch.Tensor, iters: int = 30) -> torch.Tensor:
    """Orthogonal polar factor of each F (n, 3, 3) by Newton's iteration."""
    R = F
    for _ in range(iters):
        R = 0.5 * (R + torch.linalg.inv(R).transpose(1, 2))
    return R


def substep(x, v, C, F, mu, lam, mass: float, volume: float, k: Consts):
    """One substep of every particle (x, v: (n, 3); C, F: (n, 3, 3))."""
    N = k.N
    xg = x * k.dx_inv
    base = torch.trunc(xg.detach() - 0.5).to(torch.int64)  # the reference's .cast<int>()
    outside = ((base + 3 < 0) | (base >= N)).any(1)
    fx = xg - base.to(F64)
    w = torch.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2], -1)  # (n,3,3)
    node = base[:, None, :] + k.offs[None]  # (n, 27, 3)
    valid = ((node >= 0) & (node < N)).all(-1) & ~outside[:, None]
    wt = w[:, 0, :][:, k.offs[:, 0]] * w[:, 1, :][:, k.offs[:, 1]] * w[:, 2, :][:, k.offs[:, 2]]
    wt = torch.where(valid, wt, torch.zeros_like(wt))
    dp = node.to(F64) * k.dx - x[:, None, :]
    lin = ((node[..., 0] * N + node[..., 1]) * N + node[..., 2]).clamp(0, N ** 3 - 1)
    # P2G (p2g_prepare_particle / p2g_node_contribution, TransferScheme.h:66-100)
    R = polar_rotation(F)
    J = torch.linalg.det(F)
    eye = torch.eye(3, dtype=F64)
    tau = 2.0 * mu[:, None, None] * (F - R) @ F.transpose(1, 2) + (lam * (J - 1.0) * J)[:, None, None] * eye
    A = ((-k.dinv) * k.dt * volume) * tau + mass * C
    mom = wt[..., None] * (mass * v[:, None, :] + torch.einsum("nij,nkj->nki", A, dp))
    P = torch.zeros((N ** 3, 3), dtype=F64).index_add(0, lin.reshape(-1), mom.reshape(-1, 3))
    M = torch.zeros(N ** 3, dtype=F64).index_add(0, lin.reshape(-1), (wt * mass).reshape(-1))
    # grid update (gridOpKernel, mpm.cu:76-107)
    has = M > 0
    Ms = torch.where(has, M, torch.ones_like(M))
    vel = torch.where(has[:, None], P / Ms[:, None], torch.zeros_like(P))
    vel = vel + torch.stack([torch.zeros_like(M), torch.where(has, torch.full_like(M, k.dt * k.g),
                                                               torch.zeros_like(M)), torch.zeros_like(M)], 1)
    vel = torch.where((has & k.sticky)[:, None], torch.zeros_like(vel), vel)
    vy = torch.where(has & k.floor, torch.clamp(vel[:, 1], min=0.0), vel[:, 1])
    vel = torch.stack([vel[:, 0], vy, vel[:, 2]], 1)
    # G2P (g2p_node_contribution / g2p_finish_particle, TransferScheme.h:102-142)
    gv = vel[lin] * valid[..., None]
    vn = (wt[..., None] * gv).sum(1)
    Cn = k.dinv * torch.einsum("nk,nki,nkj->nij", wt, gv, dp)
    Fn = (eye + k.dt * Cn) @ F
    xn = x + k.dt * vn
    keep = outside[:, None]
    return (torch.where(keep, x, xn), torch.where(keep, v, vn),
            torch.where(keep[..., None], C, Cn), torch.where(keep[..., None], F, Fn))


def rollout(x, v, C, F, mu, lam, mass: float, volume: float, params, steps: int, every: int = 20):
    """`steps` substeps, with every `every` substeps recomputed in the
    backward (torch.utils.checkpoint)."""
    k = Consts(params)

    def seg(x, v, C, F, mu, lam, m):
        for _ in range(m):
            x, v, C, F = substep(x, v, C, F, mu, lam, mass, volume, k)
        return x, v, C, F

    done = 0
    while done < steps:
        m = min(every, steps - done)
        x, v, C, F = checkpoint(seg, x, v, C, F, mu, lam, m, use_reentrant=False)
        done += m
    return x, v, C, F


def com_loss_gradients(x0, logE, v0, nu: float, mass: float, volume: float, params, steps: int,
                       target, every: int = 20):
    """The loss |mean(x_T) - target|^2 of a cube with uniform initial velocity
    v0 (3,) and per-particle log E, in float64, and its gradients: returns
    (loss, dL/dv0 (3,), dL/dlogE (n,), dL/dx0 (n, 3))."""
    x = torch.tensor(np.asarray(x0, np.float64), requires_grad=True)
    le = torch.tensor(np.asarray(logE, np.float64), requires_grad=True)
    v0t = torch.tensor(np.asarray(v0, np.float64), requires_grad=True)
    n = x.shape[0]
    E = torch.exp(le)
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    v = v0t.expand(n, 3)
    C = torch.zeros((n, 3, 3), dtype=F64)
    F = torch.eye(3, dtype=F64).expand(n, 3, 3).clone()
    xT, _, _, _ = rollout(x, v, C, F, mu, lam, mass, volume, params, steps, every)
    loss = ((xT.mean(0) - torch.tensor(np.asarray(target, np.float64))) ** 2).sum()
    loss.backward()
    return float(loss.detach()), v0t.grad.numpy().copy(), le.grad.numpy().copy(), x.grad.numpy().copy()
