"""Differentiable rollouts of the MLS-MPM substep on the GPU.

`rollout()` advances a particle state by k substeps of the library's
differentiable substep (mpm_diff_substep: fixed-corotated with J = det F,
per-particle Lame parameters, compat FIXED) inside a torch.autograd.Function
whose backward runs the hand-written adjoint kernels
(mpm_diff_substep_backward). Memory stays bounded by checkpointing: the
forward keeps the state every `checkpoint_every` substeps; the backward
recomputes each segment's states and grid tapes from its checkpoint, then
runs the segment's adjoints in reverse.

The reference has no gradients; the step itself follows src/mpm.cu:14-178
with the MMFixedCorotated plugin (MaterialModel.cuh:56-61).

State layout: one float32 tensor of shape (24, n) on the GPU, planes
x0 x1 x2, v0 v1 v2, C (9, row-major), F (9, row-major). `pack_state` and
`unpack_state` convert from and to per-field tensors. All work goes on
torch's current HIP stream. The forward scatters with FP32 atomics, so the
recomputed segment can differ from the first forward in the last bits.
"""
from __future__ import annotations

import ctypes as C
from typing import Tuple

import torch

from . import _lib
from ._lib import check, lib

PLANES = 24


def pack_state(x, v, C_=None, F=None) -> torch.Tensor:
    """Per-field (n,3), (n,3), (n,3,3), (n,3,3) tensors -> (24, n) state."""
    n = x.shape[0]
    dev = x.device
    if C_ is None:
        C_ = torch.zeros(n, 3, 3, device=dev, dtype=torch.float32)
    if F is None:
        F = torch.eye(3, device=dev, dtype=torch.float32).expand(n, 3, 3)
    return torch.cat([x.reshape(n, 3).T, v.reshape(n, 3).T, C_.reshape(n, 9).T,
                      F.reshape(n, 9).T], 0).to(torch.float32).contiguous()


def unpack_state(st: torch.Tensor):
    """(24, n) state -> x (n,3), v (n,3), C (n,3,3), F (n,3,3)."""
    n = st.shape[1]
    return (st[0:3].T, st[3:6].T, st[6:15].T.reshape(n, 3, 3), st[15:24].T.reshape(n, 3, 3))


def lame_from_logE(logE: torch.Tensor, nu: float) -> Tuple[torch.Tensor, torch.Tensor]:
    """mu, lambda of Young's modulus exp(logE) and Poisson ratio nu
    (the constructor of MaterialModel.cuh:50-54)."""
    E = torch.exp(logE)
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return mu, lam


def _ptr(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


class DiffStepper:
    """The substep and its adjoint for one scene configuration (grid size,
    particle count, particle mass and volume), with its scratch buffers."""

    def __init__(self, params, particle_mass: float, particle_volume: float, n: int,
                 device="cuda"):
        if params.compat != _lib.COMPAT_FIXED:
            raise ValueError("the differentiable substep implements compat FIXED only")
        self.params = params
        self._cp = params.to_c()
        self.mass = float(particle_mass)
        self.volume = float(particle_volume)
        self.n = int(n)
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        gf, tf = C.c_int64(), C.c_int64()
        check(lib().mpm_diff_sizes(C.byref(self._cp), C.byref(gf), C.byref(tf)))
        self.grid_floats, self.tape_floats = gf.value, tf.value
        self.grid = torch.empty(self.grid_floats, device=self.device, dtype=torch.float32)

    def new_tape(self) -> torch.Tensor:
        return torch.empty(self.tape_floats, device=self.device, dtype=torch.float32)

    def _check(self, *ts):
        for t in ts:
            if t.device != self.device or t.dtype != torch.float32 or not t.is_contiguous():
                raise ValueError("expected contiguous float32 tensors on " + str(self.device))

    def forward(self, st, out, mu, lam, tape):
        self._check(st, out, mu, lam, tape)
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().mpm_diff_substep(C.byref(self._cp), self.mass, self.volume, self.n,
                                     _ptr(st), _ptr(out), _ptr(mu), _ptr(lam), _ptr(self.grid),
                                     _ptr(tape), C.c_void_p(s)))

    def backward(self, st, mu, lam, tape, adj_out, adj_in, adj_mu, adj_lam):
        self._check(st, mu, lam, tape, adj_out, adj_in, adj_mu, adj_lam)
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().mpm_diff_substep_backward(
            C.byref(self._cp), self.mass, self.volume, self.n, _ptr(st), _ptr(mu), _ptr(lam),
            _ptr(tape), _ptr(adj_out), _ptr(adj_in), _ptr(adj_mu), _ptr(adj_lam), _ptr(self.grid),
            C.c_void_p(s)))


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, stepper: DiffStepper, steps: int, every: int, st0, mu, lam):
        st0 = st0.detach().contiguous()
        mu = mu.detach().contiguous()
        lam = lam.detach().contiguous()
        tape = stepper.new_tape()
        cur = st0
        checkpoints = [st0]
        for k in range(steps):
            nxt = torch.empty_like(cur)
            stepper.forward(cur, nxt, mu, lam, tape)
            cur = nxt
            if (k + 1) % every == 0 and k + 1 < steps:
                checkpoints.append(cur)
        ctx.stepper, ctx.steps, ctx.every = stepper, steps, every
        ctx.checkpoints = checkpoints
        ctx.save_for_backward(mu, lam)
        return cur

    @staticmethod
    def backward(ctx, g):
        stepper, steps, every = ctx.stepper, ctx.steps, ctx.every
        mu, lam = ctx.saved_tensors
        adj = g.detach().to(torch.float32).contiguous().clone()
        adj_next = torch.empty_like(adj)
        adj_mu = torch.zeros_like(mu)
        adj_lam = torch.zeros_like(lam)
        tapes = [stepper.new_tape() for _ in range(min(every, steps))]
        nseg = len(ctx.checkpoints)
        for seg in range(nseg - 1, -1, -1):
            k0 = seg * every
            k1 = min(steps, k0 + every)
            states = [ctx.checkpoints[seg]]
            for j in range(k1 - k0):  # recompute the segment and its tapes
                nxt = torch.empty_like(states[-1])
                stepper.forward(states[-1], nxt, mu, lam, tapes[j])
                states.append(nxt)
            for j in range(k1 - k0 - 1, -1, -1):
                stepper.backward(states[j], mu, lam, tapes[j], adj, adj_next, adj_mu, adj_lam)
                adj, adj_next = adj_next, adj
        ctx.checkpoints = None
        return None, None, None, adj, adj_mu, adj_lam


def rollout(stepper: DiffStepper, state0: torch.Tensor, mu: torch.Tensor, lam: torch.Tensor,
            steps: int, checkpoint_every: int = 20) -> torch.Tensor:
    """State after `steps` differentiable substeps; differentiable w.r.t.
    state0 (24, n), mu (n,) and lam (n,)."""
    if steps < 1 or checkpoint_every < 1:
        raise ValueError("steps and checkpoint_every must be >= 1")
    if state0.shape != (PLANES, stepper.n) or mu.shape != (stepper.n,) or lam.shape != (stepper.n,):
        raise ValueError("state0 must be (24, n) and mu, lam (n,)")
    return _Rollout.apply(stepper, int(steps), int(checkpoint_every), state0, mu, lam)
