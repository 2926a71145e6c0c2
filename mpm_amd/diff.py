"""Differentiable rollouts of the MLS-MPM substep on the GPU.

`rollout()` advances a particle state by k substeps of the library's
differentiable substep (fixed-corotated with J = det F, per-particle Lame
parameters, compat FIXED) inside a torch.autograd.Function whose backward
runs the hand-written adjoint kernels. Memory stays bounded by
checkpointing: the forward keeps the state every `checkpoint_every`
substeps; the backward recomputes each segment's states and grid tapes from
its checkpoint (mpm_diff_forward_steps), then runs the segment's adjoints in
reverse (mpm_diff_backward_steps). Each segment is one call into the
library, which enqueues all of its kernels (two per substep each way). The
recompute of segment s - 1 runs on a second stream while segment s's
adjoints run (two buffers of states and tapes): a cube of a few ten
thousand particles fills half the SIMDs, and each pass is latency-bound.

Inside the Function the particles are ordered by the grid cell of their
initial position (a stable sort, undone on the way out, gradients included):
the kernels sum the contributions of particles that share a cell before
their global atomics, so cell order makes those sums long.

The reference has no gradients; the step itself follows src/mpm.cu:14-178
with the MMFixedCorotated plugin (MaterialModel.cuh:56-61).

State layout: one float32 tensor of shape (24, n) on the GPU, planes
x0 x1 x2, v0 v1 v2, C (9, row-major), F (9, row-major). `pack_state` and
`unpack_state` convert from and to per-field tensors. All work goes on
torch's current HIP stream (the backward's recomputes on a side stream of
default priority, joined before it returns). The scatters add with FP32
atomics, so a recomputed segment can differ from the first forward in the
last bits.
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
        self._grid2 = None
        self._work = None
        # the backward's recompute stream (default priority), made here rather
        # than inside a backward that a caller may be recording into a graph
        self._side = torch.cuda.Stream(device=self.device)

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def new_tape(self) -> torch.Tensor:
        return torch.empty(self.tape_floats, device=self.device, dtype=torch.float32)

    def new_tapes(self, k: int) -> torch.Tensor:
        return torch.empty((k, self.tape_floats), device=self.device, dtype=torch.float32)

    def _check(self, *ts):
        for t in ts:
            if t.device != self.device or t.dtype != torch.float32 or not t.is_contiguous():
                raise ValueError("expected contiguous float32 tensors on " + str(self.device))

    # one substep (mpm_diff_substep / mpm_diff_substep_backward)
    def forward(self, st, out, mu, lam, tape):
        self._check(st, out, mu, lam, tape)
        check(lib().mpm_diff_substep(C.byref(self._cp), self.mass, self.volume, self.n,
                                     _ptr(st), _ptr(out), _ptr(mu), _ptr(lam), _ptr(self.grid),
                                     _ptr(tape), self._stream()))

    def backward(self, st, mu, lam, tape, adj_out, adj_in, adj_mu, adj_lam):
        self._check(st, mu, lam, tape, adj_out, adj_in, adj_mu, adj_lam)
        check(lib().mpm_diff_substep_backward(
            C.byref(self._cp), self.mass, self.volume, self.n, _ptr(st), _ptr(mu), _ptr(lam),
            _ptr(tape), _ptr(adj_out), _ptr(adj_in), _ptr(adj_mu), _ptr(adj_lam), _ptr(self.grid),
            self._stream()))

    # a run of substeps in one call (mpm_diff_forward_steps / _backward_steps)
    def forward_steps(self, st0, states, mu, lam, tapes=None):
        """len(states) substeps from st0: states[j] (a (k, 24, n) tensor) is
        the state after substep j + 1; with `tapes` ((k, tape_floats)) the
        grid tape of each substep is kept (else one scratch tape is reused)."""
        k = states.shape[0]
        self._check(st0, states, mu, lam)
        if tapes is None:
            tapes, stride = self.new_tape(), 0
        else:
            self._check(tapes)
            if tapes.shape[0] < k:
                raise ValueError("one tape per substep")
            stride = tapes.shape[1]
        check(lib().mpm_diff_forward_steps(
            C.byref(self._cp), self.mass, self.volume, self.n, int(k), _ptr(st0), _ptr(states),
            PLANES * self.n, _ptr(mu), _ptr(lam), _ptr(self.grid), _ptr(tapes), int(stride),
            self._stream()))

    def backward_steps(self, states, mu, lam, tapes, adj_out, adj_in, adj_mu, adj_lam):
        """The adjoint of len(states) substeps: states[j] is the state before
        substep j, tapes[j] its tape; adj_out the adjoint of the state after
        the last one. Writes adj_in, adds into adj_mu and adj_lam."""
        k = states.shape[0]
        self._check(states, mu, lam, tapes, adj_out, adj_in, adj_mu, adj_lam)
        if self._grid2 is None:
            self._grid2 = torch.empty(2 * self.grid_floats, device=self.device, dtype=torch.float32)
            self._work = torch.empty((PLANES, self.n), device=self.device, dtype=torch.float32)
        check(lib().mpm_diff_backward_steps(
            C.byref(self._cp), self.mass, self.volume, self.n, int(k), _ptr(states), PLANES * self.n,
            _ptr(mu), _ptr(lam), _ptr(tapes), int(tapes.shape[1]), _ptr(adj_out), _ptr(adj_in),
            _ptr(adj_mu), _ptr(adj_lam), _ptr(self._grid2), _ptr(self._work), self._stream()))


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, stepper: DiffStepper, steps: int, every: int, st0, mu, lam):
        st0 = st0.detach().contiguous()
        mu = mu.detach().contiguous()
        lam = lam.detach().contiguous()
        seg = torch.empty((min(every, steps), PLANES, stepper.n), device=st0.device, dtype=torch.float32)
        cur = st0
        checkpoints = [st0]
        done = 0
        while done < steps:
            k = min(every, steps - done)
            # the segment's last state goes straight into a tensor of its own
            # (the next checkpoint, or the output): no copy out of `seg`
            nxt = torch.empty((PLANES, stepper.n), device=st0.device, dtype=torch.float32)
            if k > 1:
                stepper.forward_steps(cur, seg[:k - 1], mu, lam)
                stepper.forward_steps(seg[k - 2], nxt.unsqueeze(0), mu, lam)
            else:
                stepper.forward_steps(cur, nxt.unsqueeze(0), mu, lam)
            cur = nxt
            done += k
            if done < steps:
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
        k_max = min(every, steps)
        nseg = len(ctx.checkpoints)
        # two buffers: segment s's adjoints read one while segment s - 1 is
        # recomputed into the other on the side stream
        nbuf = 2 if nseg > 1 else 1
        states = [torch.empty((k_max + 1, PLANES, stepper.n), device=adj.device, dtype=torch.float32)
                  for _ in range(nbuf)]
        tapes = [stepper.new_tapes(k_max) for _ in range(nbuf)]
        main = torch.cuda.current_stream(adj.device)
        side = stepper._side
        side.wait_stream(main)  # the forward's use of the scatter grid, the inputs
        ready = [torch.cuda.Event() for _ in range(nbuf)]  # recompute done (side)
        free = [torch.cuda.Event() for _ in range(nbuf)]   # adjoints done reading (main)
        freed = [False] * nbuf

        def seg_len(s):
            k0 = s * every
            return min(steps, k0 + every) - k0

        def recompute(s, b):
            # states[b][j] is the state before substep j of segment s
            k = seg_len(s)
            with torch.cuda.stream(side):
                if freed[b]:
                    side.wait_event(free[b])
                states[b][0].copy_(ctx.checkpoints[s])
                stepper.forward_steps(states[b][0], states[b][1:k + 1], mu, lam, tapes[b])
                ready[b].record(side)

        recompute(nseg - 1, 0)
        for s in range(nseg - 1, -1, -1):
            b = (nseg - 1 - s) % nbuf
            if s > 0:
                recompute(s - 1, (b + 1) % nbuf)
            main.wait_event(ready[b])
            k = seg_len(s)
            stepper.backward_steps(states[b][:k], mu, lam, tapes[b], adj, adj_next, adj_mu, adj_lam)
            free[b].record(main)
            freed[b] = True
            adj, adj_next = adj_next, adj
        main.wait_stream(side)
        for t in states + tapes:
            t.record_stream(side)
        ctx.checkpoints = None
        return None, None, None, adj, adj_mu, adj_lam


def cell_order(x: torch.Tensor, N: int) -> torch.Tensor:
    """Stable order of the particles (positions x: (3, n)) by the base cell
    of their quadratic stencil, int(x / dx - 0.5) per axis."""
    dx_inv = float(torch.tensor(1.0 / float(torch.tensor(1.0 / N, dtype=torch.float32)),
                                dtype=torch.float32))
    b = torch.trunc(x * dx_inv - 0.5).to(torch.int64) + 3
    M = N + 3
    key = (b[0] * M + b[1]) * M + b[2]
    return torch.argsort(key, stable=True)


def rollout(stepper: DiffStepper, state0: torch.Tensor, mu: torch.Tensor, lam: torch.Tensor,
            steps: int, checkpoint_every: int = 20) -> torch.Tensor:
    """State after `steps` differentiable substeps; differentiable w.r.t.
    state0 (24, n), mu (n,) and lam (n,)."""
    if steps < 1 or checkpoint_every < 1:
        raise ValueError("steps and checkpoint_every must be >= 1")
    if state0.shape != (PLANES, stepper.n) or mu.shape != (stepper.n,) or lam.shape != (stepper.n,):
        raise ValueError("state0 must be (24, n) and mu, lam (n,)")
    with torch.no_grad():
        perm = cell_order(state0[0:3].detach(), stepper.params.N)
        inv = torch.empty_like(perm)
        inv[perm] = torch.arange(perm.numel(), device=perm.device)
    out = _Rollout.apply(stepper, int(steps), int(checkpoint_every), state0[:, perm], mu[perm], lam[perm])
    return out[:, inv]
