"""mpm_diff_substep_noalloc (the differentiable forward substep with its
scatter grid in the library's static per-device array) against
mpm_diff_substep (the same kernels with a caller-provided grid), through the
C-ABI: two calls in a row, the second on a smaller grid and fewer particles
(the array keeps the first call's sums, which the second must clear); grids
past the array's size are rejected. The scatter adds with FP32 atomics in
hardware order, so the two entry points agree to FP32 summation-order noise:
norm-relative 1e-5 per plane of the state and of the tape."""
import ctypes as C

import numpy as np
import pytest

from mpm_amd import _lib, scenes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def cube(side, N, seed):
    c = scenes.diff_cube(side=side, N=N, seed=seed, lo=(0.4, 0.3, 0.4), spacing=0.2 / side)
    return c


def inputs(c, dev):
    from mpm_amd.diff import lame_from_logE, pack_state
    n = c.x.shape[0]
    rng = np.random.default_rng(7)
    v = torch.from_numpy(rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)).to(dev)
    Cm = torch.from_numpy(rng.uniform(-2, 2, (n, 3, 3)).astype(np.float32)).to(dev)
    F = torch.from_numpy((np.eye(3) + 0.05 * rng.uniform(-1, 1, (n, 3, 3))).astype(np.float32)).to(dev)
    st = pack_state(torch.from_numpy(c.x).to(dev), v, Cm, F)
    mu, lam = lame_from_logE(torch.from_numpy(c.logE).to(dev), c.nu)
    return st, mu.contiguous(), lam.contiguous()


def ptr(t):
    return C.c_void_p(t.data_ptr())


def rel(a, b):
    a = a.double()
    b = b.double()
    return float((a - b).abs().max() / max(float(b.abs().max()), 1e-30))


def run_both(c, dev):
    lib = _lib.lib()
    cp = c.params.to_c()
    n = c.x.shape[0]
    st, mu, lam = inputs(c, dev)
    gf, tf = C.c_int64(), C.c_int64()
    assert lib.mpm_diff_sizes(C.byref(cp), C.byref(gf), C.byref(tf)) == 0
    grid = torch.empty(gf.value, device=dev, dtype=torch.float32)
    out_a = torch.empty_like(st)
    tape_a = torch.empty(tf.value, device=dev, dtype=torch.float32)
    out_b = torch.empty_like(st)
    tape_b = torch.empty(tf.value, device=dev, dtype=torch.float32)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.mpm_diff_substep(C.byref(cp), c.particle_mass, c.particle_volume, n, ptr(st), ptr(out_a),
                              ptr(mu), ptr(lam), ptr(grid), ptr(tape_a), s)
    assert rc == 0, lib.mpm_last_error(None)
    rc = lib.mpm_diff_substep_noalloc(C.byref(cp), c.particle_mass, c.particle_volume, n, ptr(st),
                                      ptr(out_b), ptr(mu), ptr(lam), ptr(tape_b), s)
    assert rc == 0, lib.mpm_last_error(None)
    torch.cuda.synchronize()
    e_state = max(rel(out_b[q], out_a[q]) for q in range(24))
    # the tape: N^3 (velocity, mass) after the update, then N^3 (velocity
    # before the walls, wall flags); as 8 columns per node
    nodes = c.params.N ** 3
    ta = torch.cat([tape_a[:4 * nodes].view(-1, 4), tape_a[4 * nodes:].view(-1, 4)], 1)
    tb = torch.cat([tape_b[:4 * nodes].view(-1, 4), tape_b[4 * nodes:].view(-1, 4)], 1)
    e_tape = max(rel(tb[:, q], ta[:, q]) for q in range(8) if float(ta[:, q].abs().max()) > 0)
    print(f"N={c.params.N} n={n}: state {e_state:.2e} tape {e_tape:.2e}")
    assert e_state < 1e-5 and e_tape < 1e-5, (e_state, e_tape)
    # the flags column is exact, and so is which nodes hold mass
    assert torch.equal(ta[:, 7], tb[:, 7]) and torch.equal(ta[:, 3] > 0, tb[:, 3] > 0)
    assert bool((ta[:, 3] > 0).any())


def test_noalloc_matches_workspace_entry_point(gpu_lib):
    dev = torch.device("cuda")
    run_both(cube(28, 48, 1), dev)   # 21,952 particles on a 48^3 grid
    run_both(cube(12, 24, 2), dev)   # then 1,728 particles on a 24^3 grid


def test_noalloc_rejects_grids_past_its_array(gpu_lib):
    lib = _lib.lib()
    c = cube(4, 65, 3)
    cp = c.params.to_c()
    dev = torch.device("cuda")
    st, mu, lam = inputs(c, dev)
    out = torch.empty_like(st)
    tape = torch.empty(8 * 65 ** 3, device=dev, dtype=torch.float32)
    rc = lib.mpm_diff_substep_noalloc(C.byref(cp), c.particle_mass, c.particle_volume, c.x.shape[0],
                                      ptr(st), ptr(out), ptr(mu), ptr(lam), ptr(tape), None)
    assert rc == _lib.MPM_ERR_CAPACITY
    assert b"N = 65 exceeds" in lib.mpm_last_error(None)
