"""Parity of the HIP engine (through the C-ABI) with the CPU oracle.

Tolerances (norm-relative per field: max|gpu - ref| / max|ref|, with C
measured against max(|C|, |v| / dx), see conftest.state_errors):
  * P2G grid against the FP32 oracle: 2e-5 (FP32 atomics in a different
    order and FMA contraction on the device);
  * elastic scenes after 50 substeps: 1e-4 on x, v, F, C against the FP32
    oracle (the engine's parity bar);
  * liquid after 50 substeps: 1e-4, or -- for a field where the FP32 oracle
    itself is that far from FP64 (the liquid's C) -- an error against FP64
    within twice the FP32 oracle's own;
  * snow after 10 substeps: 1e-3, same FP32-noise allowance (the plasticity
    clamps are discontinuous, so ulp-level differences can flip a clamp);
  * the cell-key sort: bit-exact keys and permutation against numpy's stable
    argsort.
The error against the FP64 oracle is printed for every step-parity case.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import state_elem_errors, state_errors
from mpm_amd import (COMPAT_FIXED, COMPAT_REFERENCE, MATERIAL_FIXED_COROTATED, MATERIAL_JELLY,
                     MATERIAL_SNOW, MATERIAL_WC_LIQUID, Engine, Params, device_svd, scenes)

pytestmark = pytest.mark.gpu

OPTION_SETS = {
    "direct-unfused": dict(p2g_mode=0, fused=0),
    "direct-fused": dict(p2g_mode=0, fused=1),
    "tiled-b4": dict(p2g_mode=1, fused=1, sort_every=1, block=4),
    "tiled-b4-chunk64": dict(p2g_mode=1, fused=1, sort_every=1, block=4, chunk=64),
    "tiled-unfused": dict(p2g_mode=1, fused=0, sort_every=1),
    "tiled-sort8": dict(p2g_mode=1, fused=1, sort_every=8, chunk=128, block=4),
    "tiled-default": dict(p2g_mode=1),
}


def make_engine(scene, state=None, **opts):
    st = scene.state() if state is None else state
    eng = Engine(scene.params)
    eng.set_materials(scene.materials)
    for k, v in opts.items():
        eng.set_option(k, v)
    eng.add_particles(st["x"], st["material"], st["v"], st["C"], st["F"], st["Jp"])
    eng.upload()
    return eng


def run_engine(scene, steps, state=None, **opts):
    eng = make_engine(scene, state, **opts)
    eng.advance(steps)
    out = eng.download()
    eng.close()
    return out


# ---------------------------------------------------------------- linear algebra
def test_device_svd_reference_kats(gpu_lib):
    M1 = np.array([[1, 2, 1], [1, 3, 1], [1, 8, 1]], np.float32)
    M2 = np.array([[0, 1, 0], [-1, 2, -1], [-1, 0.001, -1]], np.float32)
    U, S, V = device_svd(np.stack([M1, M2, np.eye(3, dtype=np.float32)]))
    for i, M in enumerate((M1, M2)):
        u, s, v = U[i].astype(np.float64), S[i].astype(np.float64), V[i].astype(np.float64)
        R, Sym = u @ v.T, v @ np.diag(s) @ v.T
        # TestDevicePolar.Basic (tests/test_linalg.cu:57-75)
        assert np.abs(R @ Sym - M).sum() < 1e-5 * 4
        assert np.abs(R @ R.T - np.eye(3)).sum() < 1e-5 * 4
    assert np.allclose(S[2], 1.0) and np.allclose(U[2] @ V[2].T, np.eye(3), atol=1e-6)


def test_device_svd_matches_oracle(gpu_lib):
    rng = np.random.default_rng(11)
    n = 20000
    Q1 = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    Q2 = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    s = np.exp(rng.uniform(-1, 1, size=(n, 3)))
    A = (Q1 @ (s[:, :, None] * np.swapaxes(Q2, 1, 2))).astype(np.float32)
    U, S, V = device_svd(A)
    Uo, So, Vo = oracle.svd(A, np.float32)
    R = U.astype(np.float64) @ np.swapaxes(V, 1, 2)
    Ro = Uo.astype(np.float64) @ np.swapaxes(Vo, 1, 2)
    err = np.abs(R - Ro).max(axis=(1, 2))
    print(f"device vs oracle polar R: median {np.median(err):.2e} p99.9 {np.quantile(err, 0.999):.2e} max {err.max():.2e}")
    assert np.median(err) < 1e-6
    assert np.quantile(err, 0.999) < 1e-4
    sig = np.abs(S - So).max(axis=1) / So[:, 0]
    assert np.median(sig) < 1e-6
    rec = U.astype(np.float64) @ (S[:, :, None] * np.swapaxes(V, 1, 2))
    assert np.median(np.abs(rec - A).max(axis=(1, 2))) < 1e-5


def test_device_svd_fallback_near_singular(gpu_lib):
    # near-singular F goes through the FP64 fallback: U, V rotations,
    # accurate small singular value and sign(det)
    rng = np.random.default_rng(12)
    n = 2000
    Q1 = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    Q2 = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    Q1[:, :, 0] *= np.linalg.det(Q1)[:, None]
    Q2[:, :, 0] *= np.linalg.det(Q2)[:, None]
    s = np.stack([np.ones(n), rng.uniform(0.1, 1, n), rng.uniform(-1e-5, 1e-5, n)], 1)
    A = (Q1 @ (s[:, :, None] * np.swapaxes(Q2, 1, 2))).astype(np.float32)
    U, S, V = device_svd(A)
    rec = U.astype(np.float64) @ (S[:, :, None] * np.swapaxes(V, 1, 2))
    assert np.abs(rec - A).max() < 1e-5
    assert np.allclose(np.linalg.det(U.astype(np.float64)), 1, atol=1e-5)
    assert np.allclose(np.linalg.det(V.astype(np.float64)), 1, atol=1e-5)
    assert np.all(S[:, 0] >= 0) and np.all(S[:, 1] >= 0)


# ---------------------------------------------------------------- sort
def spread3(v):
    v = v.astype(np.uint64) & 0x3FF
    v = (v | (v << 16)) & 0x030000FF
    v = (v | (v << 8)) & 0x0300F00F
    v = (v | (v << 4)) & 0x030C30C3
    v = (v | (v << 2)) & 0x09249249
    return v


def cell_keys_numpy(x, N, block):
    """Blocked-Morton cell keys: blocks of block^3 cells in Morton order,
    cells row-major inside a block (the engine's sort key)."""
    lb = int(block).bit_length() - 1
    dx = np.float32(1.0 / N)
    dx_inv = np.float32(1.0 / np.float64(dx))
    xg = x.astype(np.float32) * dx_inv
    b = np.trunc(xg - np.float32(0.5)).astype(np.int64)
    b = np.clip(b, 0, N - 1)
    m = (spread3(b[:, 0] >> lb) << 2) | (spread3(b[:, 1] >> lb) << 1) | spread3(b[:, 2] >> lb)
    mask = block - 1
    local = (((b[:, 0] & mask) << lb | (b[:, 1] & mask)) << lb) | (b[:, 2] & mask)
    return ((m << (3 * lb)) | local.astype(np.uint64)).astype(np.uint32)


@pytest.mark.parametrize("block", [4])
@pytest.mark.parametrize("radix_bits", [4, 6, 8])
@pytest.mark.parametrize("which", ["snow_small", "dam_break_1M", "outside"])
def test_cell_sort_bit_exact(gpu_lib, radix_bits, which, block):
    if which == "snow_small":
        sc = scenes.snow_spheres(n=30_000, N=64)
    elif which == "dam_break_1M":
        sc = scenes.dam_break()
    else:
        sc = scenes.small_scene(MATERIAL_SNOW, 5000, 21, N=16)
        sc.x[::7] += 0.9   # some particles outside the grid (clamped keys)
        sc.x[::11] -= 0.9
    eng = make_engine(sc, radix_bits=radix_bits, block=block)
    keys, perm = eng.debug_sort()
    eng.close()
    ref_keys = cell_keys_numpy(sc.x, sc.params.N, block)
    ref_perm = np.argsort(ref_keys, kind="stable")
    assert np.array_equal(keys, ref_keys[ref_perm])
    assert np.array_equal(perm, ref_perm.astype(np.uint32))


# ---------------------------------------------------------------- P2G
def grid_rel_err(a, b):
    return max(float(np.abs(a[..., c] - b[..., c]).max() / max(np.abs(b[..., c]).max(), 1e-30))
               for c in range(4))


P2G_SCENES = {
    "rubber": lambda compat: scenes.rubber_elastic(3000, 31, N=16, compat=compat,
                                                   velocity=(0.5, -1.0, 0.2)),
    "snow": lambda compat: scenes.small_scene(MATERIAL_SNOW, 20000, 32, N=32, compat=compat,
                                              velocity=(0.0, -2.0, 0.0)),
    "liquid": lambda compat: scenes.small_scene(MATERIAL_WC_LIQUID, 20000, 33, N=32, compat=compat),
    "edges": lambda compat: scenes.small_scene(MATERIAL_SNOW, 4000, 34, N=16, compat=compat,
                                               center=(0.85, 0.9, 0.88), radius=0.12),
    # a few cells holding hundreds of particles each: items fill every lane of
    # a wavefront with one cell and the per-wavefront counts overflow 64
    "cluster": lambda compat: scenes.small_scene(MATERIAL_SNOW, 3000, 35, N=32, compat=compat,
                                                 center=(0.5, 0.5, 0.5), radius=0.02),
}


def perturbed_state(sc, seed):
    rng = np.random.default_rng(seed)
    st = sc.state()
    n = sc.n
    st["C"] = rng.uniform(-3, 3, (n, 3, 3)).astype(np.float32)
    st["F"] = (np.eye(3) + 0.05 * rng.uniform(-1, 1, (n, 3, 3))).astype(np.float32)
    st["Jp"] = rng.uniform(0.9, 1.1, n).astype(np.float32)
    return st


@pytest.mark.parametrize("opts", ["direct-fused", "tiled-b4", "tiled-b4-chunk64", "tiled-unfused"])
@pytest.mark.parametrize("compat", [COMPAT_REFERENCE, COMPAT_FIXED], ids=["ref", "fixed"])
@pytest.mark.parametrize("name", list(P2G_SCENES))
def test_p2g_matches_oracle(gpu_lib, name, compat, opts):
    sc = P2G_SCENES[name](compat)
    st = perturbed_state(sc, 40)
    o = dict(OPTION_SETS[opts])
    o.pop("sort_every", None)
    eng = make_engine(sc, st, **o)
    g = eng.debug_p2g()
    eng.close()
    ref = oracle.p2g(sc.params, sc.materials, st, np.float32)
    err = grid_rel_err(g, ref)
    print(f"{name} compat={compat} {opts}: P2G grid rel err {err:.2e}")
    assert err < 2e-5


# ---------------------------------------------------------------- full substeps
def compare_with_oracle(sc, st, out, steps, tol, fp32_noise=False, threads=1, label=""):
    """`out` (the engine's state after `steps` substeps from `st`) against
    the FP32 and FP64 oracle. Strict mode: every field within `tol` of the
    FP32 oracle (norm-relative). fp32_noise mode (liquid, snow): a field may
    instead pass if its error against FP64 is within twice the FP32 oracle's
    own error against FP64 -- the engine is then as accurate as the
    reference's FP32 arithmetic allows. The element-wise relative errors are
    printed beside the norm-relative ones."""
    ref32 = oracle.advance(sc.params, sc.materials, st, steps, np.float32, threads=threads)
    ref64 = oracle.advance(sc.params, sc.materials, st, steps, np.float64, threads=threads)
    N = sc.params.N
    e32 = state_errors(out, ref32, N)
    e64 = state_errors(out, ref64, N)
    o64 = state_errors(ref32, ref64, N)
    el32 = state_elem_errors(out, ref32, N)
    el64 = state_elem_errors(out, ref64, N)
    print(f"{label or sc.name} k={steps}: vs FP32 oracle {fmt(e32)} | vs FP64 {fmt(e64)} | "
          f"FP32 oracle vs FP64 {fmt(o64)} | element-wise vs FP32 {fmt(el32)} vs FP64 {fmt(el64)}")
    for k in ("x", "v", "F", "C"):
        ok = e32[k] < tol
        if not ok and fp32_noise:
            ok = e64[k] <= 2.0 * o64[k] and e32[k] < 10 * tol
        assert ok, (k, e32[k], e64[k], o64[k])
    assert np.all(np.isfinite(out["x"]))
    return e32


def step_case(sc, steps, opts, tol, state=None, fp32_noise=False, threads=1):
    """Run `steps` substeps on the engine (option set `opts`) and check them
    against the oracle (compare_with_oracle)."""
    st = sc.state() if state is None else state
    out = run_engine(sc, steps, st, **OPTION_SETS[opts])
    compare_with_oracle(sc, st, out, steps, tol, fp32_noise, threads, f"{sc.name} {opts}")
    return out


def fmt(e):
    return " ".join(f"{k}={v:.1e}" for k, v in e.items())


@pytest.mark.parametrize("opts", list(OPTION_SETS))
@pytest.mark.parametrize("compat", [COMPAT_REFERENCE, COMPAT_FIXED], ids=["ref", "fixed"])
def test_elastic_sphere_50_substeps(gpu_lib, compat, opts):
    # the engine's elastic workload (10,000 particles, N=16, walls 3 layers)
    sc = scenes.elastic_sphere(compat=compat)
    step_case(sc, 50, opts, 1e-4)


@pytest.mark.parametrize("opts", ["direct-fused", "tiled-b4", "tiled-b4-chunk64", "tiled-unfused", "tiled-default"])
def test_rubber_ball_moving_50_substeps(gpu_lib, opts):
    sc = scenes.rubber_elastic(8000, 41, N=32, velocity=(1.0, -2.0, 0.5))
    step_case(sc, 50, opts, 1e-4)


@pytest.mark.parametrize("opts", ["direct-fused", "tiled-b4", "tiled-b4-chunk64", "tiled-unfused", "tiled-default"])
def test_liquid_50_substeps(gpu_lib, opts):
    sc = scenes.dam_break(n=40_000, N=32)
    step_case(sc, 50, opts, 1e-4, fp32_noise=True)


def test_liquid_derived_f_is_bit_identical(gpu_lib):
    """The liquid step that keeps F = cbrt(J) I implicit gives the same bits as
    the one that reads and writes the F planes, across a download in the
    middle (which writes the planes out) and a switch to the unfused path."""
    sc = scenes.dam_break(n=40_000, N=32)
    outs = []
    for derive in (1, 0):
        eng = make_engine(sc, derive_f=derive)
        eng.advance(37)
        mid = eng.download()
        eng.advance(20)
        eng.set_option("fused", 0)
        eng.advance(5)
        eng.set_option("fused", 1)
        eng.advance(9)
        outs.append((mid, eng.download()))
        eng.close()
    for a, b in zip(outs[0], outs[1]):
        for f in ("x", "v", "C", "F", "Jp"):
            assert np.array_equal(a[f], b[f]), f
    mid = outs[0][0]
    cj = np.cbrt(mid["Jp"].astype(np.float64))
    assert np.allclose(mid["F"], cj[:, None, None] * np.eye(3), rtol=3e-7, atol=0)


def test_reduce_strategies_are_bit_identical(gpu_lib):
    """The ways the grid update reads the further work items of crowded home
    blocks -- one at a time, 8 blocks' items per batch, or after a combine
    pass that sums each crowded block's tiles into its first -- all sum in
    item order: a dense scene (several items per block) gives the same bits
    with each, and with the choices left to the engine."""
    sc = scenes.batched_scenes()[1]  # 41k particles in a N=32 grid, ~800 per occupied block
    outs = []
    for combine, mlp in ((0, 0), (0, 1), (1, 0), (-1, -1)):
        eng = make_engine(sc, reduce_mlp=mlp, combine=combine)
        assert eng.get_option("reduce_mlp") == mlp and eng.get_option("combine") == combine
        eng.advance(40)
        outs.append(eng.download())
        eng.close()
    for o in outs[1:]:
        for f in ("x", "v", "C", "F", "Jp"):
            assert np.array_equal(outs[0][f], o[f]), f


def test_markers_change_nothing(gpu_lib):
    """roctx ranges (option "markers") are host-side only: the same bits with
    and without them, timing events included."""
    sc = scenes.snow_spheres(n=20_000, N=32)
    outs = []
    for markers in (1, 0):
        eng = make_engine(sc, markers=markers, timing=1)
        assert eng.get_option("markers") == markers
        eng.advance(20)
        outs.append(eng.download())
        eng.close()
    for f in ("x", "v", "C", "F", "Jp"):
        assert np.array_equal(outs[0][f], outs[1][f]), f


@pytest.mark.parametrize("opts", ["direct-unfused", "tiled-b4", "tiled-b4-chunk64", "tiled-unfused", "tiled-default"])
def test_snow_10_substeps(gpu_lib, opts):
    sc = scenes.snow_spheres(n=40_000, N=32)
    step_case(sc, 10, opts, 1e-3, fp32_noise=True)


@pytest.mark.parametrize("kind", [MATERIAL_FIXED_COROTATED, MATERIAL_JELLY])
def test_other_material_plugins(gpu_lib, kind):
    sc = scenes.small_scene(kind, 6000, 51, N=24, velocity=(0.0, -1.0, 0.0))
    step_case(sc, 30, "tiled-b4", 1e-4)


def test_mixed_materials(gpu_lib):
    a = scenes.small_scene(MATERIAL_SNOW, 5000, 61, N=32, center=(0.35, 0.5, 0.5), radius=0.12)
    b = scenes.small_scene(MATERIAL_WC_LIQUID, 5000, 62, N=32, center=(0.65, 0.5, 0.5), radius=0.12)
    c = scenes.rubber_elastic(5000, 63, N=32)
    sc = a
    sc.x = np.concatenate([a.x, b.x, c.x * 0.5 + 0.25])
    sc.material = np.concatenate([np.zeros(5000), np.ones(5000), 2 * np.ones(5000)]).astype(np.uint8)
    sc.materials = [a.materials[0], b.materials[0], c.materials[0]]
    sc.v = None
    step_case(sc, 10, "tiled-b4", 1e-3)


# ---------------------------------------------------------------- boundary behaviour
def test_empty_scene(gpu_lib):
    sc = scenes.small_scene(MATERIAL_SNOW, 10, 1, N=8)
    eng = Engine(sc.params)
    eng.set_materials(sc.materials)
    eng.upload()
    eng.advance(3)
    out = eng.download()
    assert out["x"].shape == (0, 3)
    assert abs(eng.time() - 3 * np.float64(np.float32(1e-4))) < 1e-12
    eng.close()


def test_single_particle_free_fall(gpu_lib):
    sc = scenes.small_scene(MATERIAL_SNOW, 1, 2, N=16)
    sc.params.boundary = 0.0
    step_case(sc, 5, "tiled-b4", 1e-5)


def test_out_of_domain_particles_stay_frozen(gpu_lib):
    sc = scenes.small_scene(MATERIAL_SNOW, 2000, 3, N=16)
    st = sc.state()
    st["x"][:50] += 1.5
    out = run_engine(sc, 4, st, **OPTION_SETS["tiled-b4"])
    assert np.array_equal(out["x"][:50], st["x"][:50])
    assert np.array_equal(out["v"][:50], st["v"][:50])


def test_aos104_roundtrip(gpu_lib):
    sc = scenes.rubber_elastic(777, 4, N=16)
    st = perturbed_state(sc, 5)
    eng = make_engine(sc, st)
    raw = eng.download_aos104()
    eng2 = Engine(sc.params)
    eng2.set_materials(sc.materials)
    eng2.add_particles_aos104(raw)
    eng2.upload()
    out = eng2.download()
    for k in ("x", "v", "C", "F", "Jp"):
        assert np.array_equal(out[k], st[k].astype(np.float32)), k
    # column-major F at byte 28, as the reference's Eigen matrices
    rec = raw.reshape(-1, 104)
    F0 = np.frombuffer(rec[0, 28:64].tobytes(), np.float32).reshape(3, 3).T
    assert np.array_equal(F0, st["F"][0])
    eng.close()
    eng2.close()


def test_save_load_state_roundtrip(gpu_lib, tmp_path):
    sc = scenes.snow_spheres(n=20_000, N=32)
    eng = make_engine(sc, **OPTION_SETS["tiled-b4"])
    eng.advance(5)
    path = tmp_path / "state.bin"
    eng.save_state(path)
    a = eng.download()
    eng2 = Engine(Params(dt=1e-3, N=8))
    eng2.load_state(path)
    b = eng2.download()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert eng2.time() == eng.time()
    eng.close()
    eng2.close()


def test_grid_download_after_update(gpu_lib):
    # free fall of a uniform block: every node with mass moves at dt g
    sc = scenes.small_scene(MATERIAL_SNOW, 5000, 7, N=16)
    sc.params.boundary = 0.0
    eng = make_engine(sc, **OPTION_SETS["tiled-b4"])
    eng.advance(1)
    g = eng.download_grid()
    eng.close()
    has = g[..., 3] > 0
    assert np.allclose(g[has][:, 3], 1.0)
    dtg = np.float32(1e-4) * np.float32(-9.8)
    assert np.abs(g[has][:, 1] - dtg).max() < 1e-6


# ---------------------------------------------------------------- full size
def test_dam_break_full_size_properties(gpu_lib):
    """1,000,000 particles at N=128: sort exactness is covered above; here the
    size-independent properties: P2G mass equals the particle mass sum, the
    sorted tiled path agrees with the direct one after 10 substeps, and the
    momentum change of the first substep is the free-fall impulse."""
    sc = scenes.dam_break()
    m = sc.materials[0].particle_mass
    eng = make_engine(sc, **OPTION_SETS["tiled-b4"])
    g = eng.debug_p2g()
    assert abs(g[..., 3].astype(np.float64).sum() - sc.n * m) / (sc.n * m) < 1e-5
    eng.advance(10)
    a = eng.download()
    eng.close()
    b = run_engine(sc, 10, **OPTION_SETS["direct-fused"])
    e = state_errors(a, b, sc.params.N)
    print("dam break 1M, tiled vs direct after 10 substeps:", fmt(e))
    for k in ("x", "v", "F"):
        assert e[k] < 1e-4, (k, e[k])
    assert np.all(np.isfinite(a["x"])) and np.all(np.isfinite(a["F"]))


@pytest.mark.parametrize("opts", ["tiled-default", "direct-fused"])
def test_snow_spheres_full_size_10_substeps(gpu_lib, opts):
    """The snow workload at its full size: two stacked balls, 500,000
    particles, N=64 (compat REFERENCE: the original-index guard freezes the
    237,855 particles past N^3 = 262,144 in G2P), 10 substeps against the FP32
    and FP64 oracle, with the snow rule (1e-3, or the FP32 oracle's own
    distance from FP64)."""
    sc = scenes.snow_spheres()
    assert sc.n == 500_000 and sc.params.N == 64
    step_case(sc, 10, opts, 1e-3, fp32_noise=True, threads=oracle.check_threads())


# ---------------------------------------------------------------- interpolation plugin
@pytest.mark.parametrize("compat", [COMPAT_REFERENCE, COMPAT_FIXED], ids=["ref", "fixed"])
@pytest.mark.parametrize("name", ["rubber", "snow", "liquid", "edges"])
def test_p2g_cubic_kernel_matches_oracle(gpu_lib, name, compat):
    """The second InterpolationKernel policy (cubic B-spline, generic D^-1,
    InterpolationKernel.cuh:21-50) on the direct path: the raw P2G grid
    against the oracle's cubic kernel, 2e-5 (FP32 atomics in a different
    order); the quadratic grid of the same state differs, so the option is
    not ignored."""
    sc = P2G_SCENES[name](compat)
    st = perturbed_state(sc, 41)
    eng = make_engine(sc, st, p2g_mode=0, interpolation=1)
    assert eng.get_option("interpolation") == 1
    g = eng.debug_p2g()
    eng.close()
    ref = oracle.p2g(sc.params, sc.materials, st, np.float32, interp=oracle.INTERP_CUBIC)
    quad = oracle.p2g(sc.params, sc.materials, st, np.float32)
    err = grid_rel_err(g, ref)
    print(f"{name} compat={compat} cubic: P2G grid rel err {err:.2e}")
    assert err < 2e-5
    assert grid_rel_err(g, quad) > 1e-3


@pytest.mark.parametrize("fused", [0, 1], ids=["unfused", "fused"])
def test_cubic_kernel_substeps_match_oracle(gpu_lib, fused):
    """30 substeps of a moving elastic ball with the cubic kernel (direct
    path; p2g_mode 1 is left set and the engine runs the cubic kernel direct)
    against the FP32 and FP64 oracle with the same kernel: 1e-4."""
    sc = scenes.rubber_elastic(6000, 42, N=24, velocity=(0.8, -1.5, 0.3))
    st = sc.state()
    eng = make_engine(sc, st, fused=fused, interpolation=1)
    assert eng.get_option("p2g_mode") == 1
    eng.advance(30)
    out = eng.download()
    eng.close()
    N = sc.params.N
    ref32 = oracle.advance(sc.params, sc.materials, st, 30, np.float32, interp=oracle.INTERP_CUBIC)
    ref64 = oracle.advance(sc.params, sc.materials, st, 30, np.float64, interp=oracle.INTERP_CUBIC)
    e32 = state_errors(out, ref32, N)
    e64 = state_errors(out, ref64, N)
    print(f"cubic ball fused={fused} k=30: vs FP32 oracle {fmt(e32)} | vs FP64 {fmt(e64)} | "
          f"element-wise vs FP32 {fmt(state_elem_errors(out, ref32, N))}")
    for k in ("x", "v", "F", "C"):
        assert e32[k] < 1e-4, (k, e32[k])


def test_cubic_kernel_is_rejected_by_the_tiled_tuner(gpu_lib):
    sc = scenes.rubber_elastic(2000, 43, N=16)
    eng = make_engine(sc, interpolation=1)
    with pytest.raises(Exception, match="quadratic kernel only"):
        eng.autotune(trial_substeps=16)
    eng.set_option("interpolation", 0)
    eng.advance(2)
    eng.close()
