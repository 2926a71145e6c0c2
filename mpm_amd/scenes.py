"""Seeded synthetic scenes for the engine's workloads and tests.

Every random number comes from one counter-based generator, SplitMix64
(Steele, Lea, Flood 2014): value i of stream `seed` is mix(seed + (i+1) *
0x9E3779B97F4A7C15). Uniform floats take the top 24 bits, so they are exact
in float32. The engine and the CPU oracle consume the same arrays.

Named workloads (sizes of the engine's benchmark configurations):
  elastic_sphere  10,000 particles in a ball r=0.15 at (0.5,0.6,0.5); N=16,
                  dt=1e-4, E=1e4, nu=0.3, rho=1000, walls 3 node layers.
  snow_spheres    balls r=0.2 at (0.5,0.3,0.5) and r=0.13 at (0.5,0.62,0.5),
                  500,000 particles split by volume; N=64, dt=1e-4, snow
                  E=1.4e5, nu=0.2, rho=400, clamps [1-2.5e-2, 1+7.5e-3], xi=10.
  dam_break       box [0.05,0.45]x[0.05,0.65]x[0.05,0.95], 1,000,000
                  particles; N=128, dt=5e-5, weakly compressible liquid
                  K=4e4, rho=1000.
  elastic_cube    32^3 lattice in [0.35,0.65]^3 with +-1/4 spacing jitter;
                  N=32, dt=1e-4, E=1e4, nu=0.3, rho=1000.
Particle volume is the object volume over its particle count (the
reference's `1/particle_count` semantics, main.cu:36).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import _lib
from .engine import Params, material_as_numpy

GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, start: int, count: int) -> np.ndarray:
    """Values start..start+count-1 of the SplitMix64 stream `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, start: int, count: int) -> np.ndarray:
    """float32 uniforms in [0, 1) with 24 random bits."""
    return ((splitmix64(seed, start, count) >> np.uint64(40)).astype(np.float64) *
            (1.0 / (1 << 24))).astype(np.float32)


class Sampler:
    def __init__(self, seed: int):
        self.seed = int(seed)
        self.pos = 0

    def uniform(self, count: int) -> np.ndarray:
        u = uniform01(self.seed, self.pos, count)
        self.pos += count
        return u


def sample_ball(s: Sampler, n: int, center, radius) -> np.ndarray:
    """n points uniform in a ball, by rejection from the bounding cube."""
    out = []
    have = 0
    c = np.asarray(center, np.float64)
    while have < n:
        m = max(1024, int((n - have) * 2.2))
        u = s.uniform(3 * m).astype(np.float64).reshape(m, 3) * 2.0 - 1.0
        u = u[(u * u).sum(1) < 1.0]
        out.append(u)
        have += u.shape[0]
    p = np.concatenate(out)[:n] * radius + c
    return p.astype(np.float32)


def sample_box(s: Sampler, n: int, lo, hi) -> np.ndarray:
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    u = s.uniform(3 * n).astype(np.float64).reshape(n, 3)
    return (lo + u * (hi - lo)).astype(np.float32)


@dataclass
class Scene:
    name: str
    params: Params
    materials: list
    x: np.ndarray
    material: np.ndarray
    v: np.ndarray = None
    substeps: int = 1
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def state(self) -> dict:
        n = self.n
        F = np.zeros((n, 3, 3), np.float32)
        F[:, 0, 0] = F[:, 1, 1] = F[:, 2, 2] = 1.0
        return {
            "x": self.x.copy(),
            "v": (np.zeros((n, 3), np.float32) if self.v is None else self.v.copy()),
            "C": np.zeros((n, 3, 3), np.float32), "F": F,
            "Jp": np.ones((n,), np.float32), "material": self.material.copy(),
        }


def elastic_sphere(n: int = 10_000, seed: int = 1, N: int = 16, compat=_lib.COMPAT_REFERENCE) -> Scene:
    s = Sampler(seed)
    r = 0.15
    x = sample_ball(s, n, (0.5, 0.6, 0.5), r)
    vol = 4.0 / 3.0 * math.pi * r ** 3 / n
    mat = material_as_numpy(_lib.MATERIAL_SNOW, vol, 1000.0, E=1e4, nu=0.3, hardening=0.0,
                            clamp_lo=0.0, clamp_hi=1e30)
    p = Params(dt=1e-4, N=N, gravity=-9.8, boundary_mode=_lib.BOUNDARY_CELLS, boundary=3,
               compat=compat)
    return Scene("elastic_sphere", p, [mat], x, np.zeros(n, np.uint8), substeps=100)


def snow_spheres(n: int = 500_000, seed: int = 2, N: int = 64, compat=_lib.COMPAT_REFERENCE) -> Scene:
    s = Sampler(seed)
    r1, r2 = 0.2, 0.13
    v1, v2 = r1 ** 3, r2 ** 3
    n1 = int(round(n * v1 / (v1 + v2)))
    n2 = n - n1
    x = np.concatenate([sample_ball(s, n1, (0.5, 0.3, 0.5), r1),
                        sample_ball(s, n2, (0.5, 0.62, 0.5), r2)])
    vol = 4.0 / 3.0 * math.pi * (v1 + v2) / n
    mat = material_as_numpy(_lib.MATERIAL_SNOW, vol, 400.0, E=1.4e5, nu=0.2, hardening=10.0,
                            clamp_lo=1.0 - 2.5e-2, clamp_hi=1.0 + 7.5e-3)
    p = Params(dt=1e-4, N=N, gravity=-9.8, compat=compat)
    return Scene("snow_spheres", p, [mat], x, np.zeros(n, np.uint8), substeps=1000,
                 extra={"counts": (n1, n2)})


def dam_break(n: int = 1_000_000, seed: int = 3, N: int = 128, compat=_lib.COMPAT_REFERENCE) -> Scene:
    s = Sampler(seed)
    lo, hi = (0.05, 0.05, 0.05), (0.45, 0.65, 0.95)
    x = sample_box(s, n, lo, hi)
    vol = 0.4 * 0.6 * 0.9 / n
    mat = material_as_numpy(_lib.MATERIAL_WC_LIQUID, vol, 1000.0, E=0.0, nu=0.0, hardening=0.0,
                            clamp_lo=0.0, clamp_hi=0.0, bulk_modulus=4e4)
    p = Params(dt=5e-5, N=N, gravity=-9.8, compat=compat)
    return Scene("dam_break", p, [mat], x, np.zeros(n, np.uint8), substeps=2000)


def _jittered_lattice(s: Sampler, side: int, lo, h: float):
    """side^3 lattice of spacing h from corner `lo` (cell centres), each point
    moved by up to +-h/4 per axis; float64."""
    n = side ** 3
    g = [(np.arange(side, dtype=np.float64) + 0.5) * h + lo[a] for a in range(3)]
    X, Y, Z = np.meshgrid(*g, indexing="ij")
    x = np.stack([X, Y, Z], -1).reshape(-1, 3)
    jit = (s.uniform(3 * n).astype(np.float64).reshape(n, 3) - 0.5) * 0.5 * h
    return x + jit


def _log_e(s: Sampler, n: int, noise: float) -> np.ndarray:
    """log E = log 1e4 + noise * N(0,1) per particle (Box-Muller on the stream)."""
    u1 = s.uniform(n).astype(np.float64)
    u2 = s.uniform(n).astype(np.float64)
    z = np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * math.pi * u2)
    return (math.log(1e4) + noise * z).astype(np.float32)


def elastic_cube(side: int = 32, seed: int = 5, N: int = 32, logE_noise: float = 0.1,
                 compat=_lib.COMPAT_REFERENCE) -> Scene:
    s = Sampler(seed)
    n = side ** 3
    h = 0.3 / side
    x = _jittered_lattice(s, side, (0.35, 0.35, 0.35), h).astype(np.float32)
    vol = 0.3 ** 3 / n
    logE = _log_e(s, n, logE_noise)
    mat = material_as_numpy(_lib.MATERIAL_SNOW, vol, 1000.0, E=1e4, nu=0.3, hardening=0.0,
                            clamp_lo=0.0, clamp_hi=1e30)
    p = Params(dt=1e-4, N=N, gravity=-9.8, compat=compat)
    return Scene("elastic_cube", p, [mat], x, np.zeros(n, np.uint8), substeps=200,
                 extra={"logE": logE, "target": (0.6, 0.4, 0.5)})


@dataclass
class DiffCube:
    """A cube for the differentiable substep (fixed-corotated, compat FIXED,
    per-particle log E)."""
    params: Params
    x: np.ndarray           # (n, 3) float32
    logE: np.ndarray        # (n,) float32
    nu: float
    particle_volume: float
    particle_mass: float
    target: tuple
    substeps: int


def diff_cube(side: int = 32, seed: int = 5, N: int = 32, lo=(0.35, 0.35, 0.35),
              spacing: float = 0.3 / 32, logE_noise: float = 0.1, substeps: int = 200) -> DiffCube:
    """The differentiable elastic cube: a jittered side^3 lattice of the given
    spacing from corner `lo` (defaults: 32^3 particles in [0.35, 0.65]^3),
    N=32, dt=1e-4, rho=1000, nu=0.3, log E = log 1e4 + 0.1 N(0,1), target
    (0.6, 0.4, 0.5). The particle volume is spacing^3."""
    s = Sampler(seed)
    n = side ** 3
    x = _jittered_lattice(s, side, lo, spacing).astype(np.float32)
    logE = _log_e(s, n, logE_noise)
    vol = float(np.float32(spacing ** 3))
    mass = float(np.float32(1000.0) * np.float32(vol))
    p = Params(dt=1e-4, N=N, gravity=-9.8, compat=_lib.COMPAT_FIXED)
    return DiffCube(p, x, logE, 0.3, vol, mass, (0.6, 0.4, 0.5), substeps)


def small_scene(kind: int, n: int, seed: int, N: int = 16, compat=_lib.COMPAT_REFERENCE,
                velocity=(0.0, 0.0, 0.0), center=(0.5, 0.5, 0.5), radius=0.15,
                boundary_mode=_lib.BOUNDARY_FRACTION, boundary=0.05) -> Scene:
    """A ball of n particles of one material kind (parity-test sized)."""
    s = Sampler(seed)
    x = sample_ball(s, n, center, radius)
    vol = 4.0 / 3.0 * math.pi * radius ** 3 / n
    if kind == _lib.MATERIAL_WC_LIQUID:
        mat = material_as_numpy(kind, vol, 1000.0, E=0.0, nu=0.0, hardening=0.0, clamp_lo=0.0,
                                clamp_hi=0.0, bulk_modulus=4e4)
    elif kind == _lib.MATERIAL_SNOW:
        mat = material_as_numpy(kind, vol, 400.0, E=1.4e5, nu=0.2, hardening=10.0,
                                clamp_lo=0.975, clamp_hi=1.0075)
    else:
        mat = material_as_numpy(kind, vol, 1000.0, E=2e4, nu=0.3, hardening=10.0)
    v = np.tile(np.asarray(velocity, np.float32), (n, 1))
    p = Params(dt=1e-4, N=N, gravity=-9.8, boundary_mode=boundary_mode, boundary=boundary,
               compat=compat)
    return Scene(f"ball_k{kind}", p, [mat], x, np.zeros(n, np.uint8), v=v)


def rubber_elastic(n: int, seed: int, N: int = 16, compat=_lib.COMPAT_REFERENCE,
                   velocity=(0.0, 0.0, 0.0)) -> Scene:
    """An elastic ball with the reference's rubber parameters
    (scenes/rubber_duck.toml:1-8): snow model, hardening 0, clamps [0, 1e30]."""
    s = Sampler(seed)
    r = 0.15
    x = sample_ball(s, n, (0.5, 0.5, 0.5), r)
    vol = 4.0 / 3.0 * math.pi * r ** 3 / n
    mat = material_as_numpy(_lib.MATERIAL_SNOW, vol, 200.0, E=1.4e5, nu=0.2, hardening=0.0,
                            clamp_lo=0.0, clamp_hi=1e30)
    v = np.tile(np.asarray(velocity, np.float32), (n, 1))
    p = Params(dt=1e-4, N=N, gravity=-9.81, compat=compat)
    return Scene("rubber_ball", p, [mat], x, np.zeros(n, np.uint8), v=v)


def batched_scenes(count: int = 64, seed: int = 4, scale: float = 1.0) -> List[Scene]:
    """The batched workload: `count` independent scenes. Per scene: particle
    count log-uniform in [5e3, 2e5] (times `scale`), N from {32, 48, 64},
    material 1/3 each elastic (E log-uniform [5e3, 5e4], nu=0.3, rho=1000,
    snow model with hardening 0 and no clamps), snow (as snow_spheres) and
    weakly compressible liquid (K=4e4, rho=1000); a ball or a box with a random
    centre and a size (radius or half edge) in [0.1, 0.25]; dt=1e-4, gravity
    -9.8, walls 0.05."""
    s = Sampler(seed)
    out = []
    for j in range(count):
        u = s.uniform(8).astype(np.float64)
        n = max(64, int(round(math.exp(math.log(5e3) + u[0] * (math.log(2e5) - math.log(5e3))) * scale)))
        N = (32, 48, 64)[min(2, int(u[1] * 3))]
        kind = min(2, int(u[2] * 3))
        size = 0.1 + u[3] * 0.15
        center = tuple(0.5 + (uu - 0.5) * (0.8 - 2 * size) * 0.5 for uu in u[4:7])
        if u[7] < 0.5:
            x = sample_ball(s, n, center, size)
            volume = 4.0 / 3.0 * math.pi * size ** 3
        else:
            c = np.asarray(center)
            x = sample_box(s, n, c - size, c + size)
            volume = (2 * size) ** 3
        vol = volume / n
        if kind == 0:
            E = math.exp(math.log(5e3) + float(s.uniform(1)[0]) * (math.log(5e4) - math.log(5e3)))
            mat = material_as_numpy(_lib.MATERIAL_SNOW, vol, 1000.0, E=E, nu=0.3, hardening=0.0,
                                    clamp_lo=0.0, clamp_hi=1e30)
        elif kind == 1:
            mat = material_as_numpy(_lib.MATERIAL_SNOW, vol, 400.0, E=1.4e5, nu=0.2, hardening=10.0,
                                    clamp_lo=1.0 - 2.5e-2, clamp_hi=1.0 + 7.5e-3)
        else:
            mat = material_as_numpy(_lib.MATERIAL_WC_LIQUID, vol, 1000.0, E=0.0, nu=0.0,
                                    hardening=0.0, clamp_lo=0.0, clamp_hi=0.0, bulk_modulus=4e4)
        p = Params(dt=1e-4, N=N, gravity=-9.8)
        out.append(Scene(f"batch{j}_k{kind}_N{N}", p, [mat], x, np.zeros(n, np.uint8)))
    return out


def algorithmic_bytes(P: int, N: int) -> int:
    """Minimal field traffic of one substep (dense FP32 grid): 254 B per
    particle (P2G reads x,v,C,F,Jp,mat = 101 B; G2P reads x,F,Jp,mat = 53 B and
    writes x,v,C,F,Jp = 100 B) plus 64 B per node (P2G write, grid-update read
    and write, G2P gather, 16 B each)."""
    return 254 * P + 64 * N ** 3
