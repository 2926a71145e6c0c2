#!/usr/bin/env python3
"""Write the golden step fixtures of tests/golden/ from the CPU oracle.

Each fixture is one small scene (a few hundred particles, k = 10 substeps):
its parameters, materials and initial state, and the state after k substeps
of the FP32 oracle (the reference's arithmetic restated on the CPU,
serial scatter order) and of the FP64 oracle. Stored as .npz with plain
numeric arrays only (loadable with numpy.load(allow_pickle=False)).

The reference's own tests hold no golden vectors for the step (its only
known-answer tests are the polar-decomposition checks of
tests/test_linalg.cu:25-91), so these fixtures are the oracle's: they pin
the oracle against regressions and give the GPU engine fixed vectors to
meet, at sizes that need no oracle run at test time.

Usage: python tests/golden/make_golden.py   (needs the built oracle)
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from mpm_amd import scenes  # noqa: E402
from mpm_amd._lib import (COMPAT_FIXED, COMPAT_REFERENCE, MATERIAL_JELLY,  # noqa: E402
                          MATERIAL_SNOW, MATERIAL_WC_LIQUID, mpm_material)

K = 10
MAT_FIELDS = [name for name, _ in mpm_material._fields_]


def perturb(st, seed):
    rng = np.random.default_rng(seed)
    n = st["x"].shape[0]
    st["C"] = rng.uniform(-2, 2, (n, 3, 3)).astype(np.float32)
    st["F"] = (np.eye(3) + 0.04 * rng.uniform(-1, 1, (n, 3, 3))).astype(np.float32)
    st["Jp"] = rng.uniform(0.95, 1.05, n).astype(np.float32)
    return st


def cases():
    out = {}
    sc = scenes.rubber_elastic(300, 101, N=16, velocity=(0.4, -0.7, 0.1))
    out["rubber_ref"] = (sc, sc.state())
    sc = scenes.small_scene(MATERIAL_SNOW, 300, 102, N=16, velocity=(0.0, -1.5, 0.3))
    out["snow_ref"] = (sc, perturb(sc.state(), 12))
    sc = scenes.small_scene(MATERIAL_WC_LIQUID, 300, 103, N=16)
    out["liquid_ref"] = (sc, sc.state())
    sc = scenes.small_scene(MATERIAL_JELLY, 300, 104, N=16, compat=COMPAT_FIXED,
                            velocity=(0.8, 0.2, -0.5))
    out["jelly_fixed"] = (sc, perturb(sc.state(), 14))
    # near the top wall and the high grid faces: node bounds and row aliasing
    sc = scenes.small_scene(MATERIAL_SNOW, 300, 105, N=16, compat=COMPAT_REFERENCE,
                            center=(0.85, 0.9, 0.88), radius=0.1)
    out["edges_ref"] = (sc, sc.state())
    return out


def write(name, sc, st):
    p = sc.params
    mats = np.array([[float(getattr(m, f)) for f in MAT_FIELDS[1:]] for m in sc.materials], np.float64)
    kinds = np.array([m.kind for m in sc.materials], np.int32)
    o32 = oracle.advance(p, sc.materials, st, K, np.float32)
    o64 = oracle.advance(p, sc.materials, st, K, np.float64)
    arrs = dict(
        params_f=np.array([p.dt, p.gravity, p.boundary], np.float64),
        params_i=np.array([p.N, p.boundary_mode, p.compat, K], np.int64),
        mat_kind=kinds, mat_fields=mats,
        x=st["x"], v=st["v"], C=st["C"], F=st["F"], Jp=st["Jp"], material=st["material"],
    )
    for tag, o in (("f32", o32), ("f64", o64)):
        for k in ("x", "v", "C", "F", "Jp"):
            arrs[f"{tag}_{k}"] = o[k]
    np.savez_compressed(HERE / f"{name}.npz", **arrs)
    print(f"wrote {name}.npz ({sc.n} particles, N={p.N}, k={K})")


if __name__ == "__main__":
    oracle.build()
    for name, (sc, st) in cases().items():
        write(name, sc, st)
