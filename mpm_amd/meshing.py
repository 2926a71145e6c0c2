"""Host-side mesh steps and particle files (the C-ABI's mpm_mesh_smooth,
mpm_mesh_decimate, mpm_obj_write, mpm_bgeo_write, mpm_mesh_case_table).

Vertices are (nv, 3) float64 arrays, faces (nf, 3) int32 arrays of 0-based
vertex indices. These replace the libigl calls of the reference's
MeshBuilder (include/mesh_builder.h:196-246) and the Partio call of its
ParticleWriter (include/particle_writer.h:13-44).
"""
from __future__ import annotations

import ctypes as C
from typing import Tuple

import numpy as np

from ._lib import check, lib


def _mesh(V, F):
    V = np.ascontiguousarray(V, np.float64).reshape(-1, 3)
    F = np.ascontiguousarray(F, np.int32).reshape(-1, 3)
    return V, F


def case_table() -> Tuple[np.ndarray, np.ndarray]:
    """(ntri (256,) uint8, tri (256, 5, 3) int8): the marching-cubes cases."""
    ntri = np.empty(256, np.uint8)
    tri = np.empty((256, 5, 3), np.int8)
    check(lib().mpm_mesh_case_table(ntri.ctypes.data, tri.ctypes.data))
    return ntri, tri


def smooth(V, F) -> np.ndarray:
    """MeshBuilder::smoothMesh (mesh_builder.h:215-246): the smoothed vertices."""
    V, F = _mesh(V, F)
    U = V.copy()
    check(lib().mpm_mesh_smooth(len(U), U.ctypes.data, len(F), F.ctypes.data))
    return U


def decimate(V, F, max_faces: int) -> Tuple[np.ndarray, np.ndarray]:
    """igl::decimate with edge length and midpoint (mesh_builder.h:203-207)."""
    V, F = _mesh(V, F)
    U = np.empty_like(V)
    G = np.empty_like(F)
    nu, ng = C.c_int64(), C.c_int64()
    check(lib().mpm_mesh_decimate(len(V), V.ctypes.data, len(F), F.ctypes.data, int(max_faces), U.ctypes.data,
                                  C.byref(nu), G.ctypes.data, C.byref(ng)))
    return U[:nu.value].copy(), G[:ng.value].copy()


def write_obj(path: str, V, F) -> None:
    V, F = _mesh(V, F)
    check(lib().mpm_obj_write(str(path).encode(), len(V), V.ctypes.data, len(F), F.ctypes.data))


def write_bgeo(path: str, x, v) -> None:
    """Particles as the reference's ParticleWriter writes them (.bgeo)."""
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
    v = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
    if len(v) != len(x):
        raise ValueError("x and v need the same length")
    check(lib().mpm_bgeo_write(str(path).encode(), len(x), x.ctypes.data, v.ctypes.data))
