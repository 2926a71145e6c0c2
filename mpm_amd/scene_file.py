"""Scene files: the reference's TOML scenes with OBJ meshes, loaded and
sampled by the library (mpm_scene_* in include/mpm_hip.h; reference
src/main.cu:31-66, src/mpm.cu:331-394). Host only; no GPU needed."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List

import numpy as np

from ._lib import MPMError, lib, mpm_material


@dataclass
class SceneObject:
    material: int
    first: int
    count: int
    lifetime_begin: float
    lifetime_end: float

    def is_active(self, t: float) -> bool:
        """SimObject::isActive (include/mpm.cuh:38-40)."""
        return self.lifetime_begin <= t < self.lifetime_end


@dataclass
class SceneFile:
    materials: List[mpm_material]
    objects: List[SceneObject]
    x: np.ndarray          # (n, 3) float32, objects in file order
    v: np.ndarray          # (n, 3) float32
    material: np.ndarray   # (n,) uint8


def load_scene(path: str, particle_count: int = 500000, seed: int = 1) -> SceneFile:
    """Parse `path` and fill its objects with particles at `particle_count`
    per unit volume (the reference's --particle-count, default 500000);
    `seed` seeds the glibc rand() sequence (1: the reference's unseeded one)."""
    l = lib()
    h = C.c_void_p()
    rc = l.mpm_scene_load(str(path).encode(), int(particle_count), int(seed), C.byref(h))
    if rc:
        raise MPMError(rc, l.mpm_scene_last_error().decode())
    try:
        nm, no, n = C.c_int32(), C.c_int32(), C.c_int64()
        l.mpm_scene_info(h, C.byref(nm), C.byref(no), C.byref(n))
        mats = (mpm_material * max(1, nm.value))()
        l.mpm_scene_materials(h, mats)
        objs = []
        for i in range(no.value):
            m, f, c = C.c_int32(), C.c_int64(), C.c_int64()
            b, e = C.c_double(), C.c_double()
            l.mpm_scene_object(h, i, C.byref(m), C.byref(f), C.byref(c), C.byref(b), C.byref(e))
            objs.append(SceneObject(m.value, f.value, c.value, b.value, e.value))
        x = np.zeros((n.value, 3), np.float32)
        v = np.zeros((n.value, 3), np.float32)
        mat = np.zeros(n.value, np.uint8)
        fp = C.POINTER(C.c_float)
        l.mpm_scene_particles_soa(h, x.ctypes.data_as(fp), v.ctypes.data_as(fp),
                                  mat.ctypes.data_as(C.POINTER(C.c_uint8)))
        return SceneFile(list(mats)[:nm.value], objs, x, v, mat)
    finally:
        l.mpm_scene_destroy(h)
