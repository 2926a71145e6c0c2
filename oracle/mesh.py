"""Distance field of the particle mesher -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the reference's fillVoxelGrid_distance
(include/mesh_builder.h:103-149) with the voxel size and cut-off that
MeshBuilder::computeMesh gives it (mesh_builder.h:181-189), in FP32 like the
reference's `real`, for small inputs (a loop over particles). Pinned by
tests/test_oracle_mesh.py against the closed form (the Euclidean distance to
the nearest particle) and the reference's truncated voxel box.
"""
from __future__ import annotations

import numpy as np


def voxel_geometry(N: int, side: int, radius: int):
    """(voxel_dx as float32, cut-off as float32, iso level r * voxel_dx as
    float64): voxel_dx = double(float(N) * float(1/N)) / side
    (mesh_builder.h:182, params.N * params.dx in float), the cut-off
    (radius + 1) * voxel_dx (mesh_builder.h:189) passed as `real`."""
    dx = np.float32(1.0 / N)
    vdx = float(np.float32(N) * dx) / float(side)
    return np.float32(vdx), np.float32((radius + 1) * vdx), radius * vdx


def distance_field(x: np.ndarray, N: int, side: int, radius: int) -> np.ndarray:
    """side^3 float32 field, index [i, j, k] (mesh_builder.h:103-149)."""
    vdx, cutoff, _ = voxel_geometry(N, side, radius)
    g = np.full((side, side, side), cutoff, np.float32)
    rg = np.float32(cutoff / vdx)                       # r_gridpoints (:124)
    for p in np.asarray(x, np.float32):
        xg = p / vdx                                    # x_particle_gridpoints (:125)
        beg = np.maximum(0, (xg - rg).astype(np.int64))           # cast<int>: truncation (:128)
        end = np.minimum(side, (xg + (rg + np.float32(1.0))).astype(np.int64))  # (:129)
        if np.any(end <= beg):
            continue
        ii = np.arange(beg[0], end[0])
        jj = np.arange(beg[1], end[1])
        kk = np.arange(beg[2], end[2])
        nx = ii.astype(np.float32) * vdx                # x_node = i * dx (:139-145)
        ny = jj.astype(np.float32) * vdx
        nz = kk.astype(np.float32) * vdx
        # (particle.x - x_node).norm() (:144): Eigen sums the squares of a
        # 3-vector as e0 + (e1 + e2)
        d2 = (p[0] - nx)[:, None, None] ** 2 + ((p[1] - ny)[None, :, None] ** 2 + (p[2] - nz)[None, None, :] ** 2)
        d = np.sqrt(d2.astype(np.float32))
        sub = g[beg[0]:end[0], beg[1]:end[1], beg[2]:end[2]]
        np.minimum(sub, d, out=sub)                     # grid = min(d, grid) (:146)
    return g
