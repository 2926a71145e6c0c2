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


# ---------------------------------------------------------------- marching cubes
# The case table restated from its construction (a closed surface: on each
# face, seen from outside, a segment from every outside -> inside crossing to
# the next crossing counter-clockwise; segments chained into loops; each loop
# cut into a fan from the first vertex whose diagonals leave the cube's
# faces). Pinned by tests/test_oracle_mesh.py: closed, consistently oriented
# surfaces with the right Euler characteristic and enclosed volume.
MC_EDGES = [(0, 4), (1, 5), (2, 6), (3, 7), (0, 2), (1, 3), (4, 6), (5, 7), (0, 1), (2, 3), (4, 5), (6, 7)]
MC_FACES = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]


def _edge(a, b):
    return MC_EDGES.index((min(a, b), max(a, b)))


def _face_edges():
    return [{_edge(f[i], f[(i + 1) % 4]) for i in range(4)} for f in MC_FACES]


def case_loops(case: int):
    """The surface loops of a case (bit q: corner q inside) as edge lists."""
    inside = [(case >> q) & 1 == 1 for q in range(8)]
    succ = {}
    for f in MC_FACES:
        cr = []
        for i in range(4):
            a, b = f[i], f[(i + 1) % 4]
            if inside[a] != inside[b]:
                cr.append((_edge(a, b), not inside[a]))  # (edge, outside -> inside)
        for i, (e, start) in enumerate(cr):
            if start:
                succ[e] = cr[(i + 1) % len(cr)][0]
    loops, seen = [], set()
    for e0 in sorted(succ):
        if e0 in seen:
            continue
        loop, e = [], e0
        while True:
            loop.append(e)
            seen.add(e)
            e = succ[e]
            if e == e0:
                break
        loops.append(loop)
    return loops


def case_table():
    """(ntri (256,), tri (256, 5, 3)) as the library's table lays it out."""
    fe = _face_edges()
    share = lambda a, b: any(a in s and b in s for s in fe)
    ntri = np.zeros(256, np.uint8)
    tri = np.zeros((256, 5, 3), np.int8)
    for c in range(256):
        t = 0
        for loop in case_loops(c):
            k = len(loop)
            r = next(r for r in range(k)
                     if not any(share(loop[r], loop[(r + q) % k]) for q in range(2, k - 1)))
            for q in range(1, k - 1):
                tri[c, t] = (loop[r], loop[(r + q) % k], loop[(r + q + 1) % k])
                t += 1
        ntri[c] = t
    return ntri, tri


def marching_cubes(S: np.ndarray, vdx: float):
    """Triangles of the iso-surface S = 0 (inside S > 0) of a (side,)^3 FP64
    field, in the library's order (cubes in index order, each case's
    triangles in table order): (n, 3, 3) float32 vertices, interpolated along
    each grid edge from its lower-index node, and (n, 3) uint64 edge keys
    (lower node << 32 | upper node)."""
    side = S.shape[0]
    ntri, tri = case_table()
    flat = S.reshape(-1)
    s1 = side - 1
    ii, jj, kk = np.meshgrid(np.arange(s1), np.arange(s1), np.arange(s1), indexing="ij")
    base = (ii * side + jj) * side + kk
    offs = [((q >> 2) & 1) * side * side + ((q >> 1) & 1) * side + (q & 1) for q in range(8)]
    nodes = np.stack([base + o for o in offs], -1).reshape(-1, 8)
    case = np.zeros(nodes.shape[0], np.int64)
    for q in range(8):
        case |= (flat[nodes[:, q]] > 0.0).astype(np.int64) << q
    xyz, keys = [], []
    for cube in np.nonzero(ntri[case])[0]:
        c = case[cube]
        for t in range(ntri[c]):
            for q in range(3):
                a, b = MC_EDGES[tri[c, t, q]]
                na, nb = int(nodes[cube, a]), int(nodes[cube, b])
                if nb < na:
                    na, nb = nb, na
                Sa, Sb = flat[na], flat[nb]
                tt = Sa / (Sa - Sb)
                pa = np.array(np.unravel_index(na, S.shape), np.float64)
                pb = np.array(np.unravel_index(nb, S.shape), np.float64)
                xyz.append(((pa + tt * (pb - pa)) * vdx).astype(np.float32))
                keys.append((na << 32) | nb)
    n = len(keys) // 3
    return np.array(xyz, np.float32).reshape(n, 3, 3), np.array(keys, np.uint64).reshape(n, 3)


def index_mesh(tri: np.ndarray, keys: np.ndarray):
    """(V (nv, 3) float64, F (nf, 3)) with one vertex per edge key, in key order."""
    order, inv = np.unique(keys.reshape(-1), return_inverse=True)
    V = np.zeros((len(order), 3))
    V[inv] = tri.reshape(-1, 3)
    return V, inv.reshape(-1, 3)


# ---------------------------------------------------------------- smoothing
def cot_laplacian(V, F):
    """igl::cotmatrix restated with scipy: L_jk = (cot a + cot b) / 2 from the
    two angles opposite edge jk (each from atan2 of the corner's vectors),
    L_jj = -sum_k L_jk."""
    import scipy.sparse as sp
    rows, cols, vals = [], [], []
    for f in F:
        for q in range(3):
            o, a, b = f[q], f[(q + 1) % 3], f[(q + 2) % 3]
            u, w = V[a] - V[o], V[b] - V[o]
            cot = np.dot(u, w) / np.linalg.norm(np.cross(u, w))
            rows += [a, b]
            cols += [b, a]
            vals += [0.5 * cot, 0.5 * cot]
    n = len(V)
    L = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    return L - sp.diags(np.asarray(L.sum(axis=1)).ravel())


def mass_barycentric(V, F):
    """igl::massmatrix BARYCENTRIC: M_ii = the area of the faces at i / 3."""
    import scipy.sparse as sp
    area = 0.5 * np.linalg.norm(np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]]), axis=1)
    m = np.zeros(len(V))
    for q in range(3):
        np.add.at(m, F[:, q], area / 3.0)
    return sp.diags(m)


def smooth_mesh(V, F):
    """MeshBuilder::smoothMesh (mesh_builder.h:215-246) with a direct sparse
    solve: U = (M - 0.001 L)^-1 M V, then U / sqrt(area of U)."""
    import scipy.sparse.linalg as spl
    L = cot_laplacian(V, F)
    M = mass_barycentric(V, F)
    U = spl.spsolve((M - 0.001 * L).tocsc(), M @ V)
    area = 0.5 * np.linalg.norm(np.cross(U[F[:, 1]] - U[F[:, 0]], U[F[:, 2]] - U[F[:, 0]]), axis=1).sum()
    return U / np.sqrt(area)
