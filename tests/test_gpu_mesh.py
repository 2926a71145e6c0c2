"""The surface mesh of the particles on the GPU (mpm_mesh_*, Engine.mesh).

  * the distance field equals the oracle's restatement of the reference's
    fillVoxelGrid_distance bit for bit (oracle/mesh.py);
  * the marching-cubes triangles equal the oracle's restatement
    (oracle.mesh.marching_cubes on the downloaded field) bit for bit, in the
    same order;
  * the surface of one particle is a closed, consistently oriented sphere:
    every edge in two triangles, once each way, Euler characteristic 2,
    vertices near the iso distance r voxel_dx, enclosed volume (divergence
    theorem, positive: normals out) near 4/3 pi (r voxel_dx)^3;
  * every vertex is the zero of S = r voxel_dx - d interpolated along its grid
    edge, from the downloaded field;
  * a scene's mesh is closed, and its OBJ file lists one vertex per crossed
    grid edge and the same faces; smoothed and decimated as MeshBuilder does
    (--laplacian_smooth, --mesh-face-count) it stays closed.
The reference meshes with libigl's marching cubes (absent here, with its own
case table): its triangles are not reproduced, the surface is (parity by
these properties).
"""
import math
from collections import Counter

import numpy as np
import pytest

from mpm_amd import Engine, Params, make_material, MATERIAL_SNOW, scenes
from oracle.mesh import distance_field, voxel_geometry

pytestmark = pytest.mark.gpu


def engine_with(x, N):
    e = Engine(Params(dt=1e-4, N=N))
    e.set_materials([make_material(MATERIAL_SNOW, 1e-6, 400.0)])
    e.add_particles(np.asarray(x, np.float32), np.zeros(len(x), np.uint8))
    e.upload()
    return e


def edge_counts(tri_keys):
    """directed edges between vertex keys: (a, b) per triangle side"""
    c = Counter()
    for t in tri_keys:
        for q in range(3):
            c[(int(t[q]), int(t[(q + 1) % 3]))] += 1
    return c


def check_closed(tri_keys):
    c = edge_counts(tri_keys)
    assert all(v == 1 for v in c.values())          # each directed edge once
    assert all((b, a) in c for (a, b) in c)         # and its opposite: closed, oriented


def volume(tri):
    t = tri.astype(np.float64)
    return np.einsum("ij,ij->i", t[:, 0], np.cross(t[:, 1], t[:, 2])).sum() / 6.0


def test_distance_field_matches_oracle(gpu_lib):
    N, side, r = 32, 48, 3
    rng = np.random.default_rng(7)
    x = (0.2 + 0.6 * rng.random((300, 3))).astype(np.float32)
    e = engine_with(x, N)
    e.mesh(side, r)
    got = e.mesh_distance_field(side)
    e.close()
    want = distance_field(x, N, side, r)
    assert np.array_equal(got, want)


def test_marching_cubes_matches_oracle(gpu_lib):
    from oracle.mesh import marching_cubes
    N, side, r = 32, 40, 3
    rng = np.random.default_rng(9)
    x = (0.25 + 0.5 * rng.random((200, 3))).astype(np.float32)
    e = engine_with(x, N)
    m = e.mesh(side, r)
    d = e.mesh_distance_field(side)
    e.close()
    _, _, iso = voxel_geometry(N, side, r)
    vdx = float(N * np.float32(1.0 / N)) / side
    tri, keys = marching_cubes(iso - d.astype(np.float64), vdx)
    assert len(tri) > 500
    assert np.array_equal(m["edges"], keys)
    assert np.array_equal(m["triangles"], tri)


def test_single_particle_sphere(gpu_lib):
    N, side, r = 32, 64, 6
    x = np.array([[0.50321, 0.4871, 0.51234]], np.float32)
    e = engine_with(x, N)
    m = e.mesh(side, r)
    d = e.mesh_distance_field(side)
    e.close()
    tri, keys = m["triangles"], m["edges"]
    assert len(tri) > 200
    check_closed(keys)
    V = len(np.unique(keys))
    E = len(edge_counts(keys)) // 2
    assert V - E + len(tri) == 2
    vdx, _, iso = voxel_geometry(N, side, r)
    dist = np.linalg.norm(tri.reshape(-1, 3).astype(np.float64) - x[0], axis=1)
    # (the field is linear along each edge between exact node distances: the
    # vertices sit within a tenth of a voxel of the sphere at r = 6 voxels)
    assert np.abs(dist - iso).max() < 0.15 * float(vdx)
    vol = volume(tri)
    assert vol > 0
    assert abs(vol / (4.0 / 3.0 * math.pi * iso ** 3) - 1.0) < 0.03
    # every vertex is the zero of S along its grid edge
    S = iso - d.reshape(-1).astype(np.float64)
    a = (keys.reshape(-1) >> np.uint64(32)).astype(np.int64)
    b = (keys.reshape(-1) & np.uint64(0xffffffff)).astype(np.int64)
    t = S[a] / (S[a] - S[b])
    pa = np.stack(np.unravel_index(a, (side,) * 3), 1).astype(np.float64)
    pb = np.stack(np.unravel_index(b, (side,) * 3), 1).astype(np.float64)
    want = ((pa + t[:, None] * (pb - pa)) * (float(N * np.float32(1.0 / N)) / side)).astype(np.float32)
    assert np.allclose(tri.reshape(-1, 3), want, rtol=0, atol=1e-6)


def test_scene_mesh_is_closed_and_obj_matches(gpu_lib, tmp_path):
    sc = scenes.snow_spheres(n=40_000, N=32)
    st = sc.state()
    e = Engine(sc.params)
    e.set_materials(sc.materials)
    e.add_particles(st["x"], st["material"], st["v"], st["C"], st["F"], st["Jp"])
    e.upload()
    e.advance(20)
    m = e.mesh(64, 2)
    path = tmp_path / "mesh.obj"
    e.write_mesh_obj(str(path))
    e.close()
    tri, keys = m["triangles"], m["edges"]
    assert len(tri) > 1000
    check_closed(keys)
    assert volume(tri) > 0
    lines = path.read_text().splitlines()
    verts = np.array([[float(v) for v in l.split()[1:]] for l in lines if l.startswith("v ")], np.float32)
    faces = np.array([[int(v) for v in l.split()[1:]] for l in lines if l.startswith("f ")], np.int64)
    assert len(verts) == len(np.unique(keys)) and len(faces) == len(tri)
    assert faces.min() == 1 and faces.max() == len(verts)
    assert np.allclose(verts[faces - 1], tri, rtol=0, atol=1e-6)


def test_mesh_pipeline_smooth_and_decimate(gpu_lib, tmp_path):
    from mpm_amd import meshing
    sc = scenes.snow_spheres(n=40_000, N=32)
    st = sc.state()
    e = Engine(sc.params)
    e.set_materials(sc.materials)
    e.add_particles(st["x"], st["material"], st["v"], st["C"], st["F"], st["Jp"])
    e.upload()
    m = e.mesh(64, 2)
    V, F = e.mesh_indexed()
    e.close()
    assert len(F) == len(m["triangles"]) and len(V) == len(np.unique(m["edges"]))
    assert np.allclose(V[F], m["triangles"], rtol=0, atol=0)
    check_closed(F)
    U = meshing.smooth(V, F)
    area = 0.5 * np.linalg.norm(np.cross(U[F[:, 1]] - U[F[:, 0]], U[F[:, 2]] - U[F[:, 0]]), axis=1).sum()
    assert abs(area - 1.0) < 1e-9
    U2, G = meshing.decimate(U, F, len(F) // 4)
    assert len(F) // 4 - 2 <= len(G) <= len(F) // 4
    check_closed(G)
    meshing.write_obj(str(tmp_path / "m.obj"), U2, G)


def test_mesh_errors_and_empty_set(gpu_lib):
    from mpm_amd import MPMError
    e = Engine(Params(dt=1e-4, N=32))
    e.set_materials([make_material(MATERIAL_SNOW, 1e-6, 400.0)])
    with pytest.raises(MPMError, match="nothing uploaded"):
        e.mesh(32, 2)
    e.upload()  # no particles: an empty mesh
    m = e.mesh(32, 2)
    assert len(m["triangles"]) == 0
    with pytest.raises(MPMError, match="grid side"):
        e.mesh(1, 2)
    with pytest.raises(MPMError, match="grid side 32"):
        e.mesh_distance_field(16)
    assert np.all(e.mesh_distance_field(32) == voxel_geometry(32, 32, 2)[1])
    e.close()
