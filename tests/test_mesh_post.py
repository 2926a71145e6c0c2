"""Host-side mesh steps and particle files, on the CPU (mpm_amd.meshing over
the C-ABI; no GPU).

  * the library's marching-cubes case table equals the oracle's construction,
    and the oracle's marching cubes gives closed, consistently oriented
    surfaces: a sphere (Euler characteristic 2, enclosed volume within 3% of
    4/3 pi r^3, normals out) and random fields (every directed edge once and
    its opposite once);
  * mpm_mesh_smooth equals a direct sparse solve of MeshBuilder::smoothMesh
    (oracle.mesh.smooth_mesh: the cotangent weights from angles, not the
    library's edge-length form) to 1e-7, and the smoothed mesh has area 1;
  * mpm_mesh_decimate keeps a closed sphere closed (Euler characteristic 2),
    stops at the requested face count, keeps the vertices near the sphere and
    never collapses across the link condition;
  * mpm_bgeo_write writes the Houdini classic layout (gzip, big-endian
    header, attribute table, homogeneous positions, end bytes) and
    mpm_obj_write an OBJ that reads back.
libigl and Partio are not available here: the triangles, the smoothed
vertices and the collapse order are not compared with theirs (parity by the
properties above).
"""
import gzip
import struct
from collections import Counter

import numpy as np
import pytest

from oracle import mesh as om

meshing = pytest.importorskip("mpm_amd.meshing")


def sphere_field(side=20, r=6.3, c=(9.6, 9.2, 10.1)):
    g = np.arange(side, dtype=np.float64)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return r - np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)


def directed_edges(F):
    c = Counter()
    for f in F:
        for q in range(3):
            c[(int(f[q]), int(f[(q + 1) % 3]))] += 1
    return c


def closed_oriented(F):
    c = directed_edges(F)
    return all(v == 1 for v in c.values()) and all((b, a) in c for (a, b) in c)


def euler(V, F):
    return len(np.unique(F)) - len(directed_edges(F)) // 2 + len(F)


def volume(V, F):
    return np.einsum("ij,ij->i", V[F[:, 0]], np.cross(V[F[:, 1]], V[F[:, 2]])).sum() / 6.0


def test_case_table_matches_oracle_construction():
    ntri, tri = meshing.case_table()
    ont, otri = om.case_table()
    assert np.array_equal(ntri, ont)
    for c in range(256):
        assert np.array_equal(tri[c, :ntri[c]], otri[c, :ont[c]])
    # complementary cases cut the same edges; the empty and full cube nothing
    assert ntri[0] == ntri[255] == 0
    for c in range(256):
        cut = {e for e, (a, b) in enumerate(om.MC_EDGES) if ((c >> a) & 1) != ((c >> b) & 1)}
        assert set(tri[c, :ntri[c]].reshape(-1).tolist()) == cut


def test_oracle_marching_cubes_sphere_and_random_fields():
    S = sphere_field()
    tri, keys = om.marching_cubes(S, 1.0)
    V, F = om.index_mesh(tri, keys)
    assert closed_oriented(F) and euler(V, F) == 2
    vol = volume(V, F)
    assert vol > 0 and abs(vol / (4.0 / 3.0 * np.pi * 6.3 ** 3) - 1.0) < 0.03
    rng = np.random.default_rng(5)
    for _ in range(10):
        R = rng.standard_normal((7, 7, 7))
        R[[0, -1], :, :] = -1.0
        R[:, [0, -1], :] = -1.0
        R[:, :, [0, -1]] = -1.0
        tri, keys = om.marching_cubes(R, 1.0)
        _, F = om.index_mesh(tri, keys)
        assert closed_oriented(F)


def sphere_mesh():
    tri, keys = om.marching_cubes(sphere_field(), 0.05)
    return om.index_mesh(tri, keys)


def test_smoothing_matches_direct_solve():
    V, F = sphere_mesh()
    U = meshing.smooth(V, F)
    W = om.smooth_mesh(V, F)
    assert np.abs(U - W).max() < 1e-7 * np.abs(W).max()
    area = 0.5 * np.linalg.norm(np.cross(U[F[:, 1]] - U[F[:, 0]], U[F[:, 2]] - U[F[:, 0]]), axis=1).sum()
    assert abs(area - 1.0) < 1e-12
    assert closed_oriented(F) and volume(U, F) > 0


def test_decimation_keeps_a_closed_sphere():
    V, F = sphere_mesh()
    r = 6.3 * 0.05
    c = np.array([9.6, 9.2, 10.1]) * 0.05
    for target in (len(F) // 2, 400, 60):
        U, G = meshing.decimate(V, F, target)
        assert target - 2 <= len(G) <= target, (target, len(G))
        assert closed_oriented(G) and euler(U, G) == 2
        assert len(np.unique(G)) == len(U)
        d = np.linalg.norm(U - c, axis=1)
        assert np.abs(d - r).max() < 0.25 * r
        assert volume(U, G) > 0
    # nothing to do: the same mesh
    U, G = meshing.decimate(V, F, len(F))
    assert np.array_equal(G, F) and np.array_equal(U, V)


def test_decimation_of_a_tetrahedron_stops():
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
    F = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], np.int32)
    U, G = meshing.decimate(V, F, 0)
    assert len(G) == 4 and closed_oriented(G)


def test_bgeo_layout(tmp_path):
    x = np.array([[0.1, 0.2, 0.3], [0.4, 0.5, 0.6]], np.float32)
    v = np.array([[1.0, -2.0, 3.0], [0.0, 0.5, -0.25]], np.float32)
    p = tmp_path / "particles_0.bgeo"
    meshing.write_bgeo(str(p), x, v)
    raw = gzip.decompress(p.read_bytes())
    assert raw[:5] == b"BgeoV"
    counts = struct.unpack(">9i", raw[5:41])
    assert counts == (5, 2, 0, 0, 0, 3, 0, 0, 0)
    at = 41
    attrs = []
    for _ in range(3):
        (ln,) = struct.unpack(">H", raw[at:at + 2])
        name = raw[at + 2:at + 2 + ln].decode()
        size, typ = struct.unpack(">Hi", raw[at + 2 + ln:at + 8 + ln])
        at += 8 + ln + 4 * size
        attrs.append((name, size, typ))
    assert attrs == [("id", 1, 1), ("velocity", 3, 5), ("radius", 1, 0)]
    for i in range(2):
        rec = struct.unpack(">4fi3ff", raw[at:at + 36])
        at += 36
        assert np.allclose(rec[:3], x[i]) and rec[3] == 1.0 and rec[4] == i
        assert np.allclose(rec[5:8], v[i]) and rec[8] == np.float32(0.1)
    assert raw[at:] == b"\x00\xff"


def test_obj_round_trip(tmp_path):
    V, F = sphere_mesh()
    p = tmp_path / "m.obj"
    meshing.write_obj(str(p), V, F)
    lines = p.read_text().splitlines()
    v = np.array([[float(t) for t in l.split()[1:]] for l in lines if l.startswith("v ")])
    f = np.array([[int(t) for t in l.split()[1:]] for l in lines if l.startswith("f ")])
    assert np.array_equal(v, V) and np.array_equal(f - 1, F)


def test_bad_mesh_arguments():
    from mpm_amd import MPMError
    V = np.zeros((3, 3))
    with pytest.raises(MPMError, match="face index out of range"):
        meshing.smooth(V, np.array([[0, 1, 5]], np.int32))
    with pytest.raises(MPMError):
        meshing.decimate(V, np.array([[0, 1, 2]], np.int32), -1)
