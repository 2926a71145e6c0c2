"""The oracle's distance field of the particle mesher (oracle/mesh.py), pinned
on the CPU: every voxel of a particle's truncated box holds the Euclidean
distance to the particle, capped at the cut-off (closed form, FP64 within
FP32 rounding), voxels
outside every box hold the cut-off, several particles give the minimum, and
the voxel size follows N dx / side with dx = float(1/N)."""
import numpy as np

from oracle.mesh import distance_field, voxel_geometry


def test_voxel_geometry():
    vdx, cut, iso = voxel_geometry(64, 100, 5)
    dx = np.float32(1.0 / 64)
    assert vdx == np.float32(float(np.float32(64) * dx) / 100)
    assert cut == np.float32(6 * (float(np.float32(64) * dx) / 100))
    assert iso == 5 * (float(np.float32(64) * dx) / 100)


def test_single_particle_is_the_distance_in_its_box():
    N, side, r = 32, 40, 3
    p = np.array([[0.4123, 0.5771, 0.3009]], np.float32)
    g = distance_field(p, N, side, r)
    vdx, cut, _ = voxel_geometry(N, side, r)
    xg = p[0] / vdx
    rg = np.float32(cut / vdx)
    beg = np.maximum(0, (xg - rg).astype(int))
    end = np.minimum(side, (xg + rg + 1).astype(int))
    box = np.zeros_like(g, bool)
    box[beg[0]:end[0], beg[1]:end[1], beg[2]:end[2]] = True
    assert np.all(g[~box] == cut)
    ii, jj, kk = np.nonzero(box)
    nodes = np.stack([ii, jj, kk], 1).astype(np.float64) * float(vdx)
    exact = np.linalg.norm(nodes - p[0].astype(np.float64), axis=1)
    # min(d, cut-off): the field starts at the cut-off
    assert np.allclose(g[box], np.minimum(exact, float(cut)), rtol=2e-6, atol=0)
    assert (g[box] < cut).sum() > 50
    # the box reaches r + 1 voxels around the particle (truncated), no further
    assert (end - beg).max() <= 2 * (r + 1) + 1


def test_several_particles_take_the_minimum():
    N, side, r = 32, 32, 2
    rng = np.random.default_rng(3)
    p = (0.3 + 0.4 * rng.random((12, 3))).astype(np.float32)
    g = distance_field(p, N, side, r)
    singles = np.stack([distance_field(q[None], N, side, r) for q in p])
    assert np.array_equal(g, singles.min(0))
