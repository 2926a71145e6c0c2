"""The scene input path on the CPU (mpm_scene_* through mpm_amd.scene_file):
TOML scene files, OBJ meshes and the rejection sampler.

Pins:
  * the draws are glibc's rand() sequence: compared with the C library's own
    srand()/rand() for several seeds;
  * a cube mesh accepts every draw, so its particles are exactly
    position + u * size from consecutive triples of that sequence, and there
    are u32(u32(1/volume) * size^3) of them;
  * a sphere mesh keeps the draws inside its (polyhedral) ball, about pi/6 of
    them; a mesh with inward-facing triangles keeps none (winding number -1);
  * materials carry the reference's defaults and 1/particle_count volume;
    lifetimes default to [0, max);
  * malformed files and unknown material names are errors that say where.
"""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

import mpm_amd
from mpm_amd import MATERIAL_SNOW, MATERIAL_WC_LIQUID, MPMError, material_as_numpy
from mpm_amd.scene_file import load_scene

CUBE_V = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
# outward-facing quads (counter-clockwise seen from outside)
CUBE_F = [(1, 4, 3, 2), (5, 6, 7, 8), (1, 2, 6, 5), (2, 3, 7, 6), (3, 4, 8, 7), (4, 1, 5, 8)]


def write_obj(path: Path, V, F):
    with open(path, "w") as f:
        f.write("# test mesh\n")
        for v in V:
            f.write("v %r %r %r\n" % tuple(float(c) for c in v))
        for face in F:
            f.write("f " + " ".join(f"{i}/{i}/{i}" for i in face) + "\n")


def icosphere(levels=2):
    t = (1 + 5 ** 0.5) / 2
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [np.array(v, float) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(levels):
        mid = {}

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = V[a] + V[b]
                V.append(p / np.linalg.norm(p))
                mid[key] = len(V) - 1
            return mid[key]

        F2 = []
        for a, b, c in F:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            F2 += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = F2
    return [tuple(v) for v in V], [tuple(i + 1 for i in f) for f in F]


SCENE = """\
# two materials, three objects
[[material]]
name = "snow"
density = 700
E = 1.4e5
Nu = 0.2
hardening = 10
plast_clamp_lower = 0.975
plast_clamp_higher = 1.0075

[[material]]
name = "water"   # trailing comment
model = "wc_liquid"
density = 1000
bulk_modulus = 4e4

[[object]]
material = "snow"
mesh = "cube.obj"
size = 0.2
position = [0.3, 0.3, 0.3]
velocity = [1.0, -2.0, 0.5]

[[object]]
material = "water"
mesh = "ball.obj"
size = 0.25
position = [0.5, 0.1,
            0.5]
velocity = [0.0, 0.0, 0.0]
lifetime_begin = 0.5
lifetime_end = 2.0

[[object]]
material = "snow"
mesh = "inverted.obj"
size = 0.1
position = [0.1, 0.1, 0.1]
velocity = [0, 0, 0]
"""


@pytest.fixture()
def scene_dir(tmp_path):
    (tmp_path / "meshes").mkdir()
    write_obj(tmp_path / "meshes" / "cube.obj", CUBE_V, CUBE_F)
    V, F = icosphere(2)
    write_obj(tmp_path / "meshes" / "ball.obj", V, F)
    write_obj(tmp_path / "meshes" / "inverted.obj", CUBE_V, [tuple(reversed(f)) for f in CUBE_F])
    (tmp_path / "scene.toml").write_text(SCENE)
    return tmp_path


def libc_rand(seed, n):
    libc = C.CDLL("libc.so.6")
    libc.srand(C.c_uint(seed))
    return np.array([libc.rand() for _ in range(n)], np.int64)


@pytest.mark.parametrize("seed", [1, 0, 7, 123456789])
def test_generator_is_glibc_rand(seed):
    n = 5000
    out = (C.c_int32 * n)()
    assert mpm_amd.lib().mpm_debug_glibc_rand(seed, n, out) == 0
    assert np.array_equal(np.array(out[:], np.int64), libc_rand(seed, n))


def test_cube_particles_are_the_draws(scene_dir):
    pc = 100000
    sc = load_scene(str(scene_dir / "scene.toml"), pc, seed=1)
    o = sc.objects[0]
    vol = np.float32(1.0 / pc)
    density = int(1.0 / float(vol))
    # the unit cube placed in float32: vertex * 0.2f + 0.3f, so [0.3f, 0.2f + 0.3f]^3;
    # its bounding-box extent in double
    lo = np.float32(0.3)
    hi = np.float32(0.2) + np.float32(0.3)
    rng = float(hi) - float(lo)
    target = int(density * rng ** 3)
    assert o.first == 0 and o.count == target
    u = libc_rand(1, 3 * target).astype(np.float32) / np.float32(2147483647)
    expect = (float(lo) + u.astype(np.float64) * rng).astype(np.float32).reshape(-1, 3)
    assert np.array_equal(sc.x[:target], expect)
    assert np.all(sc.v[:target] == np.float32([1.0, -2.0, 0.5]))
    assert np.all(sc.material[:target] == 0)


def test_sphere_and_inverted_mesh(scene_dir):
    sc = load_scene(str(scene_dir / "scene.toml"), 200000, seed=3)
    ball = sc.objects[1]
    x = sc.x[ball.first:ball.first + ball.count]
    # the ball spans [0.5, 0.75] x [0.1, 0.35] x [0.5, 0.75] (size = diameter)
    c = np.array([0.625, 0.225, 0.625])
    r = np.linalg.norm(x - c, axis=1)
    assert r.max() <= 0.125 * (1 + 1e-6)
    drawn = int(int(1.0 / float(np.float32(1.0 / 200000))) * 0.25 ** 3)
    frac = ball.count / drawn
    assert 0.45 < frac < math.pi / 6, frac  # the icosphere is slightly smaller than the ball
    assert np.all(sc.material[ball.first:ball.first + ball.count] == 1)
    assert (ball.lifetime_begin, ball.lifetime_end) == (0.5, 2.0)
    assert sc.objects[2].count == 0  # faces inward: winding number -1 everywhere inside


def test_materials_and_defaults(scene_dir):
    pc = 500000
    sc = load_scene(str(scene_dir / "scene.toml"), pc)
    snow = material_as_numpy(MATERIAL_SNOW, 1.0 / pc, 700.0, E=1.4e5, nu=0.2, hardening=10.0,
                             clamp_lo=0.975, clamp_hi=1.0075)
    water = material_as_numpy(MATERIAL_WC_LIQUID, 1.0 / pc, 1000.0, E=1.4e5, nu=0.2,
                              hardening=10.0, clamp_lo=0.975, clamp_hi=1.0075, bulk_modulus=4e4)
    for got, want in zip(sc.materials, (snow, water)):
        for name, _ in mpm_amd.mpm_material._fields_:
            assert getattr(got, name) == getattr(want, name), name
    o = sc.objects[0]
    assert o.lifetime_begin == 0.0 and o.lifetime_end > 1e300
    assert o.is_active(0.0) and not sc.objects[1].is_active(0.2) and sc.objects[1].is_active(0.5)


@pytest.mark.parametrize("text,needle", [
    ("[[material]]\nname = \"a\"\n[[object]]\nmaterial = \"b\"\nposition=[0,0,0]\nvelocity=[0,0,0]\n",
     "unknown material 'b'"),
    ("[[material]]\nname = \"a\"\ndensity = \n", "line 3"),
    ("[[material]]\nname = \"a\"\nE = [1, 2\n", "line 4"),
    ("[[material]]\nname = \"a\"\nE = \"x\"\n", "'E' must be a number"),
    ("[[material]]\nname = \"a\"\n[[object]]\nmaterial = \"a\"\nmesh = \"none.obj\"\n"
     "position=[0,0,0]\nvelocity=[0,0,0]\n", "cannot open mesh"),
    ("[[material]]\nname = \"a\"\n[[object]]\nmaterial = \"a\"\nmesh = \"cube.obj\"\n"
     "position=[0,0]\nvelocity=[0,0,0]\n", "'position' must be an array of 3 numbers"),
    ("[[material]]\nname = \"a\"\nmodel = \"rubber\"\n", "unknown material model"),
])
def test_errors_name_the_problem(scene_dir, text, needle):
    p = scene_dir / "bad.toml"
    p.write_text(text)
    with pytest.raises(MPMError) as e:
        load_scene(str(p), 1000)
    assert needle in str(e.value)


def test_missing_file():
    with pytest.raises(MPMError, match="cannot open"):
        load_scene("/nonexistent/scene.toml", 1000)
