"""The `mpm` program and the C++ Simulation facade (include/mpm_simulation.hpp).

CPU: flag handling, and a scene that loads but finds no GPU fails with the
library's message. GPU: a scene whose second object appears mid-run
(lifetime_begin) advanced by the program, against the same sequence done
through the Python binding (same kernels, so the same bits): particles of
the first object alone for the substeps before the object's start time,
then the second object appended on the device behind them.
"""
import os
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from test_scene_file import CUBE_F, CUBE_V, icosphere, write_obj

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "mpm_amd" / "bin" / "mpm"

SCENE = """\
[[material]]
name = "rubber"
density = 200
E = 1.4e5
Nu = 0.2
hardening = 0
plast_clamp_lower = 0.0
plast_clamp_higher = 1.0e30

[[material]]
name = "snow"

[[object]]
material = "rubber"
mesh = "ball.obj"
size = 0.3
position = [0.3, 0.2, 0.3]
velocity = [0.5, 0.0, 0.0]

[[object]]
material = "snow"
mesh = "cube.obj"
size = 0.15
position = [0.4, 0.6, 0.4]
velocity = [0.0, -1.0, 0.0]
lifetime_begin = 0.01
"""


@pytest.fixture()
def scene(tmp_path):
    (tmp_path / "meshes").mkdir()
    V, F = icosphere(2)
    write_obj(tmp_path / "meshes" / "ball.obj", V, F)
    write_obj(tmp_path / "meshes" / "cube.obj", CUBE_V, CUBE_F)
    p = tmp_path / "scene.toml"
    p.write_text(SCENE)
    return p


def run(*args, timeout=120):
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=timeout)


def test_help_and_bad_flags():
    assert CLI.exists(), "build the package first (python -m mpm_amd.build)"
    r = run("--help")
    assert r.returncode == 0 and "--particle-count" in r.stdout
    r = run("--bogus", "1")
    assert r.returncode == 1 and "unknown option 'bogus'" in r.stdout
    r = run("--dt", "abc")
    assert r.returncode == 1 and "bad number" in r.stdout


# The reference README's first example (README.md:14), scene
# scenes/rubber_duck.toml: its material and objects, with stand-in meshes (the
# reference's OBJ files are Git-LFS pointers): an icosphere for the duck.
RUBBER_DUCK = """\
[[material]]
name = "rubber"
density = 200
E = 1.4e5
Nu = 0.2
hardening = 0
plast_clamp_lower = 0.0
plast_clamp_higher = 1.0e30

[[object]]
material = "rubber"
mesh = "rubber_duck.obj"
size = 0.4
position = [0.3, 0.4, 0.3]
velocity = [0.0, 0.0, 0.0]

[[object]]
material = "rubber"
mesh = "cube.obj"
size = 0.3
position = [0.35, 0.5, 0.35]
velocity = [0.0, 0.0, 0.0]
lifetime_begin = 2.0
"""


@pytest.fixture()
def rubber_duck(tmp_path):
    (tmp_path / "meshes").mkdir()
    V, F = icosphere(2)
    write_obj(tmp_path / "meshes" / "rubber_duck.obj", V, F)
    write_obj(tmp_path / "meshes" / "cube.obj", CUBE_V, CUBE_F)
    p = tmp_path / "rubber_duck.toml"
    p.write_text(RUBBER_DUCK)
    return p


def test_reference_short_flag_forms(rubber_duck):
    """cxxopts 2.2 gives the one-letter option "N" only its short form: the
    reference's documented runs pass `-N 16` (README.md:14-16). The program
    takes it, -N16 and --N alike; a missing or bad value fails."""
    for form in (["-N", "16"], ["-N16"], ["--N", "16"], ["--N=16"]):
        r = run("--scene", rubber_duck, *form, "--particle-count", 10000, "--steps", 1, "--no-mesh")
        assert "grid: N=16," in r.stdout, (form, r.stdout, r.stderr)
    r = run("--scene", rubber_duck, "-N")
    assert r.returncode == 1 and "needs a value" in r.stdout
    r = run("--scene", rubber_duck, "-N", "abc")
    assert r.returncode == 1 and "bad integer" in r.stdout


@pytest.mark.gpu
def test_reference_readme_command_line(rubber_duck, tmp_path, gpu_lib):
    """README.md:14's run, `--scene scenes/rubber_duck.toml -N 16
    --particle-count 10000`, for 12 substeps with frames saved: the program's
    last frame equals the same scene advanced through the Python binding on a
    16^3 grid (the second object starts at t = 2 s, outside the run)."""
    from mpm_amd import Engine, Params
    from mpm_amd.scene_file import load_scene

    out = tmp_path / "out"
    steps, dt = 12, 1e-4
    r = run("--scene", rubber_duck, "-N", 16, "--particle-count", 10000, "--steps", steps,
            "--save-dir", out, "--no-mesh")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "grid: N=16," in r.stdout
    frames = particle_frames(out)
    assert len(frames) == 1  # 240 frames per second: substep 0 only
    sc = load_scene(str(rubber_duck), 10000, seed=1)
    a = sc.objects[0]
    eng = Engine(Params(dt=dt, N=16, gravity=-9.81))
    eng.set_materials(sc.materials)
    sl = slice(a.first, a.first + a.count)
    eng.add_particles(sc.x[sl], sc.material[sl], sc.v[sl])
    eng.upload()
    eng.advance(1)
    ref = eng.download()
    eng.close()
    x0, v0, ids = read_bgeo(frames[0])
    assert x0.shape[0] == a.count and np.array_equal(ids, np.arange(a.count))
    assert np.array_equal(x0, ref["x"]) and np.array_equal(v0, ref["v"])


def test_missing_scene_fails_cleanly(tmp_path):
    r = run("--scene", tmp_path / "none.toml", "--steps", "1")
    assert r.returncode == 1 and "cannot open" in r.stderr


def particle_frames(out):
    """particles_<frame>.bgeo files (frame numbers not zero-padded, as the
    reference names them, main.cu:109) in frame order"""
    return sorted((out / "particles").glob("particles_*.bgeo"), key=lambda p: int(p.stem.split("_")[1]))


def read_bgeo(path):
    """x, v, id of a .bgeo the program wrote (id, velocity, radius attributes)."""
    import gzip
    import struct
    raw = gzip.decompress(Path(path).read_bytes())
    assert raw[:5] == b"BgeoV"
    n = struct.unpack(">i", raw[9:13])[0]
    at = 41
    for _ in range(3):  # id, velocity, radius
        (ln,) = struct.unpack(">H", raw[at:at + 2])
        size = struct.unpack(">H", raw[at + 2 + ln:at + 4 + ln])[0]
        at += 8 + ln + 4 * size
    rec = np.frombuffer(raw[at:at + 36 * n], dtype=np.dtype([("p", ">f4", 4), ("id", ">i4"), ("v", ">f4", 3),
                                                              ("r", ">f4")]))
    assert raw[at + 36 * n:] == b"\x00\xff" and np.all(rec["r"] == np.float32(0.1))
    return rec["p"][:, :3].astype(np.float32), rec["v"].astype(np.float32), rec["id"].astype(np.int32)


@pytest.mark.gpu
def test_lifetime_activation_matches_python_binding(scene, tmp_path, gpu_lib):
    from mpm_amd import Engine, Params
    from mpm_amd.scene_file import load_scene

    out = tmp_path / "out"
    pc, dt, N, steps = 60000, 1e-3, 32, 30
    r = run("--scene", scene, "--particle-count", pc, "--dt", dt, "-N", N, "--steps", steps,
            "--save-dir", out, "--sync-every", 20, "--mesh-grid", 64, "--mesh-particle-radius", 2,
            "--laplacian_smooth", 1, "--mesh-face-count", 500, "--frames-dir", tmp_path / "view")
    assert r.returncode == 0, r.stdout + r.stderr
    frames = particle_frames(out)
    save_every = int(1.0 / 240.0 / np.float32(dt))
    assert len(frames) == (steps + save_every - 1) // save_every
    # a surface mesh per frame (MeshBuilder): OBJ with vertices and faces
    meshes = sorted((out / "meshes").glob("mesh_*.obj"))
    assert len(meshes) == len(frames)
    # smoothed (area 1) and decimated to at most 500 faces
    text = meshes[-1].read_text().splitlines()
    V = np.array([[float(v) for v in l.split()[1:]] for l in text if l.startswith("v ")])
    faces = np.array([[int(v) for v in l.split()[1:]] for l in text if l.startswith("f ")]) - 1
    assert len(V) > 50 and 490 <= len(faces) <= 500 and faces.min() == 0 and faces.max() == len(V) - 1
    area = 0.5 * np.linalg.norm(np.cross(V[faces[:, 1]] - V[faces[:, 0]], V[faces[:, 2]] - V[faces[:, 0]]), axis=1).sum()
    assert 0.5 < area < 1.5  # the smoothed mesh has area 1; decimation changes it a little
    # the debug view every 20 substeps and once before the first (main.cu:76, 99-102)
    views = sorted((tmp_path / "view").glob("frame_*.ppm"))
    assert len(views) == 1 + (steps + 19) // 20
    assert views[0].read_bytes().startswith(b"P6\n800 600\n255\n")
    sc = load_scene(str(scene), pc, seed=1)
    a, b = sc.objects
    # frame k is written after substep i = k * save_every, i.e. after i + 1 substeps
    x0, _, _ = read_bgeo(frames[0])
    assert x0.shape[0] == a.count  # the second object starts at t = 0.01
    xl, vl, _ = read_bgeo(frames[-1])
    assert xl.shape[0] == a.count + b.count
    nsub = (len(frames) - 1) * save_every + 1
    # the same sequence through the Python binding
    t, k_before = 0.0, 0
    while not t >= b.lifetime_begin:
        t += float(np.float32(dt))
        k_before += 1
    eng = Engine(Params(dt=dt, N=N, gravity=-9.81))
    eng.set_materials(sc.materials)
    sl = slice(a.first, a.first + a.count)
    eng.add_particles(sc.x[sl], sc.material[sl], sc.v[sl])
    eng.upload()
    eng.advance(k_before)
    eng.clear_particles()
    sb = slice(b.first, b.first + b.count)
    eng.add_particles(sc.x[sb], sc.material[sb], sc.v[sb])
    eng.append()  # the facade appends a starting object on the device
    eng.advance(nsub - k_before)
    ref = eng.download()
    eng.close()
    assert np.array_equal(xl, ref["x"])
    assert np.array_equal(vl, ref["v"])
