"""Host code under AddressSanitizer and UBSan (CPU only).

Builds tests/sanitize/host_driver.cpp with the scene input path
(mpm_amd/csrc/scene.cpp: scene TOML parser, OBJ reader, rejection sampler;
mpm_amd/csrc/material.cpp: material construction) and
the CPU oracle (oracle/oracle_capi.cpp) using g++ -fsanitize=address,undefined,
then loads a scene, samples a mesh, runs oracle substeps on the particles and
feeds malformed scene and mesh files to the loaders. Any out-of-bounds access,
use after free or undefined behaviour aborts the driver.
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CXX = shutil.which("g++")

CUBE = """\
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 1 0 1
v 1 1 1
v 0 1 1
f 1 4 3 2
f 5 6 7 8
f 1 2 6 5
f 2 3 7 6
f 3 4 8 7
f 4 1 5 8
"""

SCENE = """\
[[material]]
name = "snow"
density = 700
E = 1.4e5
Nu = 0.2

[[material]]
name = "water"
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
mesh = "cube.obj"
size = 0.15
position = [0.55, 0.2, 0.5]
velocity = [0, 0, 0]
lifetime_begin = 0.5
"""

# scene files the loader must reject
BAD_SCENES = {
    "unterminated.toml": '[[material]]\nname = "snow\ndensity = 700\n',
    "unknown_material.toml": '[[material]]\nname = "snow"\n[[object]]\nmaterial = "ice"\nmesh = "cube.obj"\n',
    "missing_mesh.toml": ('[[material]]\nname = "snow"\n[[object]]\nmaterial = "snow"\nmesh = "nowhere.obj"\n'
                          'position = [0.1, 0.2, 0.3]\n'),
    "short_position.toml": ('[[material]]\nname = "snow"\n[[object]]\nmaterial = "snow"\nmesh = "cube.obj"\n'
                            'position = [0.1, 0.2]\n'),
    "bad_number.toml": '[[material]]\nname = "snow"\ndensity = 7e\n',
    "open_array.toml": ('[[material]]\nname = "snow"\n[[object]]\nmaterial = "snow"\nmesh = "cube.obj"\n'
                        'position = [0.1, 0.2,\n'),
    "long_line.toml": "[[material]]\nname = \"" + "s" * 100000 + "\n",
    "binary.toml": "\x00\x01\x02[[\xff\xfe",
}

# meshes the sampler may accept or reject; only memory errors fail
ODD_MESHES = {
    "index_past_end.obj": CUBE + "f 1 2 99\n",
    "index_zero.obj": CUBE.replace("f 1 4 3 2", "f 0 4 3 2"),
    "negative_index.obj": CUBE + "f -1 -2 -3\n",
    "far_negative_index.obj": CUBE + "f -100 -2 -3\n",
    "two_vertex_face.obj": CUBE + "f 1 2\n",
    "no_faces.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\n",
    "empty.obj": "",
    "garbage_vertex.obj": "v a b c\n" + CUBE,
    "flat.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n",
    "huge_index.obj": CUBE + "f 1 2 99999999999999999999\n",
}


def _sanitizers_work(tmp_path: Path) -> bool:
    src = tmp_path / "probe.cpp"
    src.write_text("int main() { return 0; }\n")
    r = subprocess.run([CXX, "-fsanitize=address,undefined", str(src), "-o", str(tmp_path / "probe")],
                       capture_output=True, text=True)
    return r.returncode == 0


@pytest.mark.skipif(CXX is None, reason="g++ not found")
def test_scene_input_and_oracle_under_asan_ubsan(tmp_path):
    if not _sanitizers_work(tmp_path):
        pytest.skip("g++ has no sanitizer runtime")
    exe = tmp_path / "host_driver"
    cmd = [CXX, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fopenmp", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           str(ROOT / "tests/sanitize/host_driver.cpp"), str(ROOT / "mpm_amd/csrc/scene.cpp"),
           str(ROOT / "mpm_amd/csrc/material.cpp"), str(ROOT / "oracle/oracle_capi.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]

    (tmp_path / "meshes").mkdir()
    (tmp_path / "meshes/cube.obj").write_text(CUBE)
    (tmp_path / "scene.toml").write_text(SCENE)
    args = [str(exe), str(tmp_path / "scene.toml"), str(tmp_path / "meshes/cube.obj")]
    for name, text in BAD_SCENES.items():
        (tmp_path / name).write_bytes(text.encode("latin-1"))
        args.append("bad:" + str(tmp_path / name))
    for name, text in ODD_MESHES.items():
        (tmp_path / "meshes" / name).write_text(text)
        args.append("obj:" + str(tmp_path / "meshes" / name))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run(args, capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-6000:]
    assert r.returncode == 0, out[-6000:]
    assert "host driver: 0 failures" in r.stdout
    # every malformed scene was reported as such
    assert r.stdout.count("bad:") == len(BAD_SCENES)
