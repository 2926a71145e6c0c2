"""Build recipe of libmpm_hip.so (the engine) for gfx950.

Compiles every translation unit in mpm_amd/csrc with hipcc into objects under
mpm_amd/_build/ and links them into mpm_amd/libmpm_hip.so, in-tree, so the
library travels with the source tree. hipcc cross-compiles for gfx950 without
a GPU present.

Run as ``python -m mpm_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libmpm_hip.so"
CLI = PKG / "bin" / "mpm"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no packed FP32 instructions are generated (plain VALU FP32
# only); -fgpu-flush-denormals-to-zero is not used so FP32 follows IEEE
# denormals like the reference's CPU restatement.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # code object v5: loadable by the HIP runtime PyTorch-ROCm ships (the
    # process's runtime), not only by the toolchain's newer one
    "-mcode-object-version=5",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    "-I", str(PKG.parent / "include"),
]

SOURCES = ["capi.cpp", "material.cpp", "engine.cpp", "batch.cpp", "substep.hip", "p2g_block.hip", "sort.hip",
           "linalg.hip", "adjoint.hip", "scene.cpp", "mesh.hip", "mesh_post.cpp", "render.hip"]


def _compile(src: Path) -> Path:
    out = OBJ / (src.name + ".o")
    lang = ["-x", "hip"]  # every unit goes through the HIP driver (host code uses the HIP runtime)
    cmd = [HIPCC, *COMMON, *lang, "-c", str(src), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if res.stderr.strip():
        # keep warnings visible but drop hipcc's note that the host pass does
        # not know the packed-fp32 target feature
        lines = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l]
        if any(l.strip() for l in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    return out


def build(verbose: bool = False) -> Path:
    """Compile and link libmpm_hip.so; returns its path."""
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found")
    OBJ.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=min(len(srcs), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    # --no-undefined: a declaration without a definition fails the link,
    # not the library's first load
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-mcode-object-version=5", "-shared", "-fPIC",
           "-Wl,--no-undefined", "-o", str(tmp), *map(str, objs), "-lrocprofiler-sdk-roctx", "-lz"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    build_cli()
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


def build_cli() -> Path:
    """The `mpm` program (host C++ over the C-ABI) next to the library."""
    CLI.parent.mkdir(exist_ok=True)
    tmp = CLI.with_suffix(".tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", str(CSRC / "mpm_cli.cpp"), "-o", str(tmp),
           f"-L{LIB.parent}", "-lmpm_hip", "-Wl,-rpath,$ORIGIN/.."]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"CLI build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if res.stderr.strip():
        sys.stderr.write(res.stderr)
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    build(verbose=True)
