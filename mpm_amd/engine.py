"""Python handle on one engine instance (wraps the C-ABI of libmpm_hip.so).

Mirrors the reference's class Simulation (include/mpm.cuh:44-106) for Python
callers: materials and particles are staged, `upload()` stands in for
`initCuda()`, `advance(k)` for k calls of `advance()`, `download()` for
`syncDevice()` + `getFullParticleList()`.

Particle arrays use the SoA layout of the C-ABI: x (n,3), v (n,3), C (n,3,3),
F (n,3,3) row-major, Jp (n,), material (n,) uint8, all float32.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, lib, mpm_material, mpm_params


def _fp(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _stream(stream) -> Optional[int]:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream or None
    # torch.cuda.Stream (HIP stream on ROCm)
    return int(stream.cuda_stream) or None


@dataclass
class Params:
    """Simulation parameters (SimulationParameters, include/TransferScheme.h:6-29,
    plus the constants src/mpm.cu:6,92 hard-codes)."""
    dt: float = 1e-4
    N: int = 60
    gravity: float = -9.81
    boundary_mode: int = _lib.BOUNDARY_FRACTION
    boundary: float = 0.05
    compat: int = _lib.COMPAT_REFERENCE

    def to_c(self) -> mpm_params:
        return mpm_params(self.dt, self.N, self.gravity, self.boundary_mode, self.boundary,
                          self.compat)


def make_material(kind: int, volume: float, density: float, E: float = 1.4e5, nu: float = 0.2,
                  hardening: float = 10.0, clamp_lo: float = 0.975, clamp_hi: float = 1.0075,
                  bulk_modulus: float = 0.0) -> mpm_material:
    """Build a material the way the reference's constructors do
    (MaterialModel.cuh:27-29,50-54,75-82) through the library's own routine."""
    m = mpm_material()
    check(lib().mpm_material_init(C.byref(m), kind, volume, density, E, nu, hardening, clamp_lo,
                                  clamp_hi, bulk_modulus))
    return m


def material_as_numpy(kind: int, volume: float, density: float, E: float = 1.4e5,
                      nu: float = 0.2, hardening: float = 10.0, clamp_lo: float = 0.975,
                      clamp_hi: float = 1.0075, bulk_modulus: float = 0.0) -> mpm_material:
    """Same construction as make_material, written in numpy float32 (usable
    without the HIP library; the tests check both agree bit for bit)."""
    f = np.float32
    E32, nu32, v32, d32 = f(E), f(nu), f(volume), f(density)
    m = mpm_material()
    m.kind = kind
    m.particle_volume = float(v32)
    m.particle_mass = float(d32 * v32)
    m.mu0 = float(E32 / (f(2) * (f(1) + nu32)))
    m.lambda0 = float(E32 * nu32 / ((f(1) + nu32) * (f(1) - f(2) * nu32)))
    m.hardening = float(f(hardening))
    m.clamp_lo = float(f(clamp_lo))
    m.clamp_hi = float(f(clamp_hi))
    m.bulk_modulus = float(f(bulk_modulus))
    return m


class Engine:
    """One simulation on the current HIP device."""

    def __init__(self, params: Params):
        self.params = params
        self._h = C.c_void_p()
        self._cp = params.to_c()
        check(lib().mpm_create(C.byref(self._cp), C.byref(self._h)), None, "mpm_create")
        self._n = 0

    # -- lifetime
    def close(self) -> None:
        if self._h:
            lib().mpm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def _chk(self, rc: int, what: str = "") -> None:
        check(rc, self._h, what)

    # -- configuration
    def set_materials(self, mats: Sequence[mpm_material]) -> None:
        arr = (mpm_material * len(mats))(*mats)
        self._chk(lib().mpm_set_materials(self._h, arr, len(mats)))

    def set_option(self, name: str, value: int) -> None:
        self._chk(lib().mpm_set_option(self._h, name.encode(), int(value)))

    def get_option(self, name: str) -> int:
        v = C.c_int64()
        self._chk(lib().mpm_get_option(self._h, name.encode(), C.byref(v)))
        return v.value

    # -- particles
    def add_particles(self, x, material, v=None, C_=None, F=None, Jp=None) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
        n = x.shape[0]
        mat = np.ascontiguousarray(material, dtype=np.uint8).reshape(n)
        arrs = []
        for a, shape in ((v, (n, 3)), (C_, (n, 3, 3)), (F, (n, 3, 3)), (Jp, (n,))):
            arrs.append(None if a is None else np.ascontiguousarray(a, dtype=np.float32).reshape(shape))
        self._chk(lib().mpm_add_particles_soa(self._h, n, _fp(x), _fp(arrs[0]), _fp(arrs[1]),
                                              _fp(arrs[2]), _fp(arrs[3]),
                                              mat.ctypes.data_as(C.POINTER(C.c_uint8))))

    def add_particles_aos104(self, raw: np.ndarray) -> None:
        raw = np.ascontiguousarray(raw, dtype=np.uint8)
        assert raw.size % 104 == 0
        self._chk(lib().mpm_add_particles_aos104(self._h, raw.ctypes.data, raw.size // 104))

    def clear_particles(self) -> None:
        self._chk(lib().mpm_clear_particles(self._h))

    def upload(self, stream=None) -> None:
        self._chk(lib().mpm_upload(self._h, _stream(stream)))
        self._n = self.count()

    def append(self, stream=None) -> None:
        """Add the particles staged since the last upload / append (or
        clear_particles) behind the ones on the device, without copying those
        back; the simulated time goes on (mpm_append_particles)."""
        self._chk(lib().mpm_append_particles(self._h, _stream(stream)))
        self._n = self.count()

    def remove(self, first: int, count: int, stream=None) -> np.ndarray:
        """Remove the device particles with original index in [first, first +
        count) (mpm_remove_particles); returns their last state as 104-byte
        records (uint8, (count * 104,)), in index order."""
        out = np.empty(int(count) * 104, np.uint8)
        self._chk(lib().mpm_remove_particles(self._h, int(first), int(count),
                                             out.ctypes.data if count else None, _stream(stream)))
        self._n = self.count()
        return out

    def mesh(self, grid_side: int = 250, radius: int = 5, stream=None) -> dict:
        """Surface mesh of the device particles (mpm_mesh_build): 'triangles'
        (n, 3, 3) float32 vertices, normals out by the right-hand rule, and
        'edges' (n, 3) uint64, the grid edge each vertex lies on."""
        n = C.c_int64()
        self._chk(lib().mpm_mesh_build(self._h, grid_side, radius, C.byref(n), _stream(stream)))
        tri = np.empty((n.value, 3, 3), np.float32)
        edges = np.empty((n.value, 3), np.uint64)
        self._chk(lib().mpm_mesh_triangles(self._h, _fp(tri), edges.ctypes.data))
        return {"triangles": tri, "edges": edges, "grid_side": grid_side}

    def mesh_distance_field(self, grid_side: int) -> np.ndarray:
        """The distance field of the last mesh_build, (side, side, side)."""
        d = np.empty((grid_side,) * 3, np.float32)
        self._chk(lib().mpm_mesh_distance_field(self._h, grid_side, _fp(d)))
        return d

    def write_mesh_obj(self, path: str) -> None:
        self._chk(lib().mpm_mesh_write_obj(self._h, str(path).encode()))

    def render(self, width: int = 800, height: int = 600, stream=None) -> np.ndarray:
        """The reference's debug view of the device particles (mpm_render_frame):
        (height, width, 3) uint8, rows top to bottom."""
        img = np.empty((height, width, 3), np.uint8)
        self._chk(lib().mpm_render_frame(self._h, width, height, img.ctypes.data, _stream(stream)))
        return img

    def mesh_indexed(self):
        """The last mesh as (V (nv, 3) float64, F (nf, 3) int32): one vertex
        per crossed grid edge (mpm_mesh_indexed)."""
        nv, nf = C.c_int64(), C.c_int64()
        self._chk(lib().mpm_mesh_indexed(self._h, C.byref(nv), C.byref(nf), None, None))
        V = np.empty((nv.value, 3), np.float64)
        F = np.empty((nf.value, 3), np.int32)
        self._chk(lib().mpm_mesh_indexed(self._h, C.byref(nv), C.byref(nf), V.ctypes.data, F.ctypes.data))
        return V, F

    def count(self) -> int:
        v = C.c_int64()
        self._chk(lib().mpm_count(self._h, C.byref(v)))
        return v.value

    # -- stepping
    def advance(self, substeps: int = 1, stream=None) -> None:
        self._chk(lib().mpm_advance(self._h, int(substeps), _stream(stream)))

    def time(self) -> float:
        v = C.c_double()
        self._chk(lib().mpm_time(self._h, C.byref(v)))
        return v.value

    # -- per-kernel timing (option "timing" = 1)
    def timing(self, phase: str):
        """(total milliseconds, launches) recorded for one kernel phase."""
        ms = C.c_double()
        n = C.c_int64()
        self._chk(lib().mpm_timing_read(self._h, phase.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def timing_reset(self) -> None:
        self._chk(lib().mpm_timing_reset(self._h))

    def autotune(self, trial_substeps: int = 64, cache: Optional[str] = None,
                 key: Optional[str] = None, stream=None) -> dict:
        """Pick the sort interval, work-item size and radix digit width by
        timing them on this scene (mpm_autotune; the state is restored). With `cache` (a JSON
        file) and `key`, a stored choice is applied without timing, and a new
        one is stored. Returns the choice (and the trials when timed)."""
        import json
        import os
        if cache and key and os.path.exists(cache):
            with open(cache) as f:
                stored = json.load(f)
            if key in stored:
                choice = stored[key]
                self.set_option("sort_every", int(choice["sort_every"]))
                self.set_option("chunk", int(choice["chunk"]))
                self.set_option("radix_bits", int(choice.get("radix_bits", 8)))
                return choice
        buf = C.create_string_buffer(8192)
        self._chk(lib().mpm_autotune(self._h, int(trial_substeps), buf, len(buf), _stream(stream)))
        choice = json.loads(buf.value.decode())
        if cache and key:
            stored = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    stored = json.load(f)
            stored[key] = {"sort_every": choice["sort_every"], "chunk": choice["chunk"],
                           "radix_bits": choice["radix_bits"]}
            with open(cache, "w") as f:
                json.dump(stored, f, indent=1, sort_keys=True)
        return choice

    # -- results
    def download(self, stream=None) -> dict:
        n = self._n
        out = {
            "x": np.empty((n, 3), np.float32), "v": np.empty((n, 3), np.float32),
            "C": np.empty((n, 3, 3), np.float32), "F": np.empty((n, 3, 3), np.float32),
            "Jp": np.empty((n,), np.float32), "material": np.empty((n,), np.uint8),
        }
        self._chk(lib().mpm_download_soa(self._h, _fp(out["x"]), _fp(out["v"]), _fp(out["C"]),
                                         _fp(out["F"]), _fp(out["Jp"]),
                                         out["material"].ctypes.data_as(C.POINTER(C.c_uint8)),
                                         _stream(stream)))
        return out

    def download_aos104(self, stream=None) -> np.ndarray:
        raw = np.empty(self._n * 104, np.uint8)
        self._chk(lib().mpm_download_aos104(self._h, raw.ctypes.data, _stream(stream)))
        return raw

    def download_grid(self, stream=None) -> np.ndarray:
        N = self.params.N
        g = np.empty((N, N, N, 4), np.float32)
        self._chk(lib().mpm_download_grid(self._h, _fp(g), _stream(stream)))
        return g

    def debug_p2g(self, stream=None) -> np.ndarray:
        N = self.params.N
        g = np.empty((N, N, N, 4), np.float32)
        self._chk(lib().mpm_debug_p2g(self._h, _fp(g), _stream(stream)))
        return g

    def debug_sort(self, stream=None):
        keys = np.empty(self._n, np.uint32)
        perm = np.empty(self._n, np.uint32)
        self._chk(lib().mpm_debug_sort(self._h, keys.ctypes.data_as(C.POINTER(C.c_uint32)),
                                       perm.ctypes.data_as(C.POINTER(C.c_uint32)), _stream(stream)))
        return keys, perm

    def save_state(self, path: str) -> None:
        self._chk(lib().mpm_save_state(self._h, str(path).encode()))

    def load_state(self, path: str) -> None:
        self._chk(lib().mpm_load_state(self._h, str(path).encode()))
        self._n = self.count()


class Batch:
    """Independent scenes advanced together (mpm_batch_*): their substeps are
    recorded into one HIP graph and replayed. The engines stay owned by the
    caller and must outlive the batch."""

    def __init__(self, engines: Sequence["Engine"] = ()):
        self._b = C.c_void_p()
        check(lib().mpm_batch_create(C.byref(self._b)), None, "mpm_batch_create")
        self._engines = []
        for e in engines:
            self.add(e)

    def _chk(self, rc: int, what: str = "") -> None:
        if rc != 0:
            msg = lib().mpm_batch_last_error(self._b)
            raise _lib.MPMError(rc, (msg.decode() if msg else "") or what)

    def add(self, engine: "Engine") -> None:
        self._chk(lib().mpm_batch_add(self._b, engine.handle), "mpm_batch_add")
        self._engines.append(engine)

    def advance(self, substeps: int = 1, stream=None) -> None:
        self._chk(lib().mpm_batch_advance(self._b, int(substeps), _stream(stream)))

    def autotune(self, trial_substeps: int = 64, stream=None) -> dict:
        """Pick the batch's settings by timing candidates through the batch
        itself (mpm_batch_autotune): every scene's own tuned choice, and common
        sort intervals and work-item sizes for all scenes, each from a snapshot
        of every scene that is put back afterwards. Returns the choice and the
        trials (microseconds per batch substep)."""
        import json
        buf = C.create_string_buffer(1 << 16)
        self._chk(lib().mpm_batch_autotune(self._b, int(trial_substeps), buf, len(buf), _stream(stream)))
        return json.loads(buf.value.decode())

    def set_option(self, name: str, value: int) -> None:
        self._chk(lib().mpm_batch_set_option(self._b, name.encode(), int(value)))

    def get_option(self, name: str) -> int:
        v = C.c_int64()
        self._chk(lib().mpm_batch_get_option(self._b, name.encode(), C.byref(v)))
        return v.value

    def close(self) -> None:
        if self._b:
            lib().mpm_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_svd(A: np.ndarray):
    """Batched device 3x3 SVD (the engine's routine). A: (n,3,3) float32."""
    A = np.ascontiguousarray(A, dtype=np.float32).reshape(-1, 3, 3)
    n = A.shape[0]
    U = np.empty_like(A)
    V = np.empty_like(A)
    S = np.empty((n, 3), np.float32)
    check(lib().mpm_debug_svd(n, _fp(A), _fp(U), _fp(S), _fp(V)))
    return U, S, V
