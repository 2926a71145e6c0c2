"""pytest configuration: the `gpu` marker, repo root on sys.path, and helpers
shared by the parity tests."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libmpm_hip.so")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_lib():
    """libmpm_hip.so on a machine with a GPU; the test fails (not skips) if the
    library is missing, since the GPU tests exist to exercise it."""
    import mpm_amd
    mpm_amd.lib()
    return mpm_amd


def field_rel_err(a: np.ndarray, b: np.ndarray, floor: float = 0.0) -> float:
    """max |a - b| / max(max |b|, floor) over one field (norm-relative error)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if b.size == 0:
        return 0.0
    scale = max(float(np.max(np.abs(b))), floor)
    diff = float(np.max(np.abs(a - b)))
    return diff / scale if scale > 0 else diff


def state_errors(a: dict, b: dict, N: int, keys=("x", "v", "F", "C", "Jp")) -> dict:
    """Per-field errors of state a against reference b. C (a velocity
    gradient) is measured against its physical scale max(|C|, |v| / dx):
    in rigid motion C is roundoff-sized and a bare relative error would be
    meaningless."""
    vmax = float(np.max(np.abs(np.asarray(b["v"], np.float64)))) if b["v"].size else 0.0
    floors = {"x": 0.0, "v": 0.0, "F": 0.0, "Jp": 0.0, "C": vmax * N}
    return {k: field_rel_err(a[k], b[k], floors[k]) for k in keys}


def elem_rel_err(a: np.ndarray, b: np.ndarray, floor: float = 0.0) -> float:
    """max over elements of |a_i - b_i| / max(|b_i|, 1e-3 s), s = max(max |b|,
    floor): the element-wise relative error, with elements far below the
    field's scale (exact zeros of F's off-diagonals, say) measured against a
    thousandth of it."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if b.size == 0:
        return 0.0
    scale = max(float(np.max(np.abs(b))), floor)
    if scale == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3 * scale)))


def state_elem_errors(a: dict, b: dict, N: int, keys=("x", "v", "F", "C", "Jp")) -> dict:
    """Element-wise relative errors per field (C with the floor of
    state_errors)."""
    vmax = float(np.max(np.abs(np.asarray(b["v"], np.float64)))) if b["v"].size else 0.0
    floors = {"x": 0.0, "v": 0.0, "F": 0.0, "Jp": 0.0, "C": vmax * N}
    return {k: elem_rel_err(a[k], b[k], floors[k]) for k in keys}
