"""The autotuner (mpm_autotune through Engine.autotune): it leaves the scene as
it found it (particles bit for bit, simulated time), sets the options it
reports, answers from its JSON cache on a second call, and the scene then
advances to the same result as without tuning, within the step's parity
tolerance (the options change only the FP32 summation order)."""
import json

import numpy as np
import pytest

from conftest import state_errors
from mpm_amd import Engine, scenes

pytestmark = pytest.mark.gpu


def make(sc):
    st = sc.state()
    eng = Engine(sc.params)
    eng.set_materials(sc.materials)
    eng.add_particles(st["x"], st["material"], st["v"], st["C"], st["F"], st["Jp"])
    eng.upload()
    return eng


def test_autotune_restores_state_and_sets_options(gpu_lib, tmp_path):
    sc = scenes.rubber_elastic(20000, 71, N=48, velocity=(0.5, -1.0, 0.0))
    eng = make(sc)
    eng.advance(20)
    before = eng.download()
    t = eng.time()
    cache = str(tmp_path / "tune.json")
    choice = eng.autotune(trial_substeps=32, cache=cache, key="rubber-20k-N48")
    after = eng.download()
    assert eng.time() == t
    for f in ("x", "v", "C", "F", "Jp", "material"):
        assert np.array_equal(before[f], after[f]), f
    # 8 intervals x 2 item sizes with 8-bit digits, then 6- and 4-bit digits at the best of those
    assert len(choice["trials"]) == 18
    grid, widths = choice["trials"][:16], choice["trials"][16:]
    assert all(r["radix_bits"] == 8 for r in grid)
    first = min(grid, key=lambda r: r["us_per_substep"])
    assert [(r["sort_every"], r["chunk"], r["radix_bits"]) for r in widths] == [
        (first["sort_every"], first["chunk"], 6), (first["sort_every"], first["chunk"], 4)]
    assert choice["sort_every"] in (8, 12, 16, 24, 32, 48, 64, 96) and choice["chunk"] in (128, 192)
    best = min(choice["trials"], key=lambda r: r["us_per_substep"])
    assert (best["sort_every"], best["chunk"], best["radix_bits"]) == (
        choice["sort_every"], choice["chunk"], choice["radix_bits"])
    assert eng.get_option("sort_every") == choice["sort_every"]
    assert eng.get_option("chunk") == choice["chunk"]
    assert eng.get_option("radix_bits") == choice["radix_bits"]
    stored = json.load(open(cache))
    assert stored["rubber-20k-N48"] == {"sort_every": choice["sort_every"], "chunk": choice["chunk"],
                                        "radix_bits": choice["radix_bits"]}
    # the same scene again: the cached choice, no trials
    other = make(sc)
    again = other.autotune(cache=cache, key="rubber-20k-N48")
    assert "trials" not in again and other.get_option("sort_every") == choice["sort_every"]
    assert other.get_option("radix_bits") == choice["radix_bits"]
    other.close()
    # advancing after tuning matches an untuned run within the parity tolerance
    eng.advance(30)
    tuned = eng.download()
    eng.close()
    ref = make(sc)
    ref.advance(50)
    plain = ref.download()
    ref.close()
    errs = state_errors(tuned, plain, sc.params.N, keys=("x", "v", "F", "C"))
    assert all(e < 1e-4 for e in errs.values()), errs
