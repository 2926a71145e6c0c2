"""mpm_remove_particles (Engine.remove): particles leave the device (an
object's lifetime ends) without the others being copied to the host and back.

Against the oracle on the same sequence: three elastic balls advance together,
the middle one is removed, the other two advance on; the removed particles'
returned state and the remaining particles' state match the FP32 oracle run
on the same split (1e-4, the elastic parity bar), downloads list the
remaining particles in their original order, and the indices close up (a
later append lands behind them). Removing everything, an empty range and a
range past the end are handled.
"""
import numpy as np
import pytest

import oracle
from conftest import state_errors
from mpm_amd import MATERIAL_SNOW, Engine, scenes

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "C", "F", "Jp")


def three_balls():
    a = scenes.rubber_elastic(3000, 71, N=32, velocity=(0.5, -1.0, 0.2))
    b = scenes.small_scene(MATERIAL_SNOW, 2500, 72, N=32, center=(0.5, 0.7, 0.5), radius=0.1,
                           velocity=(0.0, -2.0, 0.0))
    c = scenes.rubber_elastic(2000, 73, N=32, velocity=(-0.3, 0.5, 0.0))
    sc = a
    sc.x = np.concatenate([a.x * 0.6 + 0.1, b.x, c.x * 0.5 + 0.45])
    sc.v = np.concatenate([a.v, b.v, c.v])
    sc.material = np.concatenate([np.zeros(3000), np.ones(2500), np.zeros(2000)]).astype(np.uint8)
    sc.materials = [a.materials[0], b.materials[0]]
    return sc, (0, 3000, 5500, 7500)


def aos_to_state(raw):
    rec = raw.reshape(-1, 104)
    f = lambda off, k: np.frombuffer(rec[:, off:off + 4 * k].tobytes(), np.float32).reshape(-1, k)
    return {"x": f(4, 3), "v": f(16, 3), "F": f(28, 9).reshape(-1, 3, 3).transpose(0, 2, 1),
            "C": f(64, 9).reshape(-1, 3, 3).transpose(0, 2, 1), "Jp": f(100, 1)[:, 0],
            "material": rec[:, 0].copy()}


def test_remove_middle_object_matches_oracle(gpu_lib):
    sc, (a0, b0, c0, end) = three_balls()
    st = sc.state()
    k1, k2 = 12, 18
    eng = Engine(sc.params)
    eng.set_materials(sc.materials)
    eng.add_particles(st["x"], st["material"], st["v"], st["C"], st["F"], st["Jp"])
    eng.upload()
    eng.advance(k1)
    gone = aos_to_state(eng.remove(b0, c0 - b0))
    assert eng.count() == end - (c0 - b0)
    eng.advance(k2)
    out = eng.download()
    eng.close()
    mid = oracle.advance(sc.params, sc.materials, st, k1, np.float32)
    e_gone = state_errors(gone, {f: mid[f][b0:c0] for f in mid}, sc.params.N)
    print("removed particles vs FP32 oracle:", e_gone)
    for f in ("x", "v", "F", "C"):
        assert e_gone[f] < 1e-4, (f, e_gone[f])
    assert np.array_equal(gone["material"], st["material"][b0:c0])
    keep = np.r_[a0:b0, c0:end]
    rest = {f: mid[f][keep] for f in mid}
    ref = oracle.advance(sc.params, sc.materials, rest, k2, np.float32)
    e = state_errors(out, ref, sc.params.N)
    print("remaining particles vs FP32 oracle:", e)
    for f in ("x", "v", "F", "C"):
        assert e[f] < 1e-4, (f, e[f])
    assert np.array_equal(out["material"], st["material"][keep])


def test_remove_then_append_and_edge_cases(gpu_lib):
    sc, (a0, b0, c0, end) = three_balls()
    st = sc.state()
    eng = Engine(sc.params)
    eng.set_materials(sc.materials)
    eng.add_particles(st["x"], st["material"], st["v"], st["C"], st["F"], st["Jp"])
    eng.upload()
    eng.advance(5)
    before = eng.download()
    assert eng.remove(0, 0).size == 0 and eng.count() == end
    with pytest.raises(Exception, match="outside"):
        eng.remove(end - 10, 20)
    eng.remove(0, b0)  # the first object
    after = eng.download()
    for f in FIELDS:
        assert np.array_equal(after[f], before[f][b0:]), f
    # a later object lands behind the remaining ones
    eng.clear_particles()
    eng.add_particles(st["x"][:100] + 0.01, st["material"][:100])
    eng.append()
    assert eng.count() == end - b0 + 100
    eng.advance(3)
    out = eng.download()
    assert np.isfinite(out["x"]).all()
    eng.remove(0, eng.count())  # everything
    assert eng.count() == 0
    eng.advance(2)
    eng.close()
