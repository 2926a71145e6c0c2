"""The debug view on the GPU (mpm_render_frame, Engine.render) against the
oracle's rasteriser (oracle/render.py) bit for bit: random particles, some
outside the clip cube, at the reference's 800 x 600 and at a small odd size,
before and after substeps."""
import numpy as np
import pytest

from mpm_amd import MATERIAL_SNOW, Engine, Params, make_material
from oracle import render as orr

pytestmark = pytest.mark.gpu


def test_render_matches_oracle(gpu_lib):
    rng = np.random.default_rng(11)
    x = rng.uniform(0.1, 0.9, (5000, 3)).astype(np.float32)
    x[:20] = rng.uniform(-0.1, 1.1, (20, 3)).astype(np.float32)   # some outside the clip cube
    e = Engine(Params(dt=1e-4, N=32))
    e.set_materials([make_material(MATERIAL_SNOW, 1e-6, 400.0)])
    e.add_particles(x, np.zeros(len(x), np.uint8))
    e.upload()
    for W, H in ((800, 600), (97, 61)):
        img = e.render(W, H)
        want = orr.render_points(e.download()["x"], W, H)
        assert np.array_equal(img, want), (W, H, int(np.count_nonzero(np.any(img != want, axis=2))))
    e.advance(30)
    img = e.render()
    assert np.array_equal(img, orr.render_points(e.download()["x"]))
    assert np.count_nonzero(np.all(img == orr.POINT, axis=2)) > 1000
    e.close()


def test_render_errors(gpu_lib):
    from mpm_amd import MPMError
    e = Engine(Params(dt=1e-4, N=16))
    e.set_materials([make_material(MATERIAL_SNOW, 1e-6, 400.0)])
    with pytest.raises(MPMError, match="nothing uploaded"):
        e.render(8, 8)
    e.upload()
    assert np.all(e.render(8, 8) == orr.BACK)  # no particles: the clear colour
    with pytest.raises(MPMError, match="width, height"):
        e.render(0, 8)
    e.close()
