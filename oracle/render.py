"""The debug view's rasterisation -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the point drawing of the reference's Renderer
(src/renderer.cpp:80-107, shaders/cube_vertex.glsl: clip position
(x - 0.5) * 1.8, points of size 2; shaders/constant_fragment.glsl: one
colour; a 0.1 grey clear colour) under the pixel rule the library states: a
point outside the clip cube is dropped, its window position is
xw = (xn + 1) * (W / 2), yw = (yn + 1) * (H / 2) in FP32, and it covers the
pixels whose centres lie in (xw - 1, xw + 1] x (yw - 1, yw + 1]. Colours are
round(255 c); rows run top to bottom. Pinned by tests/test_oracle_render.py
against a brute-force test of every pixel centre in FP64.
"""
from __future__ import annotations

import numpy as np

BACK = 26                      # round(255 * 0.1)
POINT = (150, 138, 51)         # round(255 * (0.59, 0.54, 0.2))


def window_positions(x: np.ndarray, W: int, H: int):
    """(xw, yw, kept): FP32 window positions of the points inside the clip cube."""
    x = np.asarray(x, np.float32).reshape(-1, 3)
    c = (x - np.float32(0.5)) * np.float32(1.8)
    kept = np.all((c >= np.float32(-1.0)) & (c <= np.float32(1.0)), axis=1)
    xw = (c[:, 0] + np.float32(1.0)) * np.float32(0.5 * W)
    yw = (c[:, 1] + np.float32(1.0)) * np.float32(0.5 * H)
    return xw, yw, kept


def render_points(x: np.ndarray, W: int = 800, H: int = 600) -> np.ndarray:
    """(H, W, 3) uint8 image."""
    img = np.full((H, W, 3), BACK, np.uint8)
    xw, yw, kept = window_positions(x, W, H)
    for px, py in zip(xw[kept], yw[kept]):
        x0, x1 = int(np.floor(px - np.float32(1.5))) + 1, int(np.floor(px + np.float32(0.5)))
        y0, y1 = int(np.floor(py - np.float32(1.5))) + 1, int(np.floor(py + np.float32(0.5)))
        for y in range(max(y0, 0), min(y1, H - 1) + 1):
            for xx in range(max(x0, 0), min(x1, W - 1) + 1):
                img[H - 1 - y, xx] = POINT
    return img
