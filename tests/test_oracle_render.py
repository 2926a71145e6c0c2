"""The debug view's rasteriser restatement (oracle/render.py) against a brute
force: every pixel whose centre (f + 0.5) lies in (xw - 1, xw + 1] on both
axes, tested in FP64, for points in and out of the clip cube, on pixel
boundaries and at the image edges."""
import numpy as np

from oracle import render as orr


def brute(x, W, H):
    img = np.full((H, W, 3), orr.BACK, np.uint8)
    xw, yw, kept = orr.window_positions(x, W, H)
    fx = np.arange(W) + 0.5
    fy = np.arange(H) + 0.5
    for px, py in zip(xw[kept].astype(np.float64), yw[kept].astype(np.float64)):
        cx = (fx > px - 1.0) & (fx <= px + 1.0)
        cy = (fy > py - 1.0) & (fy <= py + 1.0)
        for y in np.nonzero(cy)[0]:
            img[H - 1 - y, cx] = orr.POINT
    return img


def test_rasteriser_matches_brute_force():
    rng = np.random.default_rng(3)
    W, H = 64, 48
    x = rng.uniform(-0.1, 1.1, (300, 3)).astype(np.float32)
    # points whose window position is a pixel edge or centre, and the clip faces
    edge = np.array([[0.5, 0.5, 0.5], [0.5 + 1 / 1.8, 0.5, 0.5], [0.5 - 1 / 1.8, 0.5, 0.5],
                     [0.5, 0.5, 0.5 + 1 / 1.8], [0.5 + 3 / (1.8 * W), 0.5 + 2 / (1.8 * H), 0.2]], np.float32)
    x = np.concatenate([x, edge])
    assert np.array_equal(orr.render_points(x, W, H), brute(x, W, H))


def test_a_point_covers_two_by_two_pixels():
    W, H = 40, 30
    # centre (0.5, 0.5) -> window (20, 15): pixel centres 19.5, 20.5 and 14.5, 15.5
    img = orr.render_points(np.array([[0.5, 0.5, 0.5]], np.float32), W, H)
    lit = np.argwhere(np.all(img == orr.POINT, axis=2))
    assert sorted(map(tuple, lit.tolist())) == sorted([(H - 1 - 14, 19), (H - 1 - 14, 20), (H - 1 - 15, 19),
                                                       (H - 1 - 15, 20)])
    assert np.count_nonzero(np.all(img == orr.BACK, axis=2)) == W * H - 4
