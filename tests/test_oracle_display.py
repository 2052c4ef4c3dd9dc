"""Pins the display-blit restatement (oracle/display.py) with closed forms that do not use its
arithmetic: row flip at equal sizes, exact linear reproduction under 2x magnification, 2x2 block
means under 2x minification, constant images, clamp-to-edge borders."""
import numpy as np

from oracle import display


def _image(h, w, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(-4.0, 4.0, size=(h, w, 4)).astype(np.float32)


def test_equal_size_is_a_row_flip():
    t = _image(24, 40)
    got = display.blit(t, 40, 24)
    np.testing.assert_array_equal(got, t[::-1].astype(np.float16))


def test_fixed_point_coordinate_closed_forms():
    # equal size: texel centres; 2x magnification: p/2 - 1/4; 2x minification: 2p + 1/2
    p = np.arange(50)
    assert np.array_equal(display.fixed_coord(p, 50, 50), 256 * p)
    assert np.array_equal(display.fixed_coord(p, 100, 50), 128 * p - 64)
    assert np.array_equal(display.fixed_coord(p[:25], 25, 50), 512 * p[:25] + 128)
    # 3x magnification: (p + 0.5)/3 - 0.5 rounded to 1/256, half up
    want = np.floor(((p + 0.5) / 3.0 - 0.5) * 256.0 + 0.5).astype(np.int64)
    assert np.array_equal(display.fixed_coord(p, 150, 50), want)


def test_magnification_reproduces_a_linear_ramp():
    h, w = 6, 10
    xs = np.arange(w, dtype=np.float32)
    ys = np.arange(h, dtype=np.float32)
    t = np.zeros((h, w, 4), np.float32)
    t[..., 0] = 1.0 + 0.5 * xs[None, :]           # ramp in x
    t[..., 1] = 2.0 - 0.25 * ys[:, None]          # ramp in y (target rows)
    t[..., 2] = 0.5 * xs[None, :] + 0.25 * ys[:, None]
    t[..., 3] = 1.0
    got = display.blit(t, 2 * w, 2 * h).astype(np.float64)
    px = np.arange(2 * w)
    x = np.clip(px / 2.0 - 0.25, 0, w - 1)        # clamp-to-edge holds the border texel
    py = np.arange(2 * h)
    y = np.clip((2 * h - 1 - py) / 2.0 - 0.25, 0, h - 1)   # surface row 0 samples the last target row
    np.testing.assert_array_equal(got[..., 0], np.broadcast_to(1.0 + 0.5 * x, (2 * h, 2 * w)))
    np.testing.assert_array_equal(got[..., 1], np.broadcast_to((2.0 - 0.25 * y)[:, None], (2 * h, 2 * w)))
    np.testing.assert_array_equal(got[..., 2], 0.5 * x[None, :] + 0.25 * y[:, None])
    assert np.all(got[..., 3] == 1.0)


def test_minification_is_the_block_mean():
    h, w = 8, 12
    rng = np.random.default_rng(3)
    t = (rng.integers(-64, 64, size=(h, w, 4)) / 4.0).astype(np.float32)   # quarters: every mean exact
    got = display.blit(t, w // 2, h // 2).astype(np.float64)
    flipped = t[::-1].astype(np.float64)
    want = flipped.reshape(h // 2, 2, w // 2, 2, 4).mean(axis=(1, 3))
    np.testing.assert_array_equal(got, want)


def test_constant_image_stays_constant():
    t = np.full((16, 16, 4), 1.5, np.float32)
    for (w, h) in [(7, 5), (16, 16), (33, 47), (100, 3)]:
        assert np.all(display.blit(t, w, h) == np.float16(1.5))


def test_clamp_to_edge_borders():
    t = _image(4, 4, seed=5)
    got = display.blit(t, 64, 64)
    # the outermost surface pixels sample within half a texel of the border: the corner texel
    np.testing.assert_array_equal(got[0, 0], t[3, 0].astype(np.float16))
    np.testing.assert_array_equal(got[-1, -1], t[0, 3].astype(np.float16))
