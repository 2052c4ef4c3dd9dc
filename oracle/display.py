"""CPU restatement of the display blit (numpy).

TEST INFRASTRUCTURE ONLY: the checker of nb_display_blit; the product never imports it.

Follows BlitRenderer (src/pathtracing/blit_renderer.rs) and blit.wgsl:27-38: a fullscreen
triangle whose texcoord is ndc * 0.5 + 0.5, `textureSample` of the RGBA32F target with the
target's linear / clamp-to-edge sampler (src/common/texture.rs:64-74) into an RGBA16F surface
(src/common/wgpu_context.rs:62).  Texture coordinate v = 0 is the target's first row in memory,
which the path tracer fills with the bottom of the view (pathtracing.wgsl:27-38), so surface row
0 (the top) samples near v = 1.

WebGPU leaves the filter weight precision to the implementation; this restatement and the kernel
both take the texel-space coordinate to 8 fractional bits, rounded half up (a texture unit's
fixed-point weights), and interpolate in binary32 as a * (1 - w) + b * w, x first.

Pinned by tests/test_oracle_display.py against closed forms that do not use this code's
arithmetic: equal sizes are an exact row flip; a 2x magnification of a linear ramp reproduces
the ramp at the sample points; a 2x minification is the mean of 2x2 texel blocks; a constant
image stays constant; clamp-to-edge repeats the border texels.
"""
import numpy as np


def fixed_coord(p, n_dst, n_src):
    """Texel-space sample coordinate ((p + 0.5) / n_dst * n_src - 0.5) in units of 1/256,
    rounded half up, in exact integer arithmetic."""
    p = np.asarray(p, dtype=np.int64)
    return ((2 * p + 1) * n_src * 256 + n_dst) // (2 * n_dst) - 128


def _lerp(a, b, w):
    w = w.astype(np.float32)
    k = np.float32(1.0) - w
    return a * k + b * w


def blit(target, width, height):
    """(height, width, 4) float16 surface of a (h, w, 4) float32 target in target row order."""
    src = np.ascontiguousarray(target, dtype=np.float32)
    sh, sw = src.shape[:2]
    px = np.arange(width)
    py = np.arange(height)
    fx = fixed_coord(px, width, sw)
    fy = fixed_coord(height - 1 - py, height, sh)
    x0, y0 = fx >> 8, fy >> 8
    wx = (fx & 255).astype(np.float32) * np.float32(1.0 / 256.0)
    wy = (fy & 255).astype(np.float32) * np.float32(1.0 / 256.0)
    xa, xb = np.clip(x0, 0, sw - 1), np.clip(x0 + 1, 0, sw - 1)
    ya, yb = np.clip(y0, 0, sh - 1), np.clip(y0 + 1, 0, sh - 1)
    wx = wx[None, :, None]
    wy = wy[:, None, None]
    t00 = src[ya][:, xa]
    t10 = src[ya][:, xb]
    t01 = src[yb][:, xa]
    t11 = src[yb][:, xb]
    c = _lerp(_lerp(t00, t10, wx), _lerp(t01, t11, wx), wy)
    return c.astype(np.float16)
