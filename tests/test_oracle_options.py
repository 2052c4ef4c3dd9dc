"""Pins of the oracle's native options, which have no counterpart in the reference:

* atan / atan2 (Cephes atanf) against float64 numpy over every quadrant, both reduction
  branches and the axes;
* the lat-long environment form (u = atan2(z, x)/2pi + 1/2, v = polar angle / pi, bilinear,
  wrapped across the u seam, clamped at the poles): constant fields, fields linear in u or v
  (bilinear reproduces them, which pins the orientation: +x at u = 1/2, +z at 3/4, -z at 1/4,
  +y at the top row), one-hot texels seen from their centres, continuity across the seam, and
  the baked sky against the cube bake of the same sky;
* Russian roulette from a later bounce (nb_render_options): with every roulette number forced
  high a path dies at exactly the first bounce where roulette applies, and before it the
  throughput is left as it is (a mirror under a constant sky returns F L exactly however small
  the contribution factor).
"""
import math

import numpy as np
import pytest

import oracle
from nbounce_amd import scenegen
from tests.test_oracle_multibounce import base_table, closed_box, rays_of
from tests.test_oracle_sampling import _plane_scene


@pytest.fixture(autouse=True)
def _reference_roulette():
    oracle.set_roulette_first(0)
    yield
    oracle.set_roulette_first(0)


def _ulps(got, ref):
    g = np.float32(got)
    r = np.float32(ref)
    ulp = np.spacing(np.abs(r)) if r != 0 else np.float32(1e-45)
    return abs(float(g) - float(ref)) / float(ulp)


# ---- atan / atan2 --------------------------------------------------------------------------

def test_atan_matches_float64_across_both_reductions():
    xs = np.concatenate([np.linspace(-50, 50, 2001), np.geomspace(1e-6, 1e6, 400), -np.geomspace(1e-6, 1e6, 400),
                         [0.4142135, 0.4142136, 2.4142134, 2.4142137, 1.0, -1.0]]).astype(np.float32)
    worst = max(_ulps(oracle.atan(x), np.arctan(np.float64(x))) for x in xs)
    assert worst <= 3.0
    assert oracle.atan(np.float32(np.inf)) == np.float32(math.pi / 2)
    assert oracle.atan(-np.float32(np.inf)) == -np.float32(math.pi / 2)
    assert oracle.atan(0.0) == 0.0


def test_atan2_matches_float64_in_every_quadrant():
    rng = np.random.default_rng(3)
    worst = 0.0
    for _ in range(4000):
        y, x = rng.normal(size=2).astype(np.float32) * np.float32(10.0) ** rng.integers(-3, 4)
        worst = max(worst, _ulps(oracle.atan2(y, x), np.arctan2(np.float64(y), np.float64(x))))
    assert worst <= 4.0
    half_pi, pi = np.float32(math.pi / 2), np.float32(math.pi)
    assert oracle.atan2(1.0, 0.0) == half_pi and oracle.atan2(-1.0, 0.0) == -half_pi
    assert oracle.atan2(0.0, 0.0) == 0.0 and oracle.atan2(0.0, 2.0) == 0.0
    assert oracle.atan2(0.0, -1.0) == pi and oracle.atan2(-0.0, -1.0) == pi    # +pi for both zeros
    assert abs(oracle.atan2(-1e-6, -1.0) + pi) < 1e-5
    # a slip in the quadrant correction (x < 0 without +-pi) is off by pi
    assert abs(oracle.atan2(1.0, -1.0) - np.float32(3 * math.pi / 4)) < 1e-6


# ---- lat-long lookup -------------------------------------------------------------------------

def _dir(u, v):
    """Direction of lookup coordinates (u, v) under the stated convention."""
    phi, theta = (u - 0.5) * 2 * math.pi, v * math.pi
    return np.array([math.sin(theta) * math.cos(phi), math.cos(theta), math.sin(theta) * math.sin(phi)], np.float32)


def test_constant_latlong_is_constant():
    env = oracle.Env.latlong(np.full((8, 16, 4), 3.0, np.float32))
    rng = np.random.default_rng(1)
    for d in rng.normal(size=(500, 3)):
        assert np.all(np.abs(env.sample(d / np.linalg.norm(d)) - 3.0) <= 3.0 * 2e-7)   # weights sum to 1 +- ulps


def test_linear_fields_pin_the_orientation():
    W, H = 32, 16
    col = np.zeros((H, W, 4), np.float32)
    col[..., 0] = np.arange(W)[None, :] + 0.5          # value = u * W at texel centres
    col[..., 1] = np.arange(H)[:, None] + 0.5          # value = v * H
    env = oracle.Env.latlong(col)
    # named directions: +x at u = 1/2, +z at 3/4, -z at 1/4; the horizon at v = 1/2
    for d, u in (((1, 0, 0), 0.5), ((0, 0, 1), 0.75), ((0, 0, -1), 0.25), ((1, 0, 1), 0.625)):
        d = np.array(d, np.float32) / np.linalg.norm(d)
        out = env.sample(d)
        assert abs(out[0] - u * W) < 1e-4 and abs(out[1] - 0.5 * H) < 1e-4
    # 60 degrees from +y is v = 1/3; +y itself clamps to the top row's value
    out = env.sample(_dir(0.4, 1.0 / 3.0))
    assert abs(out[0] - 0.4 * W) < 1e-4 and abs(out[1] - H / 3.0) < 1e-4
    assert abs(env.sample(np.array([0, 1, 0], np.float32))[1] - 0.5) < 1e-6
    assert abs(env.sample(np.array([0, -1, 0], np.float32))[1] - (H - 0.5)) < 1e-6
    rng = np.random.default_rng(2)
    for _ in range(300):
        u, v = rng.uniform(0.05, 0.95), rng.uniform(1.0 / H, 1.0 - 1.0 / H)
        out = env.sample(_dir(u, v))
        assert abs(out[0] - u * W) < 2e-4 and abs(out[1] - v * H) < 2e-4


def test_one_hot_texels_seen_from_their_centres():
    W, H = 16, 8
    dirs = scenegen.latlong_directions(W, H)
    for j, i in ((0, 0), (3, 8), (4, 12), (7, 15), (2, 4)):
        m = np.zeros((H, W, 4), np.float32)
        m[j, i, :3] = 1.0
        env = oracle.Env.latlong(m)
        assert env.sample(dirs[j, i].astype(np.float32))[0] > 0.999
        far = dirs[(j + H // 2) % H, (i + W // 2) % W].astype(np.float32)
        assert env.sample(far)[0] == 0.0


def test_latlong_is_continuous_across_the_seam():
    W, H = 64, 32
    u = (np.arange(W) + 0.5) / W
    m = np.ones((H, W, 4), np.float32)
    m[..., 0] = np.cos(2 * np.pi * u)[None, :]
    env = oracle.Env.latlong(m)
    for v in (0.2, 0.5, 0.8):
        a = env.sample(_dir(1.0 - 1e-5, v))[0]
        b = env.sample(_dir(1e-5, v))[0]
        assert abs(a - b) < 1e-3
        assert abs(a - 1.0) < (2 * np.pi / W) ** 2          # bilinear error of cos at the seam


def test_baked_sky_agrees_with_the_cube_bake():
    ll = oracle.Env.latlong(scenegen.sky_latlong(512, 256))
    cube = oracle.Env(scenegen.sky_cube(256))
    sun = np.array([0.4, 0.8, 0.45]) / np.linalg.norm([0.4, 0.8, 0.45])
    rng = np.random.default_rng(4)
    checked = 0
    for d in rng.normal(size=(3000, 3)):
        d = d / np.linalg.norm(d)
        if abs(d[1]) < 0.05 or d @ sun > 0.9:     # the horizon step and the sharp sun
            continue
        a, b = ll.sample(d.astype(np.float32)), cube.sample(d.astype(np.float32))
        truth = scenegen.sky_gradient(d[None, :])[0]
        assert np.all(np.abs(a - truth) <= 0.02 * np.maximum(truth, 0.2))
        assert np.all(np.abs(a - b) <= 0.03 * np.maximum(truth, 0.2))
        checked += 1
    assert checked > 2000


# ---- Russian roulette from a later bounce ----------------------------------------------------

@pytest.fixture(scope="module")
def mirror_white():
    return closed_box(color=(1.0, 1.0, 1.0))


@pytest.mark.parametrize("bounces", [8, 12])
def test_roulette_first_bounce_sets_the_death_bounce(mirror_white, bounces):
    """White mirror box: the throughput stays exactly 1, so with contribution factor 0.5 the
    continue probability is 0.5 (1 - (b/B)^8) <= 0.5 wherever roulette applies; with every
    roulette number 0.9 the path dies at the first such bounce and makes first + 1 rays (B + 1
    when roulette never applies within the bounce limit).  A `bounce > first` slip gives one
    ray more."""
    shift = oracle.pixel_shift(4, 4)
    t = base_table(4, shift, x=0.5, y=0.5, z=0.9)
    for first in range(0, bounces + 3):
        oracle.set_roulette_first(first)
        _, rays = rays_of(mirror_white, t, 1, bounces, 0.5)
        assert rays == min(first, bounces) + 1, (first, rays)
    oracle.set_roulette_first(0)
    _, rays = rays_of(mirror_white, t, 1, bounces, 0.5)
    assert rays == 1                                  # the reference's rule: dies at bounce 0


def test_throughput_is_unchanged_before_the_first_roulette():
    """A roulette-free first bounce: a mirror (F0 = 0.5) under a constant sky of 2 returns 2F
    exactly even at contribution factor 0.1, where the reference's roulette at bounce 0 either
    kills the path or divides it by p = 0.1 F."""
    sc = _plane_scene(0.0, 1.0, 0.5)
    env = oracle.Env(scenegen.constant_cube(4, 2.0))
    cam = oracle.camera([0.0, 0.0, 10.0], [3.0, 0.0, 0.0], (0, 1, 0), math.pi / 3, 1.0, 0.1)
    table = oracle.sobol_table(4, 17, 0)
    g = oracle.globals_(1, 1.0, 1, 0.1)
    oracle.set_roulette_first(1)
    img1, _ = oracle.render(sc, cam, g, table, 17, env, 16, 16)
    oracle.set_roulette_first(0)
    img0, _ = oracle.render(sc, cam, g, table, 17, env, 16, 16)
    o, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
    differ = 0
    for y in range(16):
        for x in range(16):
            shift = oracle.map4f(oracle.hash4u([x, y, x, y]))
            jit = np.ascontiguousarray((table[1, 0, :2] + shift[:2]) % 1.0, np.float32)
            oracle.lib.orc_generate_ray(oracle.p(cam), x, y, 16, 16, oracle.p(jit), oracle.p(o), oracle.p(d))
            F = 0.5 + 0.5 * (1 - (-d[2] / np.linalg.norm(d))) ** 5
            assert abs(img1[y, x, 0] - 2.0 * F) < 1e-5 * 2.0
            differ += abs(img0[y, x, 0] - 2.0 * F) > 1e-3
    assert differ > 128          # the reference's rule kills most of these paths or reweights them
