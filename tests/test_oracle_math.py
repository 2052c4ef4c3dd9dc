"""Pins of the oracle's scalar pieces: PCG4D, float mapping, sin/cos, mat4 inverse, camera,
Möller–Trumbore and the slab test, each against closed forms or independent mathematics."""
import math

import numpy as np
import pytest

import oracle

M32 = 0xFFFFFFFF


# ---- PCG4D (common.wgsl:27-41) ------------------------------------------------------------

def _pcg4d_inverse(h):
    """Undo hash4u step by step: the reference's updates are sequential (each uses the values
    just written), so only that order inverts; a parallel-update variant fails this round trip."""
    x, y, z, w = (int(v) for v in h)
    w = (w - y * z) & M32; z = (z - x * y) & M32; y = (y - z * x) & M32; x = (x - y * w) & M32
    x ^= x >> 16; y ^= y >> 16; z ^= z >> 16; w ^= w >> 16
    w = (w - y * z) & M32; z = (z - x * y) & M32; y = (y - z * x) & M32; x = (x - y * w) & M32
    inv_a = pow(1664525, -1, 1 << 32)
    return [((v - 1013904223) * inv_a) & M32 for v in (x, y, z, w)]


def test_pcg4d_inverts_to_its_input():
    rng = np.random.default_rng(1)
    for s in rng.integers(0, 1 << 32, size=(200, 4), dtype=np.uint64):
        assert _pcg4d_inverse(oracle.hash4u(s.astype(np.uint32))) == [int(v) for v in s]
    for x in range(4):
        for y in range(4):
            assert _pcg4d_inverse(oracle.hash4u([x, y, x, y])) == [x, y, x, y]


def test_pcg4d_avalanche():
    rng = np.random.default_rng(2)
    flips = []
    for s in rng.integers(0, 1 << 32, size=(64, 4), dtype=np.uint64).astype(np.uint32):
        a = oracle.hash4u(s)
        for bit in range(0, 32, 5):
            t = s.copy()
            t[0] ^= np.uint32(1 << bit)
            b = oracle.hash4u(t)
            flips.append(sum(bin(int(p) ^ int(q)).count("1") for p, q in zip(a, b)) / 128.0)
    assert 0.4 < np.mean(flips) < 0.6


def test_map4f_closed_forms():
    out = oracle.map4f([0, 0xFFFFFFFF, 1 << 31, 1 << 9])
    assert out[0] == 0.0
    assert out[1] == np.float32(1.0 - 2.0 ** -23)
    assert out[2] == 0.5
    assert out[3] == np.float32(2.0 ** -23)


# ---- sin / cos --------------------------------------------------------------------------

def test_sin_cos_against_float64():
    xs = np.linspace(-math.pi, 2 * math.pi, 20001, dtype=np.float32)
    s = np.array([oracle.lib.orc_sin(float(x)) for x in xs], dtype=np.float32)
    c = np.array([oracle.lib.orc_cos(float(x)) for x in xs], dtype=np.float32)
    ulp = np.float32(2.0 ** -23)
    assert np.max(np.abs(s - np.sin(xs.astype(np.float64)))) <= 2 * ulp
    assert np.max(np.abs(c - np.cos(xs.astype(np.float64)))) <= 2 * ulp
    assert oracle.lib.orc_sin(0.0) == 0.0 and oracle.lib.orc_cos(0.0) == 1.0


# ---- matrices and camera ----------------------------------------------------------------

def test_mat4_inverse_against_numpy():
    rng = np.random.default_rng(3)
    for _ in range(100):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        m = np.eye(4)
        m[:3, :3] = r * rng.uniform(0.3, 3.0)
        m[:3, 3] = rng.uniform(-100, 100, size=3)
        cols = m.T.reshape(16).astype(np.float32)
        inv = oracle.mat4_inverse(cols).reshape(4, 4).T
        ref = np.linalg.inv(m.astype(np.float32).astype(np.float64))
        assert np.allclose(inv, ref, rtol=1e-5, atol=1e-4)
        assert np.allclose(inv @ m, np.eye(4), atol=1e-4)


def test_camera_projects_as_a_perspective_camera():
    eye, target = np.array([3.0, 2.0, 5.0]), np.array([0.5, 1.0, -1.0])
    fov, aspect, near = math.pi / 3, 16 / 9, 0.1
    cam = oracle.camera(eye, target, (0, 1, 0), fov, aspect, near)
    w2c = cam["world_to_clip"][0].reshape(4, 4).T.astype(np.float64)
    c2w = cam["clip_to_world"][0].reshape(4, 4).T.astype(np.float64)

    def ndc(p):
        c = w2c @ np.append(p, 1.0)
        return c[:3] / c[3]

    f = (target - eye) / np.linalg.norm(target - eye)
    s = np.cross(f, [0, 1, 0]); s /= np.linalg.norm(s)
    u = np.cross(s, f)
    assert np.allclose(ndc(target)[:2], 0.0, atol=1e-5)
    d = 7.0
    top = eye + f * d + u * d * math.tan(fov / 2)
    right = eye + f * d + s * d * math.tan(fov / 2) * aspect
    assert np.allclose(ndc(top)[:2], [0.0, 1.0], atol=1e-4)
    assert np.allclose(ndc(right)[:2], [1.0, 0.0], atol=1e-4)
    assert abs(ndc(eye + f * near)[2]) < 1e-5                 # infinite far plane: near -> 0
    assert abs(ndc(eye + f * 1e6)[2] - 1.0) < 1e-5            # far -> 1
    assert np.allclose(c2w @ w2c, np.eye(4), atol=1e-4)


def test_camera_ray_is_unnormalized_from_the_near_plane():
    """generate_ray (pathtracing.wgsl:27-41): origin = clip_to_world column 3 (the near-plane
    centre), direction = origin - unproject(-uv, -1, 1).  With ndc z = 1 + near / z_view, clip
    z = -1 unprojects to z_view = -near/2, so the centre pixel's direction is the view axis with
    length near/2."""
    eye, target = np.array([0.0, 0.0, 5.0]), np.array([0.0, 0.0, 0.0])
    cam = oracle.camera(eye, target, (0, 1, 0), math.pi / 3, 1.0, 0.1)
    o, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
    jit = np.array([0.0, 0.0], np.float32)
    oracle.lib.orc_generate_ray(oracle.p(cam), 4, 4, 8, 8, oracle.p(jit), oracle.p(o), oracle.p(d))
    assert np.allclose(o, [0, 0, 4.9], atol=1e-5)
    assert np.allclose(d / np.linalg.norm(d), [0, 0, -1], atol=1e-6)
    assert abs(np.linalg.norm(d) - 0.05) < 1e-6


# ---- intersection primitives --------------------------------------------------------------

V0, V1, V2 = np.array([0, 0, 0]), np.array([1, 0, 0]), np.array([0, 1, 0])   # CCW seen from +z


def test_triangle_front_face_hit():
    t, u, v = oracle.intersect_triangle([0.25, 0.25, 2.0], [0, 0, -1], V0, V1, V2)
    assert (t, u, v) == (2.0, 0.25, 0.25)


def test_triangle_back_face_culled_and_bias():
    nohit = oracle.NO_HIT
    assert oracle.intersect_triangle([0.25, 0.25, -2.0], [0, 0, 1], V0, V1, V2)[0] == nohit  # back face
    assert oracle.intersect_triangle([0.25, 0.25, 0.05], [0, 0, -1], V0, V1, V2)[0] == nohit  # t < 0.1
    t = oracle.intersect_triangle([0.25, 0.25, 0.2], [0, 0, -2], V0, V1, V2)[0]              # t in |d| units
    assert t == np.float32(0.1)
    assert oracle.intersect_triangle([0.25, 0.25, 1.0], [1, 0, 0], V0, V1, V2)[0] == nohit   # parallel


def test_triangle_outside_edges():
    nohit = oracle.NO_HIT
    for x, y in [(-0.1, 0.5), (0.5, -0.1), (0.6, 0.6), (1.1, 0.0)]:
        assert oracle.intersect_triangle([x, y, 1.0], [0, 0, -1], V0, V1, V2)[0] == nohit
    assert oracle.intersect_triangle([0.5, 0.5, 1.0], [0, 0, -1], V0, V1, V2)[0] == 1.0   # on the edge u+v=1


def test_slab_closed_forms():
    bmin, bmax = [1, 1, 1], [2, 2, 2]
    inv = lambda d: np.float32(1.0) / np.asarray(d, np.float32)  # noqa: E731
    assert oracle.intersect_aabb([0, 1.5, 1.5], inv([1, 1e-30, 1e-30]), bmin, bmax) == 1.0
    assert oracle.intersect_aabb([1.5, 1.5, 1.5], inv([1, 2, 3]), bmin, bmax) < 0   # inside: negative near
    assert oracle.intersect_aabb([3, 1.5, 1.5], inv([1, 0.1, 0.1]), bmin, bmax) == oracle.NO_HIT  # behind
    with np.errstate(divide="ignore"):
        assert oracle.intersect_aabb([0, 1.5, 1.5], inv([1, 0, 0]), bmin, bmax) == 1.0       # inf slopes
        assert oracle.intersect_aabb([0, 3.0, 1.5], inv([1, 0, 0]), bmin, bmax) == oracle.NO_HIT
