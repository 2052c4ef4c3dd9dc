"""Seeded synthetic scenes and environment maps for the benchmark and parity workloads.

Two workloads:
  * `spheres_in_box(seed)`: 64 icospheres (subdivision 3, 1280 triangles each) of radius
    U[0.2, 1.0] with centres in [0, 10]^3, inside an open box [0, 10]^3 (no +Y face, two
    triangles per face, wound inward).  Base colour U[0,1]^3, roughness U[0.05, 1],
    metallic Bernoulli(0.3); 8 spheres emissive with grey radiance U[1, 10].
  * `instanced_meshes(seed)`: 16 base meshes (noise-displaced geodesic spheres and tori,
    2k-50k triangles), 4096 instances with uniform random rotation, scale U[0.3, 3] and
    translation U[0, 200]^3, per-instance materials drawn as above (1 in 8 emissive).

Draws come from numpy.random.default_rng(seed) in the order written in each function.
Triangles are wound counter-clockwise seen from the side their vertex normals point to
(back-face culling in the tracer keeps only that side).  Tangents are unit vectors
perpendicular to the normal with w = 1.
"""
import math

import numpy as np

from . import layout

# ---------------------------------------------------------------------------------------------
# meshes


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def geodesic_sphere(freq):
    """Unit geodesic sphere: every icosahedron face split into freq^2 triangles (20*freq^2)."""
    base_v, base_f = _icosahedron()
    pts, tris = [], []
    offset = 0
    for a, b, c in base_f:
        A, B, C = base_v[a], base_v[b], base_v[c]
        idx = {}
        local = []
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                idx[(i, j)] = len(local)
                local.append(A + (B - A) * (i / freq) + (C - A) * (j / freq))
        for i in range(freq):
            for j in range(freq - i):
                tris.append([offset + idx[(i, j)], offset + idx[(i + 1, j)], offset + idx[(i, j + 1)]])
                if i + j + 1 < freq:
                    tris.append([offset + idx[(i + 1, j)], offset + idx[(i + 1, j + 1)], offset + idx[(i, j + 1)]])
        pts.extend(local)
        offset += len(local)
    p = np.array(pts)
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    key = np.round(p * 1e6).astype(np.int64)
    _, first, inverse = np.unique(key, axis=0, return_index=True, return_inverse=True)
    return p[first], inverse.reshape(-1)[np.array(tris)]


def icosphere(subdivisions):
    """Loop-style icosphere: 4^s * 20 triangles, (s=3: 642 vertices, 1280 triangles)."""
    v, f = _icosahedron()
    verts = [x for x in v]
    faces = f.tolist()
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        faces = nf
    return np.array(verts), np.array(faces, dtype=np.int64)


def torus(n_major, n_minor, r_major=1.0, r_minor=0.35):
    """Torus around +Y with n_major x n_minor quads (2 * n_major * n_minor triangles)."""
    u = np.arange(n_major) * (2 * math.pi / n_major)
    w = np.arange(n_minor) * (2 * math.pi / n_minor)
    uu, ww = np.meshgrid(u, w, indexing="ij")
    x = (r_major + r_minor * np.cos(ww)) * np.cos(uu)
    z = (r_major + r_minor * np.cos(ww)) * np.sin(uu)
    y = r_minor * np.sin(ww)
    p = np.stack([x, y, z], -1).reshape(-1, 3)
    i = np.arange(n_major)[:, None]
    j = np.arange(n_minor)[None, :]
    a = i * n_minor + j
    b = ((i + 1) % n_major) * n_minor + j
    c = ((i + 1) % n_major) * n_minor + (j + 1) % n_minor
    d = i * n_minor + (j + 1) % n_minor
    f = np.concatenate([np.stack([a, d, b], -1).reshape(-1, 3), np.stack([b, d, c], -1).reshape(-1, 3)])
    return p, f


def vertex_normals(p, f):
    n = np.zeros_like(p)
    fn = np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]])
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    return n / np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)


def orient_outward(p, f, n):
    """Flip triangles whose counter-clockwise normal disagrees with the vertex normals."""
    fn = np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]])
    flip = np.einsum("ij,ij->i", fn, n[f].sum(axis=1)) < 0
    f = f.copy()
    f[flip] = f[flip][:, [0, 2, 1]]
    return f


def tangents(n):
    a = np.where(np.abs(n[:, 1:2]) < 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(a, n)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    return np.concatenate([t, np.ones((len(n), 1))], axis=1)


def make_vertices(p, n, uv=None):
    v = np.zeros(len(p), dtype=layout.VERTEX)
    v["position"] = (p.astype(np.float32) + np.float32(0.0))   # + 0 turns -0 into +0
    v["normal"] = n
    v["tangent"] = tangents(n)
    if uv is not None:
        v["u"], v["v"] = uv[:, 0], uv[:, 1]
    return v


def displace(p, rng, amp):
    """Smooth radial noise: sum of 4 random plane waves, amplitude `amp`."""
    k = rng.normal(size=(4, 3)) * 3.0
    ph = rng.uniform(0, 2 * math.pi, size=4)
    s = np.sin(p @ k.T + ph).sum(axis=1) / 4.0
    return p * (1.0 + amp * s)[:, None]


# ---------------------------------------------------------------------------------------------
# transforms


def translate(t):
    m = np.eye(4)
    m[:3, 3] = t
    return m


def scale(s):
    return np.diag([s, s, s, 1.0])


def quat_matrix(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y), 0],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x), 0],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y), 0],
                     [0, 0, 0, 1]])


def random_rotation(rng):
    """Uniform rotation (Shoemake 1992, from three uniforms)."""
    u1, u2, u3 = rng.uniform(size=3)
    q = (math.sqrt(u1) * math.cos(2 * math.pi * u3), math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
         math.sqrt(1 - u1) * math.cos(2 * math.pi * u2), math.sqrt(u1) * math.sin(2 * math.pi * u3))
    return quat_matrix(q)


def _prim(m, color, roughness, metallic, emissive, start, end):
    p = np.zeros(1, dtype=layout.PRIMITIVE)
    p["local_to_world"] = layout.mat4_cols(m)
    p["color"] = color
    p["roughness"], p["metallic"], p["emissive"] = roughness, metallic, emissive
    p["index_start"], p["index_end"] = start, end
    return p


def _material(rng, emissive):
    """Base colour, roughness, metallic; emissive primitives take their radiance as colour
    (scene.rs:191-197: colour = emissive factor, alpha 1)."""
    base = rng.uniform(0.0, 1.0, size=3)
    rough = rng.uniform(0.05, 1.0)
    metal = float(rng.uniform() < 0.3)
    if emissive:
        L = rng.uniform(1.0, 10.0)
        return (L, L, L, 1.0), rough, metal, 1.0
    return (base[0], base[1], base[2], 1.0), rough, metal, 0.0


# ---------------------------------------------------------------------------------------------
# workloads


def spheres_in_box(seed=0, n_spheres=64, n_emissive=8, subdivisions=3, box=10.0):
    """Small scene: icospheres inside an open box; returns (vertices, indices, primitives)."""
    rng = np.random.default_rng(seed)
    sp, sf = icosphere(subdivisions)
    sf = orient_outward(sp, sf, sp)
    verts = [make_vertices(sp, sp)]
    idx = [sf.reshape(-1)]
    sphere_range = (0, len(idx[0]))
    # open box [0, box]^3 without the +Y face; inward-facing normals
    L = box
    faces = {  # corners in counter-clockwise order seen from inside, inward normal
        "-x": ([[0, 0, 0], [0, 0, L], [0, L, L], [0, L, 0]], [1, 0, 0]),
        "+x": ([[L, 0, 0], [L, L, 0], [L, L, L], [L, 0, L]], [-1, 0, 0]),
        "-y": ([[0, 0, 0], [L, 0, 0], [L, 0, L], [0, 0, L]], [0, 1, 0]),
        "-z": ([[0, 0, 0], [0, L, 0], [L, L, 0], [L, 0, 0]], [0, 0, 1]),
        "+z": ([[0, 0, L], [L, 0, L], [L, L, L], [0, L, L]], [0, 0, -1]),
    }
    bp, bn, bf = [], [], []
    for corners, normal in faces.values():
        c = np.array(corners, dtype=np.float64)
        k = sum(len(x) for x in bp)
        bp.append(c)
        bn.append(np.tile(np.array(normal, dtype=np.float64), (4, 1)))
        bf.append(np.array([[0, 1, 2], [0, 2, 3]]) + k)
    bp, bn, bf = np.concatenate(bp), np.concatenate(bn), np.concatenate(bf)
    bf = orient_outward(bp, bf, bn)
    base_v = len(verts[0])
    verts.append(make_vertices(bp, bn))
    idx.append((bf + base_v).reshape(-1))
    box_range = (sphere_range[1], sphere_range[1] + bf.size)

    radii = rng.uniform(0.2, 1.0, size=n_spheres)
    centres = rng.uniform(0.0, box, size=(n_spheres, 3))
    emissive = np.zeros(n_spheres, dtype=bool)
    emissive[rng.choice(n_spheres, size=n_emissive, replace=False)] = True
    prims = [_prim(np.eye(4), (0.8, 0.8, 0.8, 1.0), 0.6, 0.0, 0.0, *box_range)]
    for i in range(n_spheres):
        color, rough, metal, emis = _material(rng, emissive[i])
        m = translate(centres[i]) @ scale(radii[i])
        prims.append(_prim(m, color, rough, metal, emis, *sphere_range))
    return np.concatenate(verts), np.concatenate(idx).astype(np.uint32), np.concatenate(prims)


def base_meshes(seed=0, n_meshes=16):
    """16 noise-displaced geodesic spheres (even slots) and tori (odd slots), 2k..50k triangles."""
    rng = np.random.default_rng(seed + 1000)
    sphere_freq = [10, 13, 16, 20, 25, 30, 38, 50]
    torus_grid = [(40, 25), (50, 40), (64, 50), (80, 64), (100, 80), (128, 96), (160, 120), (200, 125)]
    meshes = []
    for k in range(n_meshes):
        if k % 2 == 0:
            p, f = geodesic_sphere(sphere_freq[k // 2])
            p = displace(p, rng, 0.15)
        else:
            p, f = torus(*torus_grid[k // 2])
            p = displace(p, rng, 0.08)
        n = vertex_normals(p, f)
        f = orient_outward(p, f, n)
        meshes.append((p, n, f))
    return meshes


def instanced_meshes(seed=0, n_instances=4096, extent=200.0, emissive_every=8):
    """Large two-level scene; returns (vertices, indices, primitives).

    Instance k uses mesh m with probability proportional to 1 / triangles(m), which keeps the
    instanced total near 20M triangles.
    """
    meshes = base_meshes(seed)
    verts, idx, ranges = [], [], []
    vbase = ibase = 0
    for p, n, f in meshes:
        verts.append(make_vertices(p, n))
        idx.append((f + vbase).reshape(-1))
        ranges.append((ibase, ibase + f.size))
        vbase += len(p)
        ibase += f.size
    tris = np.array([(e - s) // 3 for s, e in ranges], dtype=np.float64)
    prob = (1.0 / tris) / (1.0 / tris).sum()
    rng = np.random.default_rng(seed)
    choice = rng.choice(len(meshes), size=n_instances, p=prob)
    prims = []
    for k in range(n_instances):
        rot = random_rotation(rng)
        s = rng.uniform(0.3, 3.0)
        t = rng.uniform(0.0, extent, size=3)
        color, rough, metal, emis = _material(rng, k % emissive_every == 0)
        prims.append(_prim(translate(t) @ rot @ scale(s), color, rough, metal, emis, *ranges[choice[k]]))
    return np.concatenate(verts), np.concatenate(idx).astype(np.uint32), np.concatenate(prims)


# ---------------------------------------------------------------------------------------------
# environment


def cube_directions(n):
    """Unit directions of the texel centres of a cube map (6, n, n, 3), Vulkan face order."""
    c = (np.arange(n) + 0.5) / n * 2.0 - 1.0
    sc, tc = np.meshgrid(c, c, indexing="xy")   # [j, i]
    one = np.ones_like(sc)
    faces = [
        np.stack([one, -tc, -sc], -1),    # +X: sc = -z, tc = -y
        np.stack([-one, -tc, sc], -1),    # -X: sc = +z, tc = -y
        np.stack([sc, one, tc], -1),      # +Y: sc = +x, tc = +z
        np.stack([sc, -one, -tc], -1),    # -Y: sc = +x, tc = -z
        np.stack([sc, -tc, one], -1),     # +Z: sc = +x, tc = -y
        np.stack([-sc, -tc, -one], -1),   # -Z: sc = -x, tc = -y
    ]
    d = np.stack(faces)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def sky_gradient(d):
    """Analytic sky: zenith blue to horizon white above, dark ground below, a soft sun."""
    y = d[..., 1]
    up = np.clip(y, 0.0, 1.0)[..., None]
    sky = (1.0 - np.sqrt(up)) * np.array([1.0, 1.0, 1.0]) + np.sqrt(up) * np.array([0.35, 0.55, 1.0])
    ground = np.array([0.25, 0.22, 0.2]) * np.clip(1.0 + y, 0.0, 1.0)[..., None]
    col = np.where((y >= 0.0)[..., None], sky, ground)
    sun_dir = np.array([0.4, 0.8, 0.45]) / np.linalg.norm([0.4, 0.8, 0.45])
    sun = np.clip(d @ sun_dir, 0.0, 1.0) ** 64 * 20.0
    return col + sun[..., None] * np.array([1.0, 0.9, 0.7])


def sky_cube(n=128):
    """The analytic sky baked into RGBA32F cube faces (6, n, n, 4)."""
    rgb = sky_gradient(cube_directions(n))
    out = np.ones((6, n, n, 4), dtype=np.float32)
    out[..., :3] = rgb
    return out


def latlong_directions(width, height):
    """Unit directions of the texel centres of an equirectangular map (height, width, 3): y up,
    u = atan2(z, x) / 2pi + 1/2, v = polar angle / pi, row 0 at +y."""
    u = (np.arange(width) + 0.5) / width
    v = (np.arange(height) + 0.5) / height
    phi = (u - 0.5) * 2.0 * np.pi
    theta = v * np.pi
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    return np.stack([st * np.cos(phi)[None, :], np.broadcast_to(ct, (height, width)), st * np.sin(phi)[None, :]], -1)


def sky_latlong(width=512, height=256):
    """The analytic sky baked into an RGBA32F equirectangular map (height, width, 4)."""
    out = np.ones((height, width, 4), dtype=np.float32)
    out[..., :3] = sky_gradient(latlong_directions(width, height))
    return out


def constant_cube(n=4, value=1.0):
    out = np.full((6, n, n, 4), value, dtype=np.float32)
    out[..., 3] = 1.0
    return out


# ---------------------------------------------------------------------------------------------
# workload presets (scene + camera)

SMALL = dict(width=256, height=256, spp=16, bounces=8, contribution_factor=4.0, env_face=128,
             eye=(5.0, 5.0, 9.5), target=(5.0, 4.0, 0.0))
LARGE = dict(width=1920, height=1080, spp=64, bounces=12, contribution_factor=4.0, env_face=256,
             eye=(-20.0, 110.0, -20.0), target=(100.0, 95.0, 100.0))
# A second view of the same scene from inside the instance cloud (its centre), for comparison
# with the default view from outside it, where more of the rays reach the sky.
LARGE_INTERIOR = dict(eye=(100.0, 100.0, 100.0), target=(190.0, 85.0, 150.0))

# Learned-cache workload: the spheres-in-box scene at 1920x1080, one sample per frame for 240 frames
# while the camera orbits; errors are measured at the orbit's final pose on every 8th pixel.
CACHE = dict(width=1920, height=1080, frames=240, bounces=8, contribution_factor=4.0, env_face=128,
             centre=(5.0, 3.0, 5.0), radius=3.5, height_y=6.0, eval_stride=8)


def orbit_pose(k, frames=240, centre=(5.0, 3.0, 5.0), radius=3.5, height_y=6.0):
    """Camera eye and target of orbit frame k (frame `frames` returns to frame 0)."""
    a = 2.0 * math.pi * (k % frames) / frames
    eye = (centre[0] + radius * math.cos(a), height_y, centre[2] + radius * math.sin(a))
    return eye, tuple(centre)


def eval_pixels(width=1920, height=1080, stride=8):
    """Pixel centres of every `stride`-th pixel: (stride*i + stride/2, stride*j + stride/2)."""
    xs = np.arange(stride // 2, width, stride)
    ys = np.arange(stride // 2, height, stride)
    X, Y = np.meshgrid(xs, ys)
    return np.stack([X.ravel(), Y.ravel()], 1).astype(np.uint32)
