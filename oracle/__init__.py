"""ctypes wrapper of the CPU oracle (nb_oracle.c).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py as the checker; the product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

NODE = np.dtype([("min", "<f4", 3), ("start", "<u4"), ("max", "<f4", 3), ("end", "<u4")])
VERTEX = np.dtype([("position", "<f4", 3), ("u", "<f4"), ("normal", "<f4", 3), ("v", "<f4"), ("tangent", "<f4", 4)])
INSTANCE = np.dtype([("world_to_local", "<f4", 16), ("local_to_world", "<f4", 16), ("color", "<f4", 4),
                     ("roughness", "<f4"), ("metallic", "<f4"), ("emissive", "<f4"), ("node", "<u4")])
PRIMITIVE = np.dtype([("local_to_world", "<f4", 16), ("color", "<f4", 4), ("roughness", "<f4"),
                      ("metallic", "<f4"), ("emissive", "<f4"), ("index_start", "<u4"), ("index_end", "<u4")])
CAMERA = np.dtype([("world_to_clip", "<f4", 16), ("clip_to_world", "<f4", 16)])
GLOBALS = np.dtype([("sample", "<u4"), ("weight", "<f4"), ("bounces", "<u4"), ("contribution_factor", "<f4")])
RAW_HIT = np.dtype([("instance", "<u4"), ("i0", "<u4"), ("i1", "<u4"), ("i2", "<u4"), ("bary", "<f4", 3),
                    ("dist", "<f4"), ("n_aabb", "<u4"), ("n_tri", "<u4")])
COUNTERS = np.dtype([("rays", "<u8"), ("n_aabb", "<u8"), ("n_tri", "<u8"), ("samples", "<u8")])
NO_HIT = np.float32(3.4028234663852886e38)


class _Scene(ctypes.Structure):
    _fields_ = [("blas", ctypes.c_void_p), ("n_blas", ctypes.c_uint32), ("tlas", ctypes.c_void_p),
                ("n_tlas", ctypes.c_uint32), ("instances", ctypes.c_void_p), ("n_instances", ctypes.c_uint32),
                ("vertices", ctypes.c_void_p), ("n_vertices", ctypes.c_uint32), ("indices", ctypes.c_void_p),
                ("n_indices", ctypes.c_uint32)]


class _Env(ctypes.Structure):
    _fields_ = [("texels", ctypes.c_void_p), ("face_size", ctypes.c_uint32), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32)]


def build():
    """Compile liboracle.so with the committed Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.orc_sin.restype = ctypes.c_float
    lib.orc_sin.argtypes = [ctypes.c_float]
    lib.orc_cos.restype = ctypes.c_float
    lib.orc_cos.argtypes = [ctypes.c_float]
    lib.orc_atan.restype = ctypes.c_float
    lib.orc_atan.argtypes = [ctypes.c_float]
    lib.orc_atan2.restype = ctypes.c_float
    lib.orc_atan2.argtypes = [ctypes.c_float, ctypes.c_float]
    lib.orc_set_roulette_first.argtypes = [ctypes.c_uint32]
    lib.orc_intersect_aabb.restype = ctypes.c_float
    lib.orc_intersect_aabb.argtypes = [ctypes.c_void_p] * 4
    lib.orc_sobol_unscrambled.restype = ctypes.c_uint32
    lib.orc_sobol_unscrambled.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    lib.orc_camera_compute.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_float] * 3 + [ctypes.c_void_p]
    lib.orc_render.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_uint32] * 2 + [ctypes.c_void_p] + \
        [ctypes.c_uint32] * 6 + [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib.orc_render_mean_f64.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_uint32] * 2 + [ctypes.c_void_p] + \
        [ctypes.c_uint32] * 4 + [ctypes.c_void_p, ctypes.c_int]
    lib.orc_render_pixels_mean_f64.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_uint32] * 2 + [ctypes.c_void_p] + \
        [ctypes.c_uint32] * 2 + [ctypes.c_void_p] + [ctypes.c_uint32] * 3 + [ctypes.c_void_p, ctypes.c_int]
    lib.orc_sample_pixel.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_uint32] * 2 + [ctypes.c_void_p] + \
        [ctypes.c_uint32] * 4 + [ctypes.c_void_p, ctypes.c_void_p]
    lib.orc_train_path.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_uint32] * 2 + [ctypes.c_void_p] + \
        [ctypes.c_uint32] * 4 + [ctypes.c_void_p] * 4
    lib.orc_cache_query.restype = ctypes.c_uint32
    lib.orc_cache_query.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_uint32] * 2 + [ctypes.c_void_p] + \
        [ctypes.c_uint32] * 4 + [ctypes.c_float, ctypes.c_uint32] + [ctypes.c_void_p] * 3
    lib.orc_primary_hits.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
    lib.orc_build_scene.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                    ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                    ctypes.c_void_p]
    lib.orc_build_blas.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                   ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                   ctypes.c_void_p]
    for name in ("orc_hash4u", "orc_map4f", "orc_mat4_inverse", "orc_trace", "orc_trace_brute", "orc_env_sample"):
        getattr(lib, name).argtypes = [ctypes.c_void_p] * (2 if name in ("orc_hash4u", "orc_map4f", "orc_mat4_inverse") else 3 if name == "orc_env_sample" else 4)
    lib.orc_mat4_mul.argtypes = [ctypes.c_void_p] * 3
    lib.orc_intersect_triangle.argtypes = [ctypes.c_void_p] * 6
    lib.orc_generate_ray.argtypes = [ctypes.c_void_p] + [ctypes.c_uint32] * 4 + [ctypes.c_void_p] * 3
    lib.orc_sobol_init.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
    lib.orc_sobol_table.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
    lib.orc_sobol_sample4d.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
    _init_sobol(lib)
    return lib


def _init_sobol(lib):
    import scipy
    path = os.path.join(os.path.dirname(scipy.__file__), "stats", "_sobol_direction_numbers.npz")
    d = np.load(path, allow_pickle=False)
    poly = np.ascontiguousarray(d["poly"][:256], dtype=np.int64)
    vinit = np.ascontiguousarray(d["vinit"][:256], dtype=np.int64)
    lib.orc_sobol_init(poly.ctypes.data, vinit.ctypes.data, 256, vinit.shape[1])


lib = _load()


def p(a):
    return None if a is None else a.ctypes.data


# ---- wrappers -------------------------------------------------------------------------------

def hash4u(s):
    s = np.ascontiguousarray(s, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib.orc_hash4u(p(s), p(out))
    return out


def map4f(x):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = np.zeros(4, dtype=np.float32)
    lib.orc_map4f(p(x), p(out))
    return out


def atan(x):
    return lib.orc_atan(float(x))


def atan2(y, x):
    return lib.orc_atan2(float(y), float(x))


def set_roulette_first(bounce):
    """First bounce with Russian roulette in every later oracle render (0: the reference's rule)."""
    lib.orc_set_roulette_first(int(bounce))


def mat4_inverse(m16):
    m = np.ascontiguousarray(m16, dtype=np.float32)
    out = np.zeros(16, dtype=np.float32)
    lib.orc_mat4_inverse(p(m), p(out))
    return out


def camera(eye, target, up, fov, aspect, near):
    cam = np.zeros(1, dtype=CAMERA)
    e, t, u = (np.ascontiguousarray(v, dtype=np.float32) for v in (eye, target, up))
    lib.orc_camera_compute(p(e), p(t), p(u), fov, aspect, near, p(cam))
    return cam


def build_scene(vertices, indices, prims):
    """Returns (blas, tlas, instances, permuted_indices)."""
    v = np.ascontiguousarray(vertices, dtype=VERTEX)
    idx = np.ascontiguousarray(indices, dtype=np.uint32).copy()
    pr = np.ascontiguousarray(prims, dtype=PRIMITIVE)
    tris = (pr["index_end"] // 3 - pr["index_start"] // 3).astype(np.int64)
    bcap = int((2 * tris - 1).sum())
    tcap = 2 * len(pr) - 1
    blas = np.zeros(max(bcap, 1), dtype=NODE)
    tlas = np.zeros(max(tcap, 1), dtype=NODE)
    inst = np.zeros(len(pr), dtype=INSTANCE)
    nb, nt = ctypes.c_uint32(), ctypes.c_uint32()
    rc = lib.orc_build_scene(p(v), len(v), p(idx), len(idx), p(pr), len(pr), p(blas), bcap, ctypes.byref(nb),
                             p(tlas), tcap, ctypes.byref(nt), p(inst))
    if rc != 0:
        raise RuntimeError(f"orc_build_scene failed: {rc}")
    return blas[:nb.value].copy(), tlas[:nt.value].copy(), inst, idx


def build_blas(vertices, indices, tri_start, tri_end):
    v = np.ascontiguousarray(vertices, dtype=VERTEX)
    idx = np.ascontiguousarray(indices, dtype=np.uint32).copy()
    cap = 2 * (tri_end - tri_start)
    nodes = np.zeros(cap, dtype=NODE)
    n = ctypes.c_uint32()
    rc = lib.orc_build_blas(p(v), len(v), p(idx), len(idx), tri_start, tri_end, p(nodes), cap, ctypes.byref(n))
    if rc != 0:
        raise RuntimeError(f"orc_build_blas failed: {rc}")
    return nodes[:n.value].copy(), idx


class Scene:
    """Holds the reference-layout arrays alive for the oracle's tracing calls."""

    def __init__(self, blas, tlas, instances, vertices, indices):
        self.arrays = [np.ascontiguousarray(a) for a in (blas, tlas, instances, vertices, indices)]
        b, t, i, v, x = self.arrays
        self.s = _Scene(p(b), len(b), p(t), len(t), p(i), len(i), p(v), len(v), p(x), len(x))

    @property
    def ref(self):
        return ctypes.byref(self.s)


class Env:
    def __init__(self, faces):
        self.faces = np.ascontiguousarray(faces, dtype=np.float32)
        self.e = _Env(p(self.faces), self.faces.shape[1], 0, 0)

    @classmethod
    def latlong(cls, texels):
        """The lat-long form: (height, width, 4) RGBA32F, row 0 at +y."""
        self = cls.__new__(cls)
        self.faces = np.ascontiguousarray(texels, dtype=np.float32)
        self.e = _Env(p(self.faces), 0, self.faces.shape[1], self.faces.shape[0])
        return self

    @property
    def ref(self):
        return ctypes.byref(self.e)

    def sample(self, d):
        d = np.ascontiguousarray(d, dtype=np.float32)
        out = np.zeros(3, dtype=np.float32)
        lib.orc_env_sample(self.ref, p(d), p(out))
        return out


def trace(scene, o, d, brute=False):
    o = np.ascontiguousarray(o, dtype=np.float32)
    d = np.ascontiguousarray(d, dtype=np.float32)
    h = np.zeros(1, dtype=RAW_HIT)
    (lib.orc_trace_brute if brute else lib.orc_trace)(scene.ref, p(o), p(d), p(h))
    return h[0]


def intersect_triangle(o, d, v0, v1, v2):
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (o, d, v0, v1, v2)]
    out = np.zeros(3, dtype=np.float32)
    lib.orc_intersect_triangle(*[p(a) for a in arrs], p(out))
    return out


def intersect_aabb(o, inv_d, bmin, bmax):
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (o, inv_d, bmin, bmax)]
    return np.float32(lib.orc_intersect_aabb(*[p(a) for a in arrs]))


def sobol_table(rows, dims, seed=0):
    out = np.zeros((rows, dims, 4), dtype=np.float32)
    lib.orc_sobol_table(rows, dims, seed, p(out))
    return out


def primary_hits(scene, cam, width, height, table, stride, sample):
    t = np.ascontiguousarray(table, dtype=np.float32)
    out = np.zeros((height, width, 5), dtype=np.uint32)
    lib.orc_primary_hits(scene.ref, p(cam), width, height, p(t), stride, sample, p(out))
    return out


def globals_(sample, weight, bounces, cf):
    g = np.zeros(1, dtype=GLOBALS)
    g["sample"], g["weight"], g["bounces"], g["contribution_factor"] = sample, weight, bounces, cf
    return g


def render(scene, cam, g, table, stride, env, width, height, rgba=None, rect=None, threads=1):
    """One dispatch of pathtracing.wgsl:main into rgba (height, width, 4); returns (rgba, counters)."""
    t = np.ascontiguousarray(table, dtype=np.float32)
    rows = t.shape[0]
    if rgba is None:
        rgba = np.zeros((height, width, 4), dtype=np.float32)
    x0, y0, x1, y1 = rect if rect is not None else (0, 0, width, height)
    cnt = np.zeros(1, dtype=COUNTERS)
    lib.orc_render(scene.ref, p(cam), p(g), p(t), stride, rows, env.ref, width, height, x0, y0, x1, y1, p(rgba),
                   p(cnt), threads)
    return rgba, cnt[0]


def sample_pixel(scene, cam, g, table, stride, env, width, height, x, y):
    """One path for pixel (x, y) of frame g.sample (orc_sample_pixel): returns (rgb, counters)."""
    t = np.ascontiguousarray(table, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    cnt = np.zeros(1, dtype=COUNTERS)
    lib.orc_sample_pixel(scene.ref, p(cam), p(g), p(t), stride, t.shape[0], env.ref, width, height, x, y, p(out),
                         p(cnt))
    return out, cnt[0]


def train_path(scene, cam, g, table, stride, env, width, height, x, y):
    """One path of pixel (x, y) with its training records (orc_train_path): returns
    (raw (n, 16), target (n, 3), radiance (3,))."""
    t = np.ascontiguousarray(table, dtype=np.float32)
    raw = np.zeros((3, 16), dtype=np.float32)
    target = np.zeros((3, 3), dtype=np.float32)
    rad = np.zeros(3, dtype=np.float32)
    n = ctypes.c_uint32()
    lib.orc_train_path(scene.ref, p(cam), p(g), p(t), stride, t.shape[0], env.ref, width, height, x, y, p(raw),
                       p(target), ctypes.byref(n), p(rad))
    return raw[:n.value], target[:n.value], rad


def cache_query(scene, cam, g, table, stride, env, width, height, x, y, mirror_alpha=0.05, max_cache_bounce=3):
    """The learned cache's query of pixel (x, y) (orc_cache_query): returns (queried, direct (3,),
    raw (16,), throughput (3,)); raw and throughput are zero when the path did not reach its
    query point."""
    t = np.ascontiguousarray(table, dtype=np.float32)
    direct = np.zeros(3, dtype=np.float32)
    raw = np.zeros(16, dtype=np.float32)
    thr = np.zeros(3, dtype=np.float32)
    q = lib.orc_cache_query(scene.ref, p(cam), p(g), p(t), stride, t.shape[0], env.ref, width, height, x, y,
                            mirror_alpha, max_cache_bounce, p(direct), p(raw), p(thr))
    return bool(q), direct, raw, thr


def pixel_shift(x, y):
    """The per-pixel Cranley-Patterson shift hash4f(x, y, x, y) (common.wgsl:27-41, pathtracing.wgsl:252)."""
    return map4f(hash4u(np.array([x, y, x, y], dtype=np.uint32)))


def render_mean_f64(scene, cam, g, table, stride, env, width, height, first, n, threads=1):
    t = np.ascontiguousarray(table, dtype=np.float32)
    out = np.zeros((height, width, 3), dtype=np.float64)
    lib.orc_render_mean_f64(scene.ref, p(cam), p(g), p(t), stride, t.shape[0], env.ref, width, height, first, n,
                            p(out), threads)
    return out


def render_pixels_mean_f64(scene, cam, g, table, stride, env, width, height, xy, first, n, threads=1):
    """float64 mean of samples first..first+n-1 at pixels xy (N x 2 uint32)."""
    t = np.ascontiguousarray(table, dtype=np.float32)
    xy = np.ascontiguousarray(xy, dtype=np.uint32)
    out = np.zeros((len(xy), 3), dtype=np.float64)
    lib.orc_render_pixels_mean_f64(scene.ref, p(cam), p(g), p(t), stride, t.shape[0], env.ref, width, height, p(xy),
                                   len(xy), first, n, p(out), threads)
    return out
