"""Serving at the mixed-request workload's shape: four resident instanced scenes (seeds 0-3 with
1024, 2048, 3072 and 4096 instances, shared BLASes), a Poisson stream of mixed-resolution,
mixed-spp requests that includes a 1920x1080, 64-spp request, driven in real time through
nb_server_* (continuous batching, HBM admission).

  * every served image equals nb_render_frames of its request alone, bit for bit (the caller
    contract of Pathtracer::dispatch, pathtracer.rs:225-239: sample k of a pixel is frame k);
  * a band of the 1080p request and the whole of a small request equal the CPU oracle's
    rendering of the same samples (per-pixel relative error <= 1e-3, bit-identical fraction
    printed);
  * with per-scene learned caches (scenes 0 and 2 BF16, 1 and 3 FP16), cache-assisted requests
    (inference only) equal nb_render_cached of their samples alone, and plain requests served in
    the same steps still equal nb_render_frames.
"""
import ctypes

import numpy as np
import pytest

import oracle
from tests import _workloads as W

pytestmark = pytest.mark.gpu

REL_TOL = 1e-3
BOUNCES = 8
CF = 4.0
SCENES = ((0, 1024), (1, 2048), (2, 3072), (3, 4096))


@pytest.fixture(scope="module")
def world():
    import nbounce_amd as nb
    from nbounce_amd import scenegen
    from nbounce_amd._lib import check, lib, ptr
    print("setup: 4 resident instanced scenes (1024..4096 instances)", flush=True)
    ctx = nb.Context(0)
    table = W.table()
    check(lib.nb_sobol_set(ctx.handle, ptr(np.ascontiguousarray(table)), table.shape[0], table.shape[1], 17))
    faces = scenegen.sky_cube(64)
    env = nb.EnvMap(ctx, faces)
    scenes, arrays = [], []
    for seed, n in SCENES:
        w = W.large(seed, n)
        sc = nb.Scene(w["vertices"], w["indices_in"].copy(), w["primitives"])
        b, t, inst = nb.build_scene(sc, share_blas=True)
        scenes.append(nb.SceneBuffers(ctx, b, t, inst, sc.vertices, sc.indices))
        arrays.append((b, t, inst, sc.vertices, sc.indices))
    yield dict(nb=nb, ctx=ctx, table=table, faces=faces, env=env, scenes=scenes, arrays=arrays)
    ctx.close()


def _camera(nb, r):
    return nb.CameraController(world_position=r["eye"], target=r["target"],
                               aspect_ratio=r["width"] / r["height"]).data()


def _stream(n, seed, big_spp=64):
    """A Poisson stream of the serving workload (serving.mixed_requests) with its first request
    replaced by a 1920x1080 request of `big_spp` samples that looks into the instance cloud."""
    from nbounce_amd import serving
    reqs = serving.mixed_requests(n, 40.0, 4, seed=seed)
    reqs[0].update(width=1920, height=1080, spp=big_spp, scene=3, eye=np.array([20.0, 90.0, 30.0]),
                   target=np.array([110.0, 100.0, 120.0]))
    return reqs


def _plain_alone(w, r):
    nb = w["nb"]
    from nbounce_amd._lib import check, lib, ptr
    t = ctypes.c_void_p()
    check(lib.nb_target_create(w["ctx"].handle, r["width"], r["height"], ctypes.byref(t)))
    g = nb.Globals(1, 0.0, BOUNCES, CF).array()
    check(lib.nb_render_frames(w["ctx"].handle, w["scenes"][r["scene"]].handle, t, ptr(_camera(nb, r)), ptr(g),
                               r["spp"], w["env"].handle))
    img = np.zeros((r["height"], r["width"], 4), dtype=np.float32)
    check(lib.nb_target_read(t, ptr(img)))
    lib.nb_target_destroy(t)
    return img


def _oracle_rows(w, r, y0, y1):
    """The oracle's progressive image of request r over rows [y0, y1) (samples 1..spp)."""
    nb = w["nb"]
    b, t, inst, v, idx = w["arrays"][r["scene"]]
    osc, oenv = oracle.Scene(b, t, inst, v, idx), oracle.Env(w["faces"])
    cam = _camera(nb, r)
    img = np.zeros((r["height"], r["width"], 4), dtype=np.float32)
    for k in range(1, r["spp"] + 1):
        g = oracle.globals_(k, np.float32(1.0) / np.float32(k), BOUNCES, CF)
        img, _ = oracle.render(osc, cam, g, w["table"], 17, oenv, r["width"], r["height"], rgba=img,
                               rect=(0, y0, r["width"], y1), threads=16)
    return img[y0:y1]


def _compare(tag, got, ref):
    err = W.rel_err(got, ref)
    same = (got.view(np.uint32) == ref.view(np.uint32)).all(axis=-1).mean()
    print(f"{tag}: bit-identical pixels {same:.5f}, max rel err {err.max():.3e}", flush=True)
    assert np.isfinite(got).all() and err.max() <= REL_TOL


def test_mixed_stream_with_a_1080p_request(world):
    from nbounce_amd import serving
    nb, ctx = world["nb"], world["ctx"]
    reqs = _stream(12, seed=5)
    print("case: 12-request Poisson stream over 4 resident scenes, first request 1920x1080 x 64 spp", flush=True)
    srv = serving.Server(ctx, max_paths=1 << 23, max_bounces=12)
    for sb in world["scenes"]:
        srv.add_scene(sb, world["env"])
    lat, wall, images = serving.run_open_loop(srv, reqs, lambda r: _camera(nb, r), bounces=BOUNCES,
                                              contribution_factor=CF, keep_images=True)
    st = srv.stats()
    print(f"served {len(images)} requests in {wall:.3f} s, {st['steps']} steps, p50 {np.percentile(lat, 50) * 1e3:.1f} ms, "
          f"p99 {np.percentile(lat, 99) * 1e3:.1f} ms", flush=True)
    assert len(images) == len(reqs) and st["steps"] >= 2
    assert st["samples"] == sum(r["width"] * r["height"] * r["spp"] for r in reqs)
    srv.close()
    # each request as it was served equals the plain renderer of that request alone
    for k, r in enumerate(reqs):
        np.testing.assert_array_equal(images[k].view(np.uint32), _plain_alone(world, r).view(np.uint32))
    # anchored to the CPU oracle: a band through the middle of the 1080p request (64 samples)
    # and the whole of the smallest request
    _compare("1080p request, rows 536..543", images[0][536:544], _oracle_rows(world, reqs[0], 536, 544))
    k = min(range(1, len(reqs)), key=lambda j: reqs[j]["width"] * reqs[j]["height"] * reqs[j]["spp"])
    r = reqs[k]
    _compare(f"request {k} ({r['width']}x{r['height']}, {r['spp']} spp, scene {r['scene']})",
             images[k], _oracle_rows(world, r, 0, r["height"]))


def test_mixed_stream_with_per_scene_caches(world):
    """Cache-assisted and plain requests of one stream: scenes 0 and 2 keep BF16 caches, 1 and 3
    FP16 caches (each trained for a few frames first); every cache-assisted request (inference
    only) equals nb_render_cached of its samples alone on its scene's cache, every plain request
    equals nb_render_frames alone."""
    from nbounce_amd import serving
    from nbounce_amd._lib import check, lib, ptr
    from nbounce_amd.cache import RadianceCache
    nb, ctx = world["nb"], world["ctx"]
    print("case: stream with per-scene BF16/FP16 caches, half the requests cache-assisted", flush=True)
    caches = [RadianceCache(ctx, [0, 0, 0], [200, 200, 200], seed=s, fp16=bool(s % 2)) for s, _ in SCENES]
    reqs = _stream(10, seed=8, big_spp=4)
    rng = np.random.default_rng(2)
    for k, r in enumerate(reqs):
        r["cached"] = k == 0 or bool(rng.random() < 0.5)   # the 1080p request is cache-assisted
    # a few training frames per cache from the stream's cameras
    for s, (sb, cache) in enumerate(zip(world["scenes"], caches)):
        cams = [r for r in reqs if r["scene"] == s] or reqs[:1]
        for f in range(6):
            r = cams[f % len(cams)]
            cam = nb.CameraController(world_position=r["eye"], target=r["target"], aspect_ratio=1.0)
            pt = nb.Pathtracer(ctx, sb, cam, world["env"], 128, 128, resolution_factor=1.0, table=world["table"],
                               sobol_stride=17)
            cache.render(pt, train_paths=4096, lr=1e-3)
    srv = serving.Server(ctx, max_paths=1 << 23, max_bounces=12)
    for s, sb in enumerate(world["scenes"]):
        srv.add_scene(sb, world["env"])
        srv.set_cache(s, caches[s])
    lat, wall, images = serving.run_open_loop(srv, reqs, lambda r: _camera(nb, r), bounces=BOUNCES,
                                              contribution_factor=CF, keep_images=True, train_paths=0)
    st = srv.stats()
    n_cached = sum(r["cached"] for r in reqs)
    print(f"served {len(images)} requests ({n_cached} cache-assisted) in {wall:.3f} s, {st['steps']} steps", flush=True)
    assert st["cached_samples"] == sum(r["width"] * r["height"] * r["spp"] for r in reqs if r["cached"])
    srv.close()
    for k, r in enumerate(reqs):
        if not r["cached"]:
            np.testing.assert_array_equal(images[k].view(np.uint32), _plain_alone(world, r).view(np.uint32))
            continue
        t = ctypes.c_void_p()
        check(lib.nb_target_create(ctx.handle, r["width"], r["height"], ctypes.byref(t)))
        cam = _camera(nb, r)
        for j in range(1, r["spp"] + 1):
            g = nb.Globals(j, float(np.float32(1.0) / np.float32(j)), BOUNCES, CF).array()
            check(lib.nb_render_cached(ctx.handle, world["scenes"][r["scene"]].handle, t, ptr(cam), ptr(g),
                                       world["env"].handle, caches[r["scene"]].handle, 0, 0.0, None))
        want = np.zeros((r["height"], r["width"], 4), dtype=np.float32)
        check(lib.nb_target_read(t, ptr(want)))
        lib.nb_target_destroy(t)
        assert np.isfinite(images[k]).all() and images[k][..., :3].max() > 0
        np.testing.assert_array_equal(images[k].view(np.uint32), want.view(np.uint32))
    for c in caches:
        c.close()
