"""Benchmark: Mrays/s of the path tracer at 1080p on one MI355X.

Workload (default): the instanced-mesh scene (16 base meshes, 4096 instances, ~26M instanced
triangles, seed 0) at 1920x1080, 12 bounces, contribution factor 4, one sample per pixel per step
(one reference dispatch, pathtracer.rs:225-239), 64 steps = the 64-spp image.  A ray is one
intersect_scene call (primary or continuation).  Inputs are resident in HBM before timing.

Prints one JSON line with the metric, a roofline object for the render kernel (algorithmic bytes
per launch / average launch time from HIP events on the render stream) and a CPU baseline (the
C oracle on a bounded band of rows of the same frame, OpenMP threads as stated).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def algorithmic_bytes(c, pipelined=False):
    """Bytes the traversal and shading must read/write: 32 per box test, 48 per triangle test,
    64 per instance entry, 304 per hit (3 x 48-B vertices + 160-B instance), 64 per miss (four
    RGBA32F texels), 32 per ray (two Sobol entries), per sample the 16-B jitter entry and the
    RGBA32F read-modify-write of the target (48 B), or, for the path kernel of a pipelined
    nb_render, the 16-B jitter entry and the sample's 16-B write (32 B; the mix kernel moves
    the other 48 B per pixel)."""
    misses = c["rays"] - c["hits"]
    per_sample = 32 if pipelined else 48
    return (32 * c["n_aabb"] + 48 * c["n_tri"] + 64 * c["n_inst"] + 304 * c["hits"] + 64 * misses
            + 32 * c["rays"] + per_sample * c["samples"])


def build_workload(name, share_blas=True):
    from nbounce_amd import scenegen
    import nbounce_amd as nb
    if name == "large":
        P = scenegen.LARGE
        v, i, pr = scenegen.instanced_meshes(0)
        desc = "instanced meshes: 16 base meshes, 4096 instances"
    else:
        P = scenegen.SMALL
        v, i, pr = scenegen.spheres_in_box(0)
        desc = "64 icospheres in an open box"
    scene = nb.Scene(v, i, pr)
    blas, tlas, inst = nb.build_scene(scene, share_blas=share_blas)
    n_inst_tris = int(((pr["index_end"] - pr["index_start"]) // 3).sum())
    return P, desc, scene, (blas, tlas, inst), n_inst_tris


def cpu_baseline(P, scene, arrays, env_faces, cam, table, target_s=15.0):
    """Times the C oracle (the CPU restatement of the reference shader) on the same workload:
    whole frames 1, 2, ... until about `target_s` seconds, or a band of rows of frame 1 when one
    frame would take longer than that."""
    import oracle
    blas, tlas, inst = arrays
    osc = oracle.Scene(blas, tlas, inst, scene.vertices, scene.indices)
    oenv = oracle.Env(env_faces)
    threads = max(1, min(16, os.cpu_count() or 1))
    W, H = P["width"], P["height"]
    g1 = oracle.globals_(1, 1.0, P["bounces"], P["contribution_factor"])
    band = 16
    y0 = H // 2 - band // 2
    t = time.perf_counter()
    oracle.render(osc, cam, g1, table, 17, oenv, W, H, rect=(0, y0, W, y0 + band), threads=threads)
    frame_s = (time.perf_counter() - t) * H / band
    rays = 0
    t0 = time.perf_counter()
    if frame_s < target_s:
        frames = 0
        while time.perf_counter() - t0 < target_s and frames < 64:
            g = oracle.globals_(frames + 1, 1.0, P["bounces"], P["contribution_factor"])
            _, cnt = oracle.render(osc, cam, g, table, 17, oenv, W, H, threads=threads)
            rays += int(cnt["rays"])
            frames += 1
        what = f"frames 1..{frames} of {W}x{H}"
    else:
        rows = int(min(H, max(8, band * target_s / max(frame_s * band / H, 1e-3))))
        y0 = max(0, H // 2 - rows // 2)
        _, cnt = oracle.render(osc, cam, g1, table, 17, oenv, W, H, rect=(0, y0, W, y0 + rows), threads=threads)
        rays = int(cnt["rays"])
        what = f"frame 1, rows {y0}..{y0 + rows} of {W}x{H}"
    dt = time.perf_counter() - t0
    return {"value": round(rays / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"C oracle, {what} ({rays} rays in {dt:.1f} s, {threads} OpenMP threads)"}


def calibration_copy(mib):
    """One device-to-device copy of `mib` MiB through the HIP runtime (its copy kernel), so that
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE can be scaled against a kernel of known traffic."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), n) == 0 and hip.hipMalloc(ctypes.byref(b), n) == 0
    assert hip.hipMemcpy(b, a, n, 3) == 0          # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def serving_scenes(nb, ctx, srv, env):
    """The serving workload's 4 resident instanced scenes (seeds 0-3, 1024..4096 instances)."""
    from nbounce_amd import scenegen
    keep = []
    for seed, n_inst in zip(range(4), (1024, 2048, 3072, 4096)):
        v, i, pr = scenegen.instanced_meshes(seed, n_instances=n_inst)
        sc = nb.Scene(v, i, pr)
        b, t, inst = nb.build_scene(sc, share_blas=True)
        sb = nb.SceneBuffers(ctx, b, t, inst, sc.vertices, sc.indices)
        srv.add_scene(sb, env)
        keep.append(sb)
    return keep


def serve_streams(qps_list, requests, max_paths=1 << 23):
    """Serving workload: 4 resident instanced scenes (seeds 0-3 with 1024..4096 instances),
    Poisson arrivals at each QPS in turn (a fresh stream each), resolution from {256^2, 512^2,
    1024x768, 1920x1080} with p = 0.4/0.3/0.2/0.1, spp uniform over {1, 4, 16, 64}, random
    cameras, 8 bounces, HBM admission budget 200 GB.  Returns per QPS: request latency p50/p99,
    aggregate Mrays/s and the GPU's busy fraction."""
    import nbounce_amd as nb
    from nbounce_amd import scenegen, serving
    from nbounce_amd._lib import check, lib, ptr
    ctx = nb.Context(0)
    table = nb.sobol_table(1024, 17, 0)
    check(lib.nb_sobol_set(ctx.handle, ptr(table), 1024, 17, 17))
    srv = serving.Server(ctx, hbm_budget_bytes=200 << 30, max_paths=max_paths, max_bounces=12)
    env = nb.EnvMap(ctx, scenegen.sky_cube(256))
    keep = serving_scenes(nb, ctx, srv, env)

    def camera(r):
        return nb.CameraController(world_position=r["eye"], target=r["target"],
                                   aspect_ratio=r["width"] / r["height"]).data()

    # warm-up: two small requests so first-use costs stay out of the timed runs
    warm = serving.mixed_requests(2, 20.0, 4, seed=99)
    serving.run_open_loop(srv, [dict(r, arrival=0.0) for r in warm], camera)
    out = []
    for qps in qps_list:
        reqs = serving.mixed_requests(requests, qps, 4, seed=0)
        s0 = srv.stats()
        lat, wall, _ = serving.run_open_loop(srv, reqs, camera)
        s1 = srv.stats()
        out.append({
            "qps": qps,
            "requests": requests,
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
            "mean_ms": round(float(lat.mean()) * 1e3, 2),
            "aggregate_mrays_s": round((s1["rays"] - s0["rays"]) / wall / 1e6, 2),
            "gpu_busy_frac": round((s1["gpu_busy_s"] - s0["gpu_busy_s"]) / wall, 3),
            "offered_samples_per_s_M": round(sum(r["width"] * r["height"] * r["spp"] for r in reqs) / reqs[-1]["arrival"] / 1e6, 2),
            "steps": s1["steps"] - s0["steps"],
            "wall_s": round(wall, 3),
        })
    srv.close()
    del keep
    ctx.close()
    return out


SERVE_CONFIG = {"workload": "4 resident instanced scenes (1024-4096 instances), Poisson arrivals, mixed resolution "
                            "(256^2..1920x1080) and spp (1..64), 8 bounces", "hbm_budget_gb": 200}


def serve_bench(args):
    """The serving workload at --qps (a comma-separated list gives a sweep, one line each)."""
    for r in serve_streams([float(q) for q in str(args.qps).split(",")], args.requests, args.max_paths):
        line = {"metric": "serving request latency p50/p99 and aggregate Mrays/s", "value": r["p99_ms"],
                "unit": "ms (p99 latency)", "higher_is_better": False, "n_gpus": 1}
        line.update(r)
        line["config"] = dict(SERVE_CONFIG, max_paths_per_step=args.max_paths)
        print(json.dumps(line), flush=True)


def view_rate(nb, ctx, sb, env, P, eye, target, frames=32):
    """Mrays/s of pipelined 1-spp nb_render calls at 1920x1080 from another camera: an untimed
    counted pass over the frames gives the rays, a timed pass renders them again."""
    cam = nb.CameraController(world_position=eye, target=target, aspect_ratio=P["width"] / P["height"])
    pt = nb.Pathtracer(ctx, sb, cam, env, P["width"], P["height"], resolution_factor=1.0)
    pt.globals.bounces = P["bounces"]
    pt.globals.contribution_factor = P["contribution_factor"]
    ctx.enable_counters(True)
    ctx.reset_counters()
    for _ in range(frames):
        pt.dispatch()
    ctx.synchronize()
    c = ctx.counters()
    pt.invalidate()
    ctx.enable_counters(False)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        pt.dispatch()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.enable_counters(True)
    rays = c["rays"]
    return {"value": round(rays / wall / 1e6, 3), "unit": "Mrays/s", "ms_per_frame": round(wall * 1e3 / frames, 4),
            "rays_per_sample": round(rays / max(1, c["samples"]), 4),
            "per_ray": {k: round(c[k] / max(1, rays), 3) for k in ("n_aabb", "n_tri", "n_inst", "hits")},
            "config": {"workload": f"the instanced scene, {P['width']}x{P['height']}, {P['bounces']} bounces, "
                                   f"{frames} frames of 1 spp, camera at {tuple(eye)} looking at {tuple(target)} "
                                   "(inside the instance cloud)"}}


def moving_camera_rate(nb, ctx, sb, env, P, frames=32):
    """Mrays/s of 1-spp nb_render calls at 1920x1080 while the camera moves every frame (each move
    invalidates the accumulation, as CameraController updates do), pipelined and, for comparison,
    one launch after another (NB_PIPELINE=0).  An untimed counted pass gives the rays."""
    def views():
        for k in range(frames):
            eye = (P["eye"][0] + 0.25 * k, P["eye"][1] + 0.1 * k, P["eye"][2] + 0.15 * k)
            yield nb.CameraController(world_position=eye, target=P["target"], aspect_ratio=P["width"] / P["height"])

    def run(pt):
        for cam in views():
            pt.update(cam)
            pt.dispatch()

    pt = nb.Pathtracer(ctx, sb, next(views()), env, P["width"], P["height"], resolution_factor=1.0)
    pt.globals.bounces = P["bounces"]
    pt.globals.contribution_factor = P["contribution_factor"]
    ctx.enable_counters(True)
    ctx.reset_counters()
    run(pt)
    ctx.synchronize()
    rays = ctx.counters()["rays"]
    ctx.enable_counters(False)
    out = {}
    for mode in ("1", "0"):
        old = os.environ.get("NB_PIPELINE")
        os.environ["NB_PIPELINE"] = mode
        try:
            ctx.synchronize()
            t0 = time.perf_counter()
            run(pt)
            ctx.synchronize()
            out[mode] = time.perf_counter() - t0
        finally:
            if old is None:
                del os.environ["NB_PIPELINE"]
            else:
                os.environ["NB_PIPELINE"] = old
    ctx.enable_counters(True)
    return {"value": round(rays / out["1"] / 1e6, 3), "unit": "Mrays/s", "ms_per_frame": round(out["1"] * 1e3 / frames, 4),
            "one_launch_after_another": round(rays / out["0"] / 1e6, 3),
            "config": {"workload": f"the default workload, {frames} frames of 1 spp, the camera moving (and the "
                                   "accumulation restarting) every frame"}}


def small_workload_rate(frames=16, native=False):
    """The small workload (64 icospheres in a box, 256x256, 8 bounces) at `frames` samples per
    pixel in one nb_render_frames call (sample-parallel): Mrays/s over an untimed counted pass and
    a timed one of the same frames.  native: as the configuration is written (Russian roulette
    from bounce 3, the sky as a 512x256 RGBA32F lat-long map) instead of the reference's rules
    (roulette from bounce 0, the sky baked into a cube)."""
    import nbounce_amd as nb
    from nbounce_amd import scenegen
    P, desc, scene, arrays, _ = build_workload("small")
    blas, tlas, inst = arrays
    ctx = nb.Context(0)
    sb = nb.SceneBuffers(ctx, blas, tlas, inst, scene.vertices, scene.indices)
    if native:
        env = nb.EnvMap.latlong(ctx, scenegen.sky_latlong(512, 256))
        ctx.set_render_options(roulette_first_bounce=3)
        desc += ", Russian roulette from bounce 3, 512x256 lat-long sky"
    else:
        env = nb.EnvMap(ctx, scenegen.sky_cube(P["env_face"]))
    cam = nb.CameraController(world_position=P["eye"], target=P["target"], aspect_ratio=1.0)
    pt = nb.Pathtracer(ctx, sb, cam, env, P["width"], P["height"], resolution_factor=1.0)
    pt.globals.bounces, pt.globals.contribution_factor = P["bounces"], P["contribution_factor"]
    pt.dispatch_frames(frames)                    # warm-up
    pt.invalidate()
    ctx.reset_counters()
    pt.dispatch_frames(frames)
    ctx.synchronize()
    c = ctx.counters()
    pt.invalidate()
    ctx.enable_counters(False)
    ctx.synchronize()
    t0 = time.perf_counter()
    pt.dispatch_frames(frames)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.close()
    return {"value": round(c["rays"] / wall / 1e6, 2), "unit": "Mrays/s", "ms_per_image": round(wall * 1e3, 3),
            "rays": int(c["rays"]),
            "config": {"workload": f"{desc}, {P['width']}x{P['height']}, {frames} spp in one call, {P['bounces']} bounces"}}


def serve_cache_bench(args):
    """Serving with learned caches: the serving workload (4 resident instanced scenes, Poisson
    arrivals at --qps, mixed resolution and spp, 8 bounces) where every scene keeps its own
    device-resident cache and Adam state (scenes 0 and 2 BF16, 1 and 3 FP16 with loss scaling)
    and each request is plain or cache-assisted with probability 1/2.  A cache-assisted request
    asks for the samples that reach the plain request's relative MSE: its spp divided by the
    plain-spp equivalent of one cached sample, measured first on scene 0 against a 1024-spp plain
    reference.  Cached requests train their scene's cache online.  Reports p50/p99 latency and
    Mrays/s per mode."""
    import nbounce_amd as nb
    from nbounce_amd import scenegen, serving
    from nbounce_amd.cache import RadianceCache, rel_mse
    from nbounce_amd._lib import check, lib, ptr
    ctx = nb.Context(0)
    table = nb.sobol_table(1024, 17, 0)
    check(lib.nb_sobol_set(ctx.handle, ptr(table), 1024, 17, 17))
    srv = serving.Server(ctx, hbm_budget_bytes=200 << 30, max_paths=args.max_paths, max_bounces=12)
    env = nb.EnvMap(ctx, scenegen.sky_cube(256))
    keep, caches = [], []
    for seed, n_inst in zip(range(4), (1024, 2048, 3072, 4096)):
        v, i, pr = scenegen.instanced_meshes(seed, n_instances=n_inst)
        sc = nb.Scene(v, i, pr)
        b, t, inst = nb.build_scene(sc, share_blas=True)
        sb = nb.SceneBuffers(ctx, b, t, inst, sc.vertices, sc.indices)
        idx = srv.add_scene(sb, env)
        cache = RadianceCache(ctx, [0, 0, 0], [200, 200, 200], seed=seed, fp16=bool(seed % 2))
        srv.set_cache(idx, cache)
        keep.append(sb)
        caches.append(cache)
    qps = float(str(args.qps).split(",")[0])
    reqs = serving.mixed_requests(args.requests, qps, 4, seed=0)

    def camera(r):
        return nb.CameraController(world_position=r["eye"], target=r["target"],
                                   aspect_ratio=r["width"] / r["height"]).data()

    # warm-up training of every cache: cached frames at 256x256 from request cameras
    rng = np.random.default_rng(1)
    for k, (sb, cache) in enumerate(zip(keep, caches)):
        for f in range(args.cache_warmup):
            r = reqs[int(rng.integers(0, len(reqs)))]
            cam = nb.CameraController(world_position=r["eye"], target=r["target"], aspect_ratio=1.0)
            pt = nb.Pathtracer(ctx, sb, cam, env, 256, 256, resolution_factor=1.0, table=table, sobol_stride=17)
            cache.render(pt, train_paths=8192, lr=args.lr)

    # equal-error calibration on scene 0 (128x128, a request camera)
    r0 = next(r for r in reqs if r["scene"] == 0)
    cam = nb.CameraController(world_position=r0["eye"], target=r0["target"], aspect_ratio=1.0)
    ref_pt = nb.Pathtracer(ctx, keep[0], cam, env, 128, 128, resolution_factor=1.0, table=table, sobol_stride=17)
    ref_pt.dispatch_frames(1023)
    ref = ref_pt.output_texture()
    plain_err = []
    for n in (1, 2, 4, 8):
        pt = nb.Pathtracer(ctx, keep[0], cam, env, 128, 128, resolution_factor=1.0, table=table, sobol_stride=17)
        pt.dispatch_frames(n)
        plain_err.append((n, rel_mse(pt.output_texture(), ref)))
    pt = nb.Pathtracer(ctx, keep[0], cam, env, 128, 128, resolution_factor=1.0, table=table, sobol_stride=17)
    caches[0].render(pt, train_paths=0)
    e1 = rel_mse(pt.output_texture(), ref)
    xs, ys = np.log([p[0] for p in plain_err]), np.log([p[1] for p in plain_err])
    factor = float(np.exp(np.interp(-np.log(e1), -ys, xs)))
    factor = max(1.0, factor)
    for r in reqs:
        r["cached"] = bool(rng.random() < 0.5)
        if r["cached"]:
            r["spp_plain"] = r["spp"]
            r["spp"] = max(1, int(round(r["spp"] / factor)))

    serving.run_open_loop(srv, [dict(r, arrival=0.0, cached=False) for r in reqs[:2]], camera)
    s0 = srv.stats()
    ctx.reset_counters()
    lat, wall, _ = serving.run_open_loop(srv, reqs, camera, train_paths=args.serve_train_paths, lr=args.lr)
    s1 = srv.stats()
    cc = ctx.counters()
    cached = np.array([r["cached"] for r in reqs])
    plain_rays = s1["rays"] - s0["rays"]
    plain_busy = (s1["gpu_busy_s"] - s0["gpu_busy_s"]) - (s1["cached_busy_s"] - s0["cached_busy_s"])
    cached_busy = s1["cached_busy_s"] - s0["cached_busy_s"]

    def pct(a, q):
        return round(float(np.percentile(a, q)) * 1e3, 2) if len(a) else None

    out = {
        "metric": "serving with learned caches: equal-relative-MSE request latency p50/p99 per mode",
        "value": pct(lat[cached], 99),
        "unit": "ms (p99 latency, cache-assisted requests)",
        "higher_is_better": False,
        "plain": {"p50_ms": pct(lat[~cached], 50), "p99_ms": pct(lat[~cached], 99), "requests": int((~cached).sum()),
                  "mrays_s_busy": round(plain_rays / max(plain_busy, 1e-9) / 1e6, 2)},
        "cached": {"p50_ms": pct(lat[cached], 50), "p99_ms": pct(lat[cached], 99), "requests": int(cached.sum()),
                   "mrays_s_busy": round(cc["rays"] / max(cached_busy, 1e-9) / 1e6, 2),
                   "samples": int(s1["cached_samples"] - s0["cached_samples"])},
        "equal_error_plain_spp_per_cached_spp": round(factor, 3),
        "calibration": {"plain": [[n, round(e, 5)] for n, e in plain_err], "cached_1spp": round(e1, 5)},
        "loss_scale_fp16": [c.loss_scale() for c in caches if c.fp16],
        "aggregate_mrays_s": round((plain_rays + cc["rays"]) / wall / 1e6, 2),
        "gpu_busy_frac": round((s1["gpu_busy_s"] - s0["gpu_busy_s"]) / wall, 3),
        "requests": args.requests,
        "qps": qps,
        "wall_s": round(wall, 3),
        "n_gpus": 1,
        "config": {"workload": "4 resident instanced scenes (1024-4096 instances) with per-scene caches "
                               "(BF16, FP16, BF16, FP16), Poisson arrivals, mixed resolution and spp, 8 bounces, "
                               "half the requests cache-assisted", "cache_warmup_frames": args.cache_warmup,
                   "serve_train_paths": args.serve_train_paths, "hbm_budget_gb": 200},
    }
    print(json.dumps(out), flush=True)


def cache_equal_error(train_paths=1 << 16, lr=1e-3, max_batch=None, frames=None, seed=0):
    """Learned radiance cache on the spheres-in-box scene at 1920x1080, 8 bounces: `frames` (240)
    frames of one sample while the camera orbits, each frame training the cache on `train_paths`
    full paths (up to three records each; max_batch records per frame, by default all of them);
    then, at the orbit's final pose, relative MSE (every 8th pixel, against the committed float64
    4096-spp CPU-oracle reference) and device time of the plain tracer at 1..128 spp and of the
    cache-assisted tracer (trained network, no further training) at 1..32 spp.  Returns the
    equal-error frame-time ratios and the measurements behind them."""
    import nbounce_amd as nb
    from nbounce_amd import scenegen
    from nbounce_amd.cache import RadianceCache, rel_mse
    P = scenegen.CACHE
    frames = P["frames"] if frames is None else frames
    if max_batch is None:
        max_batch = min(3 * train_paths, 1 << 20)
    ref = np.load(os.path.join(ROOT, "tests", "golden", "cache_reference_spheres_1080p.npz"), allow_pickle=False)
    xy, ref_mean = ref["xy"], ref["mean"].astype(np.float64)
    v, i, pr = scenegen.spheres_in_box(0)
    scene = nb.Scene(v, i, pr)
    blas, tlas, inst = nb.build_scene(scene)            # the reference's per-primitive BLAS, as the oracle
    ctx = nb.Context(0)
    ctx.enable_counters(False)     # both tracers timed as the renderer is: counter-free kernels
    sb = nb.SceneBuffers(ctx, blas, tlas, inst, scene.vertices, scene.indices)
    env = nb.EnvMap(ctx, scenegen.sky_cube(P["env_face"]))
    W, H = P["width"], P["height"]

    def tracer(k):
        eye, target = scenegen.orbit_pose(k, P["frames"], P["centre"], P["radius"], P["height_y"])
        cam = nb.CameraController(world_position=eye, target=target, aspect_ratio=W / H)
        pt = nb.Pathtracer(ctx, sb, cam, env, W, H, resolution_factor=1.0)
        pt.globals.bounces, pt.globals.contribution_factor = P["bounces"], P["contribution_factor"]
        return pt

    def err(pt):
        img = pt.output_texture()
        return rel_mse(img[xy[:, 1], xy[:, 0]], ref_mean)

    cache = RadianceCache(ctx, [0, 0, 0], [10, 10, 10], seed=seed, max_batch=max_batch)
    ctx.enable_render_timing()
    train_ms, losses = [], []
    plain_orbit_ms = []
    for k in range(frames):
        pt = tracer(k)
        losses.append(cache.render(pt, train_paths=train_paths, lr=lr))
        train_ms.append(ctx.last_render_ms())
        if k % 20 == 0:
            pq = tracer(k)
            pq.dispatch()
            plain_orbit_ms.append(ctx.last_render_ms())
    final = P["frames"]
    plain = []
    for n in (1, 2, 4, 8, 16, 32, 64, 128):
        pt = tracer(final)
        ctx.timer_start()
        pt.dispatch_frames(n)
        ms = ctx.timer_stop()
        plain.append({"spp": n, "rel_mse": err(pt), "ms": round(ms, 3)})
    cached = []
    for n in (1, 2, 4, 8, 16, 32):
        pt = tracer(final)
        ctx.timer_start()
        for _ in range(n):
            cache.render(pt, train_paths=0)
        ms = ctx.timer_stop()
        cached.append({"spp": n, "rel_mse": err(pt), "ms": round(ms, 3)})
    # plain spp (log-log interpolation) that reaches the 1-spp cache-assisted error
    e1, t1 = cached[0]["rel_mse"], cached[0]["ms"]
    xs = np.log([p["spp"] for p in plain])
    ys = np.log([p["rel_mse"] for p in plain])
    spp_eq = float(np.exp(np.interp(-np.log(e1), -ys, xs))) if e1 > plain[-1]["rel_mse"] else float("inf")
    ms_per_spp = plain[-1]["ms"] / plain[-1]["spp"]
    t_plain_eq = spp_eq * ms_per_spp
    t_cache_train = float(np.mean(train_ms[-20:]))
    cache.close()
    ctx.close()
    return {
        "ratio": round(t1 / t_plain_eq, 4) if np.isfinite(t_plain_eq) else None,
        "with_training_ratio": round(t_cache_train / t_plain_eq, 4) if np.isfinite(t_plain_eq) else None,
        "cache_1spp": {"rel_mse": e1, "ms": t1},
        "plain_spp_at_equal_error": round(spp_eq, 3),
        "plain_ms_per_spp": round(ms_per_spp, 4),
        "plain_ms_per_frame_orbit": round(float(np.mean(plain_orbit_ms)), 4),
        "cache_frame_with_training_ms": round(t_cache_train, 4),
        "loss_first_last": [round(losses[0], 4), round(float(np.mean(losses[-10:])), 4)],
        "plain": plain,
        "cached": cached,
        "config": {"workload": "spheres-in-box scene, 1920x1080, 8 bounces, %d-frame orbit, 1 spp per frame" % frames,
                   "train_paths_per_frame": train_paths, "lr": lr, "max_records_per_frame": max_batch,
                   "reference": "CPU oracle float64 mean of 4096 spp at every 8th pixel"},
    }


def cache_bench(args):
    """Equal relative-MSE frame time of the cache-assisted renderer over the plain one
    (cache_equal_error) as one JSON line."""
    r = cache_equal_error(args.train_paths, args.lr, args.max_batch or None)
    out = {
        "metric": "equal relative-MSE frame time, cache-assisted / plain",
        "value": r.pop("ratio"),
        "unit": "ratio (lower is better)",
        "higher_is_better": False,
        "n_gpus": 1,
    }
    out.update(r)
    print(json.dumps(out), flush=True)


def kernel_name(frames_per_launch, pixels):
    """The kernel a launch of the timed region runs: nb_render_frames of two or more frames on a
    target of at most 2^19 pixels renders one path per (frame, pixel) with the serving kernel
    (unless NB_SAMPLE_PARALLEL=0 or NB_SAMPLE_PARALLEL_PIXELS lowers the limit)."""
    if frames_per_launch >= 2 and os.environ.get("NB_SAMPLE_PARALLEL", "1") != "0":
        limit = int(os.environ.get("NB_SAMPLE_PARALLEL_PIXELS", str(1 << 19)))
        if pixels <= limit:
            return "serve_pt_kernel<16, 4> (sample-parallel frames)"
    return "render_pt_kernel" if os.environ.get("NB_RENDER", "pt") != "tile" else "render_kernel"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["large", "small"], default="large")
    ap.add_argument("--frames-per-launch", type=int, default=1,
                    help="frames per kernel launch (nb_render_frames); 1 = one nb_render per step")
    ap.add_argument("--reference-blas", action="store_true", help="rebuild the BLAS per instance as the reference")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--mode", choices=["render", "serve", "cache", "serve-cache"], default="render")
    ap.add_argument("--cache-warmup", type=int, default=40, help="serve-cache: training frames per scene cache")
    ap.add_argument("--serve-train-paths", type=int, default=2048,
                    help="serve-cache: training paths per sample of a cache-assisted request")
    ap.add_argument("--train-paths", type=int, default=1 << 16, help="cache: training paths per frame")
    ap.add_argument("--lr", type=float, default=1e-3, help="cache: Adam learning rate")
    ap.add_argument("--max-batch", type=int, default=0,
                    help="cache: training records per frame (0: all, three per training path)")
    ap.add_argument("--qps", default="20", help="serve: Poisson arrival rate(s), comma-separated for a sweep")
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--no-secondary", action="store_true",
                    help="render: skip the secondary workloads (small scene, serving stream, learned cache)")
    ap.add_argument("--max-paths", type=int, default=1 << 23)
    ap.add_argument("--count-in-timed", action="store_true", help="keep the kernel counters on in the timed pass")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by image tiles or Sobol rows, which this build does not run)", file=sys.stderr)
        sys.exit(2)
    if args.mode == "serve":
        serve_bench(args)
        return
    if args.mode == "cache":
        cache_bench(args)
        return
    if args.mode == "serve-cache":
        serve_cache_bench(args)
        return

    import nbounce_amd as nb
    from nbounce_amd import scenegen

    P, desc, scene, arrays, n_inst_tris = build_workload(args.workload, share_blas=not args.reference_blas)
    blas, tlas, inst = arrays
    ctx = nb.Context(0)
    sb = nb.SceneBuffers(ctx, blas, tlas, inst, scene.vertices, scene.indices)
    env_faces = scenegen.sky_cube(P["env_face"])
    env = nb.EnvMap(ctx, env_faces)
    cam = nb.CameraController(world_position=P["eye"], target=P["target"], aspect_ratio=P["width"] / P["height"])
    pt = nb.Pathtracer(ctx, sb, cam, env, P["width"], P["height"], resolution_factor=1.0)
    pt.globals.bounces = P["bounces"]
    pt.globals.contribution_factor = P["contribution_factor"]
    F = max(1, args.frames_per_launch)

    def step_block(n_steps):
        done = 0
        while done < n_steps:
            if F == 1:
                pt.dispatch()
                done += 1
            else:
                done += pt.dispatch_frames(min(F, n_steps - done))

    if args.calibrate_copy_mib:
        calibration_copy(args.calibrate_copy_mib)
    step_block(args.warmup)
    ctx.synchronize()
    # The counters (the reference's n_aabb / n_tri debug variables plus rays, instance entries,
    # hits) are taken over an untimed pass of the same frames; the timed pass renders the same
    # frames with counting off.  Paths depend only on (pixel, sample, scene): the counts match.
    pt.invalidate()
    ctx.reset_counters()
    step_block(args.steps)
    ctx.synchronize()
    c = ctx.counters()
    pt.invalidate()
    ctx.enable_counters(args.count_in_timed)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.timer_start()
    step_block(args.steps)
    dev_ms = ctx.timer_stop()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    # the same frames again in one nb_render_frames call (the image is the same bit for bit)
    pt.invalidate()
    ctx.synchronize()
    t1 = time.perf_counter()
    pt.dispatch_frames(args.steps)
    ctx.synchronize()
    wall_one_call = time.perf_counter() - t1
    ctx.enable_counters(True)
    launches = args.steps if F == 1 else math.ceil(args.steps / F)
    rays = c["rays"]
    value = rays / wall / 1e6
    kernel_ms = dev_ms / launches
    pipelined = F == 1 and os.environ.get("NB_PIPELINE", "1") != "0"
    bytes_per_launch = algorithmic_bytes(c, pipelined) / launches
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9

    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            key = "render_pt_kernel instanced 1920x1080 12 bounces shared BLAS" + (" pipelined" if pipelined else "")
            if args.workload == "large" and not args.reference_blas and F == 1:
                traffic = json.load(f).get(key, {}).get("hbm_bytes_per_launch")
    except OSError:
        pass

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(P, scene, arrays, env_faces, pt._camera_data, pt.table, target_s=args.cpu_seconds)

    # the HBM traffic the counters measured, as a fraction of peak at this kernel time
    hbm_frac = round(traffic / (kernel_ms * 1e-3) / (PEAK_HBM_GBS * 1e9), 4) if traffic else None

    secondary = None
    if not args.no_secondary and args.workload == "large":
        interior = view_rate(nb, ctx, sb, env, P, scenegen.LARGE_INTERIOR["eye"], scenegen.LARGE_INTERIOR["target"])
        moving = moving_camera_rate(nb, ctx, sb, env, P)
        ll_env = nb.EnvMap.latlong(ctx, scenegen.sky_latlong(1024, 512))
        latlong = view_rate(nb, ctx, sb, ll_env, P, P["eye"], P["target"])
        latlong["config"]["workload"] = (f"the default workload with the sky as a 1024x512 RGBA32F lat-long map "
                                         f"instead of the {P['env_face']}^2 cube, 32 frames of 1 spp")
        del pt, sb, env, ll_env
        ctx.close()
        secondary = {
            "interior_view": interior,
            "moving_camera": moving,
            "latlong_env": latlong,
            "plain_one_call": {
                "value": round(rays / wall_one_call / 1e6, 3), "unit": "Mrays/s",
                "ms_per_image": round(wall_one_call * 1e3, 3),
                "config": {"workload": f"the same {args.steps} frames of the same workload in one nb_render_frames "
                                       "call (the same image bit for bit)"}},
            "small": small_workload_rate(16),
            "small_as_written": small_workload_rate(16, native=True),
            "serving": dict(serve_streams([20.0], 100, args.max_paths)[0], config=SERVE_CONFIG),
        }
        c4 = cache_equal_error(1 << 16, 1e-3)
        secondary["learned_cache"] = {
            "value": c4["ratio"], "unit": "equal relative-MSE frame time, cache-assisted / plain (lower is better)",
            "with_training_ratio": c4["with_training_ratio"], "cache_1spp": c4["cache_1spp"],
            "plain_spp_at_equal_error": c4["plain_spp_at_equal_error"], "plain_ms_per_spp": c4["plain_ms_per_spp"],
            "config": c4["config"]}

    out = {
        "metric": "Mrays/s at 1080p" if args.workload == "large" else "Mrays/s at 256x256",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, seed 0)",
        "config": {
            "workload": f"{desc}, {n_inst_tris} instanced triangles, {P['width']}x{P['height']}, "
                        f"{args.steps} spp (1 spp per step), {P['bounces']} bounces",
            "blas": "per instance (reference)" if args.reference_blas else "shared per mesh",
            "frames_per_launch": F,
            "pipelined": pipelined,
            "rays": int(rays),
            "rays_per_sample": round(rays / max(1, c["samples"]), 4),
            "per_ray": {k: round(c[k] / max(1, rays), 3) for k in ("n_aabb", "n_tri", "n_inst", "hits")},
            "lds_stack_entries": int(os.environ.get("NB_LDS_STACK", "16")),
            "samples_per_s_M": round(c["samples"] / wall / 1e6, 3),
        },
        "roofline": {
            # the ceiling frac is priced against (no matrix work); the counters show HBM itself is
            # far from saturated: the limiter is the latency of the per-lane node fetches through
            # the texture-data path, with VALU issue beside it
            "bound": "hbm",
            "limiter": "dependent node fetches through the texture-data path (TD busy 0.79, half of it stalled "
                       "on the L1) and VALU issue (about 0.6 busy); HBM counter traffic is hbm_counter_frac of peak",
            "hbm_counter_frac": hbm_frac,
            "achieved": round(achieved, 2),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 5),
            "traffic": traffic,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (L2 to memory, Infinity Cache hits "
                              "included: an upper bound on HBM bytes), scaled by a 256 MiB copy "
                              "(profiles/render_pt_kernel_instanced_1080p_traffic.md)" if traffic else None,
            "kernel": kernel_name(F, P["width"] * P["height"]),
            "kernel_ms": round(kernel_ms, 4),
            "kernel_ms_is": ("the timed region / launches: pipelined path kernels of consecutive frames overlap, "
                             "so each launch's own duration is longer; the mix kernels are inside")
                            if pipelined else "the timed region / launches",
            "algorithmic_bytes_per_launch": int(bytes_per_launch),
        },
        "cpu_baseline": cpu,
    }
    if secondary is not None:
        out["secondary"] = secondary
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
