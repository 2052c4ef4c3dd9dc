"""GPU tests of the radiance-cache MLP kernels (BF16 or FP16 MFMA, FP32 accumulation) against the
float64 oracle (oracle/mlp.py) that rounds the same values to the same 16-bit format (FP16 with
the cache's loss scale)."""
import ctypes

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu

BOX_MIN, BOX_MAX = np.array([0, 0, 0], np.float32), np.array([10, 10, 10], np.float32)


def _batch(rng, n):
    raw = np.zeros((n, 16), dtype=np.float32)
    raw[:, 0:3] = rng.uniform(0, 10, size=(n, 3))
    d = rng.normal(size=(n, 3))
    raw[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nrm = rng.normal(size=(n, 3))
    raw[:, 6:9] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    raw[:, 9] = rng.uniform(0.05, 1, n)
    raw[:, 10:13] = rng.uniform(0, 1, size=(n, 3))
    raw[:, 13] = rng.uniform(0, 1, n) < 0.3
    target = rng.uniform(0, 3, size=(n, 3)).astype(np.float32)
    return raw, target


@pytest.fixture(scope="module", params=["bf16", "fp16"])
def cache(request):
    import nbounce_amd as nb
    from nbounce_amd._lib import check, lib, ptr
    ctx = nb.Context(0)
    h = ctypes.c_void_p()
    flags = 1 if request.param == "fp16" else 0
    check(lib.nb_cache_create_ex(ctx.handle, ptr(BOX_MIN), ptr(BOX_MAX), 7, 1 << 16, flags, ctypes.byref(h)))
    yield lib, ptr, check, h, request.param
    lib.nb_cache_destroy(h)
    ctx.close()


def _fmt(cache):
    return cache[4]


def _round(cache, x):
    return mlp.fp16(x) if _fmt(cache) == "fp16" else mlp.bf16(x)


def _encode(cache, raw):
    lib, ptr, check, h = cache[:4]
    out = np.zeros((len(raw), 96), dtype=np.uint16)
    check(lib.nb_cache_encode(h, ptr(raw), len(raw), ptr(out)))
    if _fmt(cache) == "fp16":
        return out.view(np.float16).astype(np.float64)
    return (out.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _params(cache):
    lib, ptr, check, h = cache[:4]
    p = np.zeros(lib.nb_cache_param_count(), dtype=np.float32)
    check(lib.nb_cache_get_params(h, ptr(p)))
    return p


def test_encoding_matches_oracle(cache):
    print("case: MLP input encoding, 4096 records", flush=True)
    raw, _ = _batch(np.random.default_rng(0), 4096)
    got = _encode(cache, raw)
    want = _round(cache, mlp.encode(raw, BOX_MIN, BOX_MAX))
    ulp = np.abs(want) * (2.0 ** -10 if _fmt(cache) == "fp16" else 2.0 ** -7) + 1e-30
    assert np.mean(got == want) > 0.995
    assert np.all(np.abs(got - want) <= ulp + 1e-7)


def test_fused_inference_matches_oracle(cache):
    lib, ptr, check, h = cache[:4]
    print("case: fused MFMA inference, 5000 rows", flush=True)
    raw, _ = _batch(np.random.default_rng(1), 5000)
    params = mlp.init_params(3)
    params[np.arange(len(params)) % 7 == 0] += 0.05      # nonzero biases everywhere
    check(lib.nb_cache_set_params(h, ptr(params)))
    got = np.zeros((5000, 3), dtype=np.float32)
    check(lib.nb_cache_infer(h, ptr(raw), 5000, ptr(got)))
    x = _encode(cache, raw)
    want, _, _ = mlp.forward(params.astype(np.float64), x, fmt=_fmt(cache))
    rel = np.abs(got - want) / (np.abs(want) + 1e-2)
    assert np.median(rel) < 1e-4 and rel.max() < 2e-2


def test_gradient_matches_oracle(cache):
    lib, ptr, check, h = cache[:4]
    print("case: MFMA forward + backward gradient, 3000 rows", flush=True)
    raw, target = _batch(np.random.default_rng(2), 3000)
    params = mlp.init_params(4)
    params[np.arange(len(params)) % 5 == 0] += 0.03
    check(lib.nb_cache_set_params(h, ptr(params)))
    loss = ctypes.c_float()
    grad = np.zeros(len(params), dtype=np.float32)
    check(lib.nb_cache_gradient(h, ptr(raw), ptr(target), 3000, ctypes.byref(loss), ptr(grad)))
    x = _encode(cache, raw)
    scale = ctypes.c_float()
    check(lib.nb_cache_loss_scale(h, ctypes.byref(scale)))
    want_loss, want_grad, _ = mlp.loss_and_grad(params.astype(np.float64), x, target.astype(np.float64),
                                                fmt=_fmt(cache), scale=scale.value)
    assert abs(loss.value - want_loss) <= 1e-4 * want_loss
    off = 0
    for L, (i, o) in enumerate(mlp.DIMS):
        for size in (i * o, o):
            g, w = grad[off:off + size].astype(np.float64), want_grad[off:off + size]
            off += size
            err = np.linalg.norm(g - w) / (np.linalg.norm(w) + 1e-20)
            assert err < 1e-2, (L, size, err)


def test_training_reduces_the_loss(cache):
    lib, ptr, check, h = cache[:4]
    print("case: Adam training on a smooth radiance field", flush=True)
    rng = np.random.default_rng(5)
    raw, _ = _batch(rng, 1 << 15)
    p = raw[:, 0:3] / 10.0
    target = np.stack([0.5 + 0.4 * np.sin(3 * p[:, 0]) * np.cos(2 * p[:, 2]),
                       0.3 + 0.2 * p[:, 1],
                       0.8 * raw[:, 10]], 1).astype(np.float32)
    p0 = mlp.init_params(6)
    check(lib.nb_cache_set_params(h, ptr(p0)))
    l0 = ctypes.c_float()
    g = np.zeros(lib.nb_cache_param_count(), dtype=np.float32)
    check(lib.nb_cache_gradient(h, ptr(raw[:4096]), ptr(target[:4096]), 4096, ctypes.byref(l0), ptr(g)))
    last = ctypes.c_float()
    for _ in range(8):
        check(lib.nb_cache_train(h, ptr(raw), ptr(target), len(raw), 1024, 3e-3, ctypes.byref(last)))
    l1 = ctypes.c_float()
    check(lib.nb_cache_gradient(h, ptr(raw[:4096]), ptr(target[:4096]), 4096, ctypes.byref(l1), ptr(g)))
    scale = ctypes.c_float()
    check(lib.nb_cache_loss_scale(h, ctypes.byref(scale)))
    print(f"loss {l0.value:.4f} -> {l1.value:.5f}, loss scale {scale.value}")
    assert l1.value < 0.05 * l0.value
    assert 1.0 <= scale.value <= 65536.0


def test_fp16_overflow_skips_the_step_and_halves_the_scale(cache):
    """A training step whose loss-scaled gradient is not finite in FP16 leaves every parameter
    as it was (the Adam step is skipped on the device) and halves the loss scale.  BF16 has the
    FP32 exponent range: the same batch trains, and its loss scale stays 1."""
    lib, ptr, check, h = cache[:4]
    print(f"case: {_fmt(cache)} training step with a gradient beyond the FP16 range", flush=True)
    raw, _ = _batch(np.random.default_rng(9), 1024)
    target = np.full((1024, 3), 1e30, dtype=np.float32)   # (y - t) / (y^2 + 0.01) overflows FP16
    s0, s1 = ctypes.c_float(), ctypes.c_float()
    check(lib.nb_cache_loss_scale(h, ctypes.byref(s0)))
    p0 = _params(cache)
    loss = ctypes.c_float()
    check(lib.nb_cache_train(h, ptr(raw), ptr(target), 1024, 1024, 1e-3, ctypes.byref(loss)))
    p1 = _params(cache)
    check(lib.nb_cache_loss_scale(h, ctypes.byref(s1)))
    print(f"loss scale {s0.value} -> {s1.value}, loss {loss.value}")
    if _fmt(cache) == "bf16":
        assert s0.value == s1.value == 1.0
        assert np.isfinite(p1).all() and not np.array_equal(p1, p0)
        return
    np.testing.assert_array_equal(p1.view(np.uint32), p0.view(np.uint32))
    assert s1.value == max(s0.value * 0.5, 1.0)
    # the next finite step trains again
    raw2, target2 = _batch(np.random.default_rng(10), 1024)
    check(lib.nb_cache_train(h, ptr(raw2), ptr(target2), 1024, 1024, 1e-3, ctypes.byref(loss)))
    assert np.isfinite(loss.value)
    assert not np.array_equal(_params(cache), p0)


@pytest.mark.parametrize("fp16", [False, True])
def test_trained_cache_beats_the_plain_tracer_at_one_sample(fp16):
    """What the trained network gains: at one sample per pixel the cache-assisted image is
    closer to a 1024-spp reference (plain tracer) than the plain 1-spp image, and than the same
    method with an untrained network (which predicts near zero, so its error sits near the relative
    error of a black image, about 1)."""
    import math
    import nbounce_amd as nb
    from nbounce_amd import scenegen
    from nbounce_amd.cache import RadianceCache, rel_mse
    from tests import _workloads as Wl
    print(f"case: cache-assisted vs plain at 1 spp, 192x112 spheres scene, {'fp16' if fp16 else 'bf16'}", flush=True)
    w_, h_ = 192, 112
    ctx = nb.Context(0)
    wl = Wl.small()
    sb = nb.SceneBuffers(ctx, wl["blas"], wl["tlas"], wl["instances"], wl["vertices"], wl["indices"])
    env = nb.EnvMap(ctx, scenegen.sky_cube(64))
    eye, target = scenegen.orbit_pose(0)
    cam = nb.CameraController(world_position=eye, target=target, aspect_ratio=w_ / h_)

    def tracer():
        return nb.Pathtracer(ctx, sb, cam, env, w_, h_, resolution_factor=1.0)

    ref_pt = tracer()
    ref_pt.dispatch_frames(1023)
    ref = ref_pt.output_texture()
    plain = []
    for k in range(4):
        pt = tracer()
        pt.globals.sample = 100 + 10 * k
        pt.dispatch(weight=1.0)
        img = pt.output_texture()
        plain.append(rel_mse(img, ref))

    def cache_err(cache):
        errs = []
        for k in range(4):
            pt = tracer()
            pt.globals.sample = 300 + 10 * k
            cache.render(pt, train_paths=0, weight=1.0)
            errs.append(rel_mse(pt.output_texture(), ref))
        return float(np.mean(errs))

    untrained = RadianceCache(ctx, [0, 0, 0], [10, 10, 10], seed=1, fp16=fp16)
    e_untrained = cache_err(untrained)
    cache = RadianceCache(ctx, [0, 0, 0], [10, 10, 10], seed=1, fp16=fp16)
    train_pt = tracer()
    losses = []
    for f in range(160):
        if train_pt.sample_count() >= 1000:
            train_pt.invalidate()
        losses.append(cache.render(train_pt, train_paths=8192, lr=2e-3))
    e_cache = cache_err(cache)
    e_plain = float(np.mean(plain))
    print(f"relMSE plain 1spp {e_plain:.4f}, cache 1spp {e_cache:.4f}, untrained cache {e_untrained:.4f}, "
          f"loss {losses[0]:.3f} -> {np.mean(losses[-10:]):.3f}")
    assert math.isfinite(e_cache)
    assert e_cache < 0.7 * e_plain
    assert e_cache < 0.8 * e_untrained
    cache.close(); untrained.close()


def test_training_records_match_oracle():
    """The training records of a cache-assisted frame (nb_cache_training_records: the first three
    shading points of each training path, target = the path's radiance / the throughput that
    arrived there) equal the CPU oracle's restatement of the same paths (orc_train_path), bit for
    bit, as multisets (the device writes them in no particular order)."""
    import nbounce_amd as nb
    import oracle
    from nbounce_amd import scenegen
    from nbounce_amd._lib import check, lib, ptr
    from tests import _workloads as Wl
    print("case: 512 training paths of frame 5, 192x112 spheres scene, 8 bounces", flush=True)
    w_, h_, n_paths, sample, bounces, cf = 192, 112, 512, 5, 8, 4.0
    ctx = nb.Context(0)
    wl = Wl.small()
    sb = nb.SceneBuffers(ctx, wl["blas"], wl["tlas"], wl["instances"], wl["vertices"], wl["indices"])
    faces = scenegen.sky_cube(64)
    env = nb.EnvMap(ctx, faces)
    table = Wl.table()
    check(lib.nb_sobol_set(ctx.handle, ptr(np.ascontiguousarray(table)), table.shape[0], table.shape[1], 17))
    eye, target = scenegen.orbit_pose(0)
    cam = nb.CameraController(world_position=eye, target=target, aspect_ratio=w_ / h_).data()
    h = ctypes.c_void_p()
    check(lib.nb_cache_create_ex(ctx.handle, ptr(BOX_MIN), ptr(BOX_MAX), 3, 1 << 16, 0, ctypes.byref(h)))
    g = nb.Globals(sample, 1.0, bounces, cf).array()
    raw = np.zeros((3 * n_paths, 16), dtype=np.float32)
    tgt = np.zeros((3 * n_paths, 3), dtype=np.float32)
    n = ctypes.c_uint32()
    check(lib.nb_cache_training_records(ctx.handle, sb.handle, w_, h_, ptr(cam), ptr(g), env.handle, h, n_paths,
                                        ptr(raw), ptr(tgt), 3 * n_paths, ctypes.byref(n)))
    got = np.concatenate([raw[:n.value], tgt[:n.value]], axis=1)
    # the oracle's paths: the pixels and Sobol row the header states
    npix = w_ * h_
    stride = max(npix // n_paths, 1)
    offset = (sample * 2654435761) % npix
    row = 1 + (sample + 511) % 1023
    osc, oenv = Wl.oracle_scene(wl), oracle.Env(faces)
    og = oracle.globals_(row, 1.0, bounces, cf)
    want = []
    for i in range(n_paths):
        pix = (i * stride + offset) % npix
        r, t, _ = oracle.train_path(osc, cam, og, table, 17, oenv, w_, h_, pix % w_, pix // w_)
        if len(r):
            want.append(np.concatenate([r, t], axis=1))
    want = np.concatenate(want)

    def canon(a):
        u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
        return u[np.lexsort(u.T[::-1])]

    print(f"records: gpu {len(got)}, oracle {len(want)}", flush=True)
    assert len(got) == len(want) > n_paths
    np.testing.assert_array_equal(canon(got), canon(want))
    lib.nb_cache_destroy(h)
    ctx.close()


def test_query_records_match_oracle():
    """The queries of a cache-assisted frame (nb_cache_query_records: each pixel's path up to its
    query point, the radiance met before it, the point's record and the throughput arriving
    there) equal the CPU oracle's restatement of the same paths (orc_cache_query), pixel by pixel
    and bit for bit: which pixels query, their records and throughputs, and every pixel's direct
    radiance."""
    import nbounce_amd as nb
    import oracle
    from nbounce_amd import scenegen
    from nbounce_amd._lib import check, lib, ptr
    from tests import _workloads as Wl
    print("case: the queries of frame 3, 96x64 spheres scene, 8 bounces", flush=True)
    w_, h_, sample, bounces, cf = 96, 64, 3, 8, 4.0
    ctx = nb.Context(0)
    wl = Wl.small()
    sb = nb.SceneBuffers(ctx, wl["blas"], wl["tlas"], wl["instances"], wl["vertices"], wl["indices"])
    faces = scenegen.sky_cube(64)
    env = nb.EnvMap(ctx, faces)
    table = Wl.table()
    check(lib.nb_sobol_set(ctx.handle, ptr(np.ascontiguousarray(table)), table.shape[0], table.shape[1], 17))
    eye, target = scenegen.orbit_pose(0)
    cam = nb.CameraController(world_position=eye, target=target, aspect_ratio=w_ / h_).data()
    h = ctypes.c_void_p()
    check(lib.nb_cache_create_ex(ctx.handle, ptr(BOX_MIN), ptr(BOX_MAX), 3, 1 << 16, 0, ctypes.byref(h)))
    g = nb.Globals(sample, 1.0, bounces, cf).array()
    npix = w_ * h_
    direct = np.zeros((h_, w_, 4), dtype=np.float32)
    raw = np.zeros((npix, 16), dtype=np.float32)
    thr = np.zeros((npix, 4), dtype=np.float32)
    n = ctypes.c_uint32()
    check(lib.nb_cache_query_records(ctx.handle, sb.handle, w_, h_, ptr(cam), ptr(g), env.handle, h, ptr(direct),
                                     ptr(raw), ptr(thr), npix, ctypes.byref(n)))
    pix = thr[:n.value, 3].view(np.uint32)
    assert len(set(pix.tolist())) == n.value
    got_raw = {int(p): raw[i] for i, p in enumerate(pix)}
    got_thr = {int(p): thr[i, :3] for i, p in enumerate(pix)}
    osc, oenv = Wl.oracle_scene(wl), oracle.Env(faces)
    og = oracle.globals_(sample, 1.0, bounces, cf)
    queried = 0
    for y in range(h_):
        for x in range(w_):
            p = y * w_ + x
            q, d, r, t = oracle.cache_query(osc, cam, og, table, 17, oenv, w_, h_, x, y)
            assert (p in got_raw) == q, (x, y)
            np.testing.assert_array_equal(direct[y, x, :3].view(np.uint32), d.view(np.uint32))
            if q:
                queried += 1
                np.testing.assert_array_equal(got_raw[p].view(np.uint32), r.view(np.uint32))
                np.testing.assert_array_equal(got_thr[p].view(np.uint32), t.view(np.uint32))
    assert (direct[..., 3] == 1.0).all()
    print(f"queries: {queried} of {npix} pixels", flush=True)
    assert queried == n.value and 0 < queried < npix
    lib.nb_cache_destroy(h)
    ctx.close()


def test_cached_frame_is_built_from_its_queries():
    """A cache-assisted frame's image (nb_render_cached, inference only, weight 1) is exactly
    direct + throughput * max(prediction, 0) of its queries: the query records
    (nb_cache_query_records, pinned to the oracle above) run through the network
    (nb_cache_infer, pinned to the oracle above) and combined in binary32 in the kernel's order."""
    import nbounce_amd as nb
    from nbounce_amd import scenegen
    from nbounce_amd._lib import check, lib, ptr
    from tests import _workloads as Wl
    print("case: a cache-assisted frame rebuilt from its queries, 96x64 spheres scene", flush=True)
    w_, h_, sample, bounces, cf = 96, 64, 4, 8, 4.0
    ctx = nb.Context(0)
    wl = Wl.small()
    sb = nb.SceneBuffers(ctx, wl["blas"], wl["tlas"], wl["instances"], wl["vertices"], wl["indices"])
    env = nb.EnvMap(ctx, scenegen.sky_cube(64))
    table = Wl.table()
    check(lib.nb_sobol_set(ctx.handle, ptr(np.ascontiguousarray(table)), table.shape[0], table.shape[1], 17))
    eye, target = scenegen.orbit_pose(7)
    cam = nb.CameraController(world_position=eye, target=target, aspect_ratio=w_ / h_).data()
    h = ctypes.c_void_p()
    check(lib.nb_cache_create_ex(ctx.handle, ptr(BOX_MIN), ptr(BOX_MAX), 3, 1 << 16, 0, ctypes.byref(h)))
    params = mlp.init_params(11)
    params[np.arange(len(params)) % 7 == 0] += 0.05      # nonzero predictions
    check(lib.nb_cache_set_params(h, ptr(params)))
    g = nb.Globals(sample, 1.0, bounces, cf).array()
    t = ctypes.c_void_p()
    check(lib.nb_target_create(ctx.handle, w_, h_, ctypes.byref(t)))
    check(lib.nb_render_cached(ctx.handle, sb.handle, t, ptr(cam), ptr(g), env.handle, h, 0, 0.0, None))
    img = np.zeros((h_, w_, 4), dtype=np.float32)
    check(lib.nb_target_read(t, ptr(img)))
    npix = w_ * h_
    direct = np.zeros((h_, w_, 4), dtype=np.float32)
    raw = np.zeros((npix, 16), dtype=np.float32)
    thr = np.zeros((npix, 4), dtype=np.float32)
    n = ctypes.c_uint32()
    check(lib.nb_cache_query_records(ctx.handle, sb.handle, w_, h_, ptr(cam), ptr(g), env.handle, h, ptr(direct),
                                     ptr(raw), ptr(thr), npix, ctypes.byref(n)))
    k = n.value
    rgb = np.zeros((k, 3), dtype=np.float32)
    check(lib.nb_cache_infer(h, ptr(np.ascontiguousarray(raw[:k])), k, ptr(rgb)))
    want = direct.reshape(npix, 4)[:, :3].copy()
    pix = thr[:k, 3].view(np.uint32)
    want[pix] = want[pix] + thr[:k, :3] * np.maximum(rgb, np.float32(0.0))
    assert (rgb > 0).mean() > 0.2 and k > 0.2 * npix
    np.testing.assert_array_equal(img.reshape(npix, 4)[:, :3].view(np.uint32), want.view(np.uint32))
    lib.nb_target_destroy(t)
    lib.nb_cache_destroy(h)
    ctx.close()


def test_adam_steps_match_oracle():
    """Three Adam steps of a BF16 cache (nb_cache_train with one batch per call, so one step
    each): every parameter's update equals the oracle's adam_step (oracle/mlp.py, pinned to
    torch.optim.Adam) applied to the GPU's own gradient at the same parameters
    (nb_cache_gradient).  float32 arithmetic and the gradient atomics' summation order are the
    differences, hence a relative 1e-3 on the update (plus two units in the last place of the
    float32 parameter it is read back from), where |g| > 1e-6."""
    import nbounce_amd as nb
    from nbounce_amd._lib import check, lib, ptr
    print("case: 3 Adam steps of a BF16 cache on 4096 records", flush=True)
    ctx = nb.Context(0)
    h = ctypes.c_void_p()
    check(lib.nb_cache_create_ex(ctx.handle, ptr(BOX_MIN), ptr(BOX_MAX), 5, 1 << 16, 0, ctypes.byref(h)))
    raw, target = _batch(np.random.default_rng(21), 4096)
    params = mlp.init_params(13)
    params[np.arange(len(params)) % 5 == 0] += 0.02
    check(lib.nb_cache_set_params(h, ptr(params)))
    m = np.zeros(len(params))
    v = np.zeros(len(params))
    loss = ctypes.c_float()
    for t in range(1, 4):
        before = np.zeros(lib.nb_cache_param_count(), dtype=np.float32)
        check(lib.nb_cache_get_params(h, ptr(before)))
        g = np.zeros(len(params), dtype=np.float32)
        check(lib.nb_cache_gradient(h, ptr(raw), ptr(target), 4096, ctypes.byref(loss), ptr(g)))
        check(lib.nb_cache_train(h, ptr(raw), ptr(target), 4096, 4096, 1e-3, ctypes.byref(loss)))
        after = np.zeros(len(params), dtype=np.float32)
        check(lib.nb_cache_get_params(h, ptr(after)))
        w, m, v = mlp.adam_step(before, m, v, g, t, 1e-3)
        mask = np.abs(g) > 1e-6
        upd_gpu = after.astype(np.float64) - before
        upd_or = w - before
        print(f"step {t}: {mask.sum()} parameters compared, max |update| {np.abs(upd_gpu).max():.3e}", flush=True)
        assert mask.mean() > 0.5
        # the updates are read back as differences of float32 parameters: two units in the last
        # place of the parameter on top of the relative bound
        tol = 1e-3 * np.abs(upd_or) + 2.0 * np.spacing(np.abs(before).astype(np.float32)).astype(np.float64)
        bad = np.abs(upd_gpu - upd_or) > tol
        assert not bad[mask].any(), (np.flatnonzero(mask & bad)[:5], upd_gpu[mask & bad][:5], upd_or[mask & bad][:5])
    lib.nb_cache_destroy(h)
    ctx.close()
