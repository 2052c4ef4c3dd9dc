"""GPU parity: libnbounce (through the C ABI) against the CPU oracle on the same inputs."""
import numpy as np
import pytest

import oracle
from tests import _workloads as W

pytestmark = pytest.mark.gpu

REL_TOL = 1e-3   # per-pixel relative error bound for images at matched spp and Sobol rows


@pytest.fixture(scope="module")
def nb():
    import nbounce_amd
    return nbounce_amd


@pytest.fixture(scope="module")
def ctx(nb):
    c = nb.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def small_gpu(nb, ctx):
    w = W.small()
    sb = nb.SceneBuffers(ctx, w["blas"], w["tlas"], w["instances"], w["vertices"], w["indices"])
    env = nb.EnvMap(ctx, __import__("nbounce_amd.scenegen", fromlist=["x"]).sky_cube(128))
    return w, sb, env


def _pathtracer(nb, ctx, sb, env, size, eye, target, bounces=8):
    cam = nb.CameraController(world_position=eye, target=target, aspect_ratio=size[0] / size[1])
    pt = nb.Pathtracer(ctx, sb, cam, env, size[0], size[1], resolution_factor=1.0, table=W.table(), sobol_stride=17)
    pt.globals.bounces = bounces
    return pt, cam


def test_primary_hits_identical(nb, ctx, small_gpu):
    w, sb, env = small_gpu
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    pt, cam = _pathtracer(nb, ctx, sb, env, (128, 128), P["eye"], P["target"])
    print("case: small scene primary hits 128x128 sample 1", flush=True)
    got = pt.primary_hits(sample=1)
    ref = oracle.primary_hits(W.oracle_scene(w), cam.data(), 128, 128, W.table(), 17, 1)
    assert (got[..., 0] != 0xFFFFFFFF).mean() > 0.5
    np.testing.assert_array_equal(got, ref)


def test_one_frame_matches_oracle(nb, ctx, small_gpu):
    w, sb, env = small_gpu
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    pt, cam = _pathtracer(nb, ctx, sb, env, (128, 128), P["eye"], P["target"])
    print("case: small scene one frame 128x128", flush=True)
    ctx.reset_counters()
    pt.dispatch()
    got = pt.output_texture()
    g = oracle.globals_(1, 1.0, 8, 4.0)
    ref, cnt = oracle.render(W.oracle_scene(w), cam.data(), g, W.table(), 17, oracle.Env(env.faces), 128, 128,
                             threads=8)
    err = W.rel_err(got, ref)
    same = (got.view(np.uint32) == ref.view(np.uint32)).all(axis=-1).mean()
    print(f"bit-identical pixels {same:.5f}, max rel err {err.max():.3e}")
    assert err.max() <= REL_TOL
    c = ctx.counters()
    assert c["rays"] == cnt["rays"]
    assert c["samples"] == 128 * 128


def test_frames_equal_sequential_dispatches(nb, ctx, small_gpu):
    w, sb, env = small_gpu
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    a, _ = _pathtracer(nb, ctx, sb, env, (64, 64), P["eye"], P["target"])
    b, _ = _pathtracer(nb, ctx, sb, env, (64, 64), P["eye"], P["target"])
    print("case: 4 frames in one launch vs 4 dispatches", flush=True)
    for _ in range(4):
        a.dispatch()
    b.dispatch_frames(4)
    np.testing.assert_array_equal(a.output_texture().view(np.uint32), b.output_texture().view(np.uint32))
    assert a.sample_count() == b.sample_count() == 4


@pytest.mark.parametrize("size,frames", [((256, 256), 16), ((512, 512), 20), ((8, 8), 3)])
def test_sample_parallel_frames_equal_sequential_dispatches(nb, ctx, small_gpu, size, frames):
    """Small targets render nb_render_frames one path per (frame, pixel) over the whole GPU (the
    serving kernels, ordered accumulation): the image and the ray counts equal `frames`
    sequential dispatches.  512x512 x 20 frames exceeds one launch's result slots (2^22 paths),
    so it also crosses a launch boundary; frames start at sample 3 on a non-empty target."""
    w, sb, env = small_gpu
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    print(f"case: {size[0]}x{size[1]}, frames 3..{frames + 2} in one call vs dispatches", flush=True)
    a, _ = _pathtracer(nb, ctx, sb, env, size, P["eye"], P["target"])
    b, _ = _pathtracer(nb, ctx, sb, env, size, P["eye"], P["target"])
    for p in (a, b):
        p.dispatch()
        p.dispatch()
    ctx.reset_counters()
    for _ in range(frames):
        a.dispatch()
    ca = ctx.counters()
    ctx.reset_counters()
    b.dispatch_frames(frames)
    cb = ctx.counters()
    np.testing.assert_array_equal(a.output_texture().view(np.uint32), b.output_texture().view(np.uint32))
    assert a.sample_count() == b.sample_count() == frames + 2
    for k in ("rays", "n_aabb", "n_tri", "n_inst", "samples", "hits"):
        assert ca[k] == cb[k], k
    assert cb["samples"] == size[0] * size[1] * frames


def test_progressive_frames_match_oracle(nb, ctx, small_gpu):
    w, sb, env = small_gpu
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    pt, cam = _pathtracer(nb, ctx, sb, env, (64, 64), P["eye"], P["target"])
    print("case: 8 progressive frames 64x64", flush=True)
    pt.dispatch_frames(8)
    got = pt.output_texture()
    ref = np.zeros((64, 64, 4), dtype=np.float32)
    osc, oenv = W.oracle_scene(w), oracle.Env(env.faces)
    for k in range(1, 9):
        g = oracle.globals_(k, np.float32(1.0) / np.float32(k), 8, 4.0)
        ref, _ = oracle.render(osc, cam.data(), g, W.table(), 17, oenv, 64, 64, rgba=ref, threads=8)
    assert W.rel_err(got, ref).max() <= REL_TOL


@pytest.fixture(scope="module")
def large_gpu(nb, ctx):
    """The instanced-mesh workload (4096 rotated, scaled instances of 16 meshes), built twice:
    once as the reference builds it (one BLAS per instance) and once with shared BLASes."""
    from nbounce_amd import scenegen
    w = W.large()
    ref_scene = nb.Scene(w["vertices"], w["indices_in"].copy(), w["primitives"])
    rb, rt, ri = nb.build_scene(ref_scene)
    shr_scene = nb.Scene(w["vertices"], w["indices_in"].copy(), w["primitives"])
    sb_, st_, si_ = nb.build_scene(shr_scene, share_blas=True)
    ref = nb.SceneBuffers(ctx, rb, rt, ri, ref_scene.vertices, ref_scene.indices)
    shr = nb.SceneBuffers(ctx, sb_, st_, si_, shr_scene.vertices, shr_scene.indices)
    env = nb.EnvMap(ctx, scenegen.sky_cube(64))
    arrays = dict(ref=(rb, rt, ri, ref_scene.indices), shared=(sb_, st_, si_, shr_scene.indices))
    return w, ref, shr, env, arrays


def _large_camera(nb, w_px, h_px):
    from nbounce_amd import scenegen
    P = scenegen.LARGE
    return nb.CameraController(world_position=P["eye"], target=P["target"], aspect_ratio=w_px / h_px)


def test_large_scene_primary_hits_identical(nb, ctx, large_gpu):
    w, ref, shr, env, arrays = large_gpu
    cam = _large_camera(nb, 128, 72)
    print("case: instanced scene primary hits 128x72, reference BLAS layout", flush=True)
    pt = nb.Pathtracer(ctx, ref, cam, env, 128, 72, resolution_factor=1.0, table=W.table(), sobol_stride=17)
    got = pt.primary_hits(sample=1)
    rb, rt, ri, idx = arrays["ref"]
    osc = oracle.Scene(rb, rt, ri, w["vertices"], idx)
    want = oracle.primary_hits(osc, cam.data(), 128, 72, W.table(), 17, 1)
    assert (want[..., 0] != 0xFFFFFFFF).mean() > 0.2
    np.testing.assert_array_equal(got, want)


def test_large_scene_frames_match_oracle_at_12_bounces(nb, ctx, large_gpu):
    """12 bounces with the reference's fixed 17-entry Sobol row stride (bounces past 8 read
    the next row's entries, pathtracing.wgsl:2-4,151-153)."""
    w, ref, shr, env, arrays = large_gpu
    cam = _large_camera(nb, 128, 72)
    print("case: instanced scene 4 frames 128x72, 12 bounces", flush=True)
    pt = nb.Pathtracer(ctx, ref, cam, env, 128, 72, resolution_factor=1.0, table=W.table(), sobol_stride=17)
    pt.globals.bounces = 12
    pt.dispatch_frames(4)
    got = pt.output_texture()
    rb, rt, ri, idx = arrays["ref"]
    osc, oenv = oracle.Scene(rb, rt, ri, w["vertices"], idx), oracle.Env(env.faces)
    img = np.zeros((72, 128, 4), dtype=np.float32)
    for k in range(1, 5):
        img, _ = oracle.render(osc, cam.data(), oracle.globals_(k, np.float32(1) / np.float32(k), 12, 4.0), W.table(),
                               17, oenv, 128, 72, rgba=img, threads=8)
    assert W.rel_err(got, img).max() <= REL_TOL


def test_shared_blas_layout_matches_oracle_and_reference_layout(nb, ctx, large_gpu):
    """Sharing one BLAS per mesh is a layout choice: the GPU image on the shared arrays equals the
    oracle's on the same arrays, and on this view every pixel equals the per-instance layout the
    reference builds.  (Pixels could differ only where the reference's own rebuilds leave the
    first instance of a mesh with a tree whose leaves index re-permuted triangles,
    scene.rs:285-305; this view meets none.)"""
    w, ref, shr, env, arrays = large_gpu
    cam = _large_camera(nb, 128, 72)
    print("case: shared vs per-instance BLAS 128x72, 4 frames", flush=True)
    a = nb.Pathtracer(ctx, ref, cam, env, 128, 72, resolution_factor=1.0, table=W.table(), sobol_stride=17)
    b = nb.Pathtracer(ctx, shr, cam, env, 128, 72, resolution_factor=1.0, table=W.table(), sobol_stride=17)
    for p in (a, b):
        p.globals.bounces = 12
        p.dispatch_frames(4)
    ia, ib = a.output_texture(), b.output_texture()
    sb_, st_, si_, idx = arrays["shared"]
    osc, oenv = oracle.Scene(sb_, st_, si_, w["vertices"], idx), oracle.Env(env.faces)
    img = np.zeros((72, 128, 4), dtype=np.float32)
    for k in range(1, 5):
        img, _ = oracle.render(osc, cam.data(), oracle.globals_(k, np.float32(1) / np.float32(k), 12, 4.0), W.table(),
                               17, oenv, 128, 72, rgba=img, threads=8)
    assert W.rel_err(ib, img).max() <= REL_TOL
    same = (ia.view(np.uint32) == ib.view(np.uint32)).all(-1).mean()
    print(f"shared vs per-instance layout: bit-identical pixels {same:.6f}")
    np.testing.assert_array_equal(ia.view(np.uint32), ib.view(np.uint32))


def test_render_timing_is_opt_in(nb, ctx, small_gpu):
    """Per-render HIP events are recorded only after nb_render_timing(ctx, 1); before that
    nb_last_render_ms reports that no render was timed."""
    w, sb, env = small_gpu
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    print("case: per-render timing switched off, then on", flush=True)
    pt, _ = _pathtracer(nb, ctx, sb, env, (64, 64), P["eye"], P["target"])
    ctx.enable_render_timing(False)
    pt.dispatch()
    pt.output_texture()
    with pytest.raises(nb.NbError):
        ctx.last_render_ms()
    ctx.enable_render_timing(True)
    pt.dispatch()
    assert ctx.last_render_ms() > 0.0
    ctx.enable_render_timing(False)


def test_pathtracer_surface_caps_and_invalidates(nb, ctx, small_gpu):
    """Pathtracer::dispatch stops at max_sample_count (pathtracer.rs:226) and invalidate()
    restarts accumulation at sample 1 (pathtracer.rs:221-223), overwriting the old image."""
    w, sb, env = small_gpu
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    pt, cam = _pathtracer(nb, ctx, sb, env, (32, 32), P["eye"], P["target"])
    pt.max_sample_count = 3
    for _ in range(5):
        pt.dispatch()
    assert pt.sample_count() == 3
    three = pt.output_texture()
    pt.invalidate()
    pt.dispatch()
    one = pt.output_texture()
    fresh, _ = _pathtracer(nb, ctx, sb, env, (32, 32), P["eye"], P["target"])
    fresh.dispatch()
    np.testing.assert_array_equal(one.view(np.uint32), fresh.output_texture().view(np.uint32))
    assert not np.array_equal(one, three)
    pt.resize(64, 48)
    assert pt.size == (64, 48)


def test_unit_order_across_camera_moves_changes_nothing(nb, ctx, small_gpu):
    """Cost-ordered work units (costs halved on a camera move, the order kept, rebuilt after two
    frames of a view) render, frame by frame, what band order renders."""
    import os
    w, sb, env = small_gpu
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    eyes = [np.asarray(P["eye"], dtype=np.float64) + np.array([0.3 * k, 0.0, -0.2 * k]) for k in range(3)]
    print("case: 256x256, 7 frames over 3 camera positions, cost order vs band order", flush=True)

    def run(order):
        old = os.environ.get("NB_UNIT_ORDER")
        os.environ["NB_UNIT_ORDER"] = order
        try:
            pt, cam = _pathtracer(nb, ctx, sb, env, (256, 256), eyes[0], P["target"])
            images = []
            for step, eye_k in enumerate([0, 0, 0, 1, 2, 2, 2]):
                if step and eye_k != [0, 0, 0, 1, 2, 2, 2][step - 1]:
                    pt.update(camera=nb.CameraController(world_position=eyes[eye_k], target=P["target"],
                                                         aspect_ratio=1.0))
                pt.dispatch()
                images.append(pt.output_texture())
            return images
        finally:
            if old is None:
                del os.environ["NB_UNIT_ORDER"]
            else:
                os.environ["NB_UNIT_ORDER"] = old

    a, b = run("1"), run("0")
    for x, y in zip(a, b):
        assert np.isfinite(x).all()
        np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32))
