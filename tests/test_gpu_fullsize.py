"""Full-size parity of the benchmark workload (instanced meshes, 4096 instances, 1920x1080,
12 bounces, shared BLAS): the persistent kernel at the size it is measured at, where every
XCD band, the work stealing at the tail and many waves per CU are in play.

* two whole frames equal the CPU oracle on the same arrays (per-pixel relative error <= 1e-3,
  bit-identical fraction printed), and so does one frame from inside the instance cloud;
* the persistent kernel and the one-tile-per-wave kernel render the same image bit for bit
  (two independent schedules of the same samples);
* 4 frames in one launch equal 4 (pipelined) launches bit for bit;
* the counted and the counter-free kernel render the same image, and the counts are the same
  in two passes (a checksum of the work done);
* cost-ordered work units (rebuilt after 1, 2 and 4 frames) and band order render the same image.
* pipelined frames (side streams + ordered mix) equal launches on the context stream.
"""
import os

import numpy as np
import pytest

import oracle
from tests import _workloads as W

pytestmark = pytest.mark.gpu

WIDTH, HEIGHT = 1920, 1080
REL_TOL = 1e-3


@pytest.fixture(scope="module")
def full():
    import nbounce_amd as nb
    from nbounce_amd import scenegen
    P = scenegen.LARGE
    v, i, pr = scenegen.instanced_meshes(0)
    scene = nb.Scene(v, i, pr)
    blas, tlas, inst = nb.build_scene(scene, share_blas=True)
    ctx = nb.Context(0)
    sb = nb.SceneBuffers(ctx, blas, tlas, inst, scene.vertices, scene.indices)
    faces = scenegen.sky_cube(P["env_face"])
    env = nb.EnvMap(ctx, faces)
    cam = nb.CameraController(world_position=P["eye"], target=P["target"], aspect_ratio=WIDTH / HEIGHT)
    osc = oracle.Scene(blas, tlas, inst, scene.vertices, scene.indices)
    yield nb, ctx, sb, env, cam, osc, faces, P
    ctx.close()


def _tracer(nb, ctx, sb, env, cam, P):
    pt = nb.Pathtracer(ctx, sb, cam, env, WIDTH, HEIGHT, resolution_factor=1.0, table=W.table(), sobol_stride=17)
    pt.globals.bounces = P["bounces"]
    pt.globals.contribution_factor = P["contribution_factor"]
    return pt


def _render(pt, frames, mode=None, env=None):
    env = dict(env or {})
    if mode:
        env["NB_RENDER"] = mode
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        for _ in range(frames):
            pt.dispatch()
        return pt.output_texture()
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def test_fullsize_frames_match_oracle(full):
    nb, ctx, sb, env, cam, osc, faces, P = full
    print("case: 1920x1080 instanced scene, 2 frames, the whole image against the oracle", flush=True)
    got = _render(_tracer(nb, ctx, sb, env, cam, P), 2)
    ref = None
    oenv = oracle.Env(faces)
    for k in (1, 2):
        g = oracle.globals_(k, np.float32(1) / np.float32(k), P["bounces"], P["contribution_factor"])
        ref, _ = oracle.render(osc, cam.data(), g, W.table(), 17, oenv, WIDTH, HEIGHT, rgba=ref, threads=8)
    assert (ref[..., :3].sum(-1) > 0).mean() > 0.5
    same = (got.view(np.uint32) == ref.view(np.uint32)).all(-1).mean()
    err = W.rel_err(got, ref)
    print(f"max rel err {err.max():.3e}, bit-identical pixels {same:.6f}")
    assert err.max() <= REL_TOL


def test_fullsize_interior_view_matches_oracle(full):
    """The same scene from inside the instance cloud (more hits, longer BLAS descents)."""
    nb, ctx, sb, env, cam, osc, faces, P = full
    from nbounce_amd import scenegen
    V = scenegen.LARGE_INTERIOR
    icam = nb.CameraController(world_position=V["eye"], target=V["target"], aspect_ratio=WIDTH / HEIGHT)
    print("case: 1920x1080 instanced scene from inside the cloud, 1 frame, the whole image against the oracle",
          flush=True)
    ctx.reset_counters()
    got = _render(_tracer(nb, ctx, sb, env, icam, P), 1)
    c = ctx.counters()
    g = oracle.globals_(1, 1.0, P["bounces"], P["contribution_factor"])
    ref, _ = oracle.render(osc, icam.data(), g, W.table(), 17, oracle.Env(faces), WIDTH, HEIGHT, threads=8)
    same = (got.view(np.uint32) == ref.view(np.uint32)).all(-1).mean()
    err = W.rel_err(got, ref)
    print(f"hits per ray {c['hits'] / c['rays']:.3f}, max rel err {err.max():.3e}, bit-identical pixels {same:.6f}")
    assert err.max() <= REL_TOL


def test_fullsize_persistent_equals_tile_kernel(full):
    nb, ctx, sb, env, cam, osc, faces, P = full
    print("case: 1920x1080, 3 frames, persistent vs one-tile-per-wave kernel", flush=True)
    a = _render(_tracer(nb, ctx, sb, env, cam, P), 3)
    b = _render(_tracer(nb, ctx, sb, env, cam, P), 3, mode="tile")
    assert np.isfinite(a).all()
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_fullsize_frames_in_one_launch_equal_launches(full):
    nb, ctx, sb, env, cam, osc, faces, P = full
    print("case: 1920x1080, 4 frames in one launch vs 4 launches", flush=True)
    a = _render(_tracer(nb, ctx, sb, env, cam, P), 4)
    pt = _tracer(nb, ctx, sb, env, cam, P)
    assert pt.dispatch_frames(4) == 4
    b = pt.output_texture()
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_fullsize_counting_changes_nothing(full):
    nb, ctx, sb, env, cam, osc, faces, P = full
    print("case: 1920x1080, 2 frames counted twice and once uncounted", flush=True)
    counts, images = [], []
    for on in (True, True, False):
        ctx.enable_counters(on)
        ctx.reset_counters()
        images.append(_render(_tracer(nb, ctx, sb, env, cam, P), 2))
        counts.append(ctx.counters())
    ctx.enable_counters(True)
    assert counts[0] == counts[1]
    assert counts[0]["samples"] == 2 * WIDTH * HEIGHT
    assert counts[0]["rays"] >= counts[0]["samples"]
    assert counts[2]["rays"] == 0
    for im in images[1:]:
        np.testing.assert_array_equal(images[0].view(np.uint32), im.view(np.uint32))


def test_fullsize_unit_order_changes_nothing(full):
    nb, ctx, sb, env, cam, osc, faces, P = full
    print("case: 1920x1080, 6 frames, cost-ordered work units vs band order", flush=True)
    a = _render(_tracer(nb, ctx, sb, env, cam, P), 6)
    b = _render(_tracer(nb, ctx, sb, env, cam, P), 6, env={"NB_UNIT_ORDER": "0"})
    assert np.isfinite(a).all()
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_fullsize_pipelined_equals_sequential(full):
    """Pipelined nb_render (path kernels of consecutive frames side by side, the mix on the
    context stream) against one launch after another on the context stream."""
    nb, ctx, sb, env, cam, osc, faces, P = full
    print("case: 1920x1080, 6 frames, pipelined vs sequential nb_render", flush=True)
    counts, images = [], []
    for pipe in ("1", "0"):
        ctx.reset_counters()
        images.append(_render(_tracer(nb, ctx, sb, env, cam, P), 6, env={"NB_PIPELINE": pipe}))
        counts.append(ctx.counters())
    assert np.isfinite(images[0]).all()
    np.testing.assert_array_equal(images[0].view(np.uint32), images[1].view(np.uint32))
    assert counts[0] == counts[1]


def test_fullsize_pipelined_moving_camera_equals_sequential(full):
    """A camera that moves every frame: cost halving and order rebuilds overlap the previous
    frame's paths (atomic costs, two order buffers)."""
    nb, ctx, sb, env, cam, osc, faces, P = full
    print("case: 1920x1080, 6 frames with a moving camera, pipelined vs sequential", flush=True)
    images = []
    for pipe in ("1", "0"):
        os.environ["NB_PIPELINE"] = pipe
        try:
            pt = _tracer(nb, ctx, sb, env, cam, P)
            for k in range(6):
                eye = (P["eye"][0] + 0.5 * k, P["eye"][1], P["eye"][2] + 0.3 * k)
                c = nb.CameraController(world_position=eye, target=P["target"], aspect_ratio=WIDTH / HEIGHT)
                pt.camera, pt._camera_data = c, c.data()   # no invalidation: every frame counts
                pt.dispatch()
            images.append(pt.output_texture())
        finally:
            del os.environ["NB_PIPELINE"]
    assert np.isfinite(images[0]).all()
    np.testing.assert_array_equal(images[0].view(np.uint32), images[1].view(np.uint32))


def test_fullsize_tlas_refit_changes_no_pixel(full):
    """The TLAS the kernels traverse has its boxes refitted to the instances' transformed
    geometry (NB_TLAS_REFIT, read at scene creation): fewer instance entries and box tests,
    the same closest hits, so the same image and ray count as the reference's boxes."""
    nb, ctx, sb, env, cam, osc, faces, P = full
    from nbounce_amd import scenegen
    print("case: 1920x1080, 3 frames, refitted TLAS vs the reference's TLAS boxes", flush=True)
    v, i, pr = scenegen.instanced_meshes(0)
    scene = nb.Scene(v, i, pr)
    blas, tlas, inst = nb.build_scene(scene, share_blas=True)
    os.environ["NB_TLAS_REFIT"] = "0"
    try:
        sb_ref = nb.SceneBuffers(ctx, blas, tlas, inst, scene.vertices, scene.indices)
    finally:
        del os.environ["NB_TLAS_REFIT"]
    counts, images = [], []
    for buffers in (sb, sb_ref):
        ctx.reset_counters()
        images.append(_render(_tracer(nb, ctx, buffers, env, cam, P), 3))
        counts.append(ctx.counters())
    np.testing.assert_array_equal(images[0].view(np.uint32), images[1].view(np.uint32))
    assert counts[0]["rays"] == counts[1]["rays"] and counts[0]["hits"] == counts[1]["hits"]
    assert counts[0]["n_inst"] < 0.7 * counts[1]["n_inst"]
    assert counts[0]["n_aabb"] < counts[1]["n_aabb"]
    sb_ref.close()
