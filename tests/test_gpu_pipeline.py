"""Pipelined nb_render: each frame's path kernel runs on a side stream and writes the frame's
samples, a mix kernel on the context stream accumulates them in frame order, so consecutive
frames' paths overlap.  These tests pin it against the same calls rendered one launch after
another on the context stream (NB_PIPELINE=0, read at every call) — bit for bit and count for
count — across every case where the ordering rules matter: the work-unit order rebuilt between
frames, a camera change (costs halved), a counter reset, two targets interleaved, a one-launch
nb_render_frames and a target write between pipelined frames, custom and zero weights; and
destroy calls right after pipelined frames."""
import os

import numpy as np
import pytest

from tests import _workloads as W

pytestmark = pytest.mark.gpu


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
    from nbounce_amd import scenegen
    w = W.small()
    sb = nb.SceneBuffers(ctx, w["blas"], w["tlas"], w["instances"], w["vertices"], w["indices"])
    env = nb.EnvMap(ctx, scenegen.sky_cube(128))
    return w, sb, env


def _pathtracer(nb, ctx, sb, env, size, eye, target, bounces=8):
    cam = nb.CameraController(world_position=eye, target=target, aspect_ratio=size[0] / size[1])
    pt = nb.Pathtracer(ctx, sb, cam, env, size[0], size[1], resolution_factor=1.0, table=W.table(), sobol_stride=17)
    pt.globals.bounces = bounces
    return pt, cam


class _Env:
    def __init__(self, **kv):
        self.kv = kv

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kv}
        os.environ.update(self.kv)

    def __exit__(self, *exc):
        for k, v in self.old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def _sequence(nb, ctx, sb, env, pipeline):
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    snaps, counts = [], []
    with _Env(NB_PIPELINE="1" if pipeline else "0"):
        a, _ = _pathtracer(nb, ctx, sb, env, (128, 96), P["eye"], P["target"])
        b, _ = _pathtracer(nb, ctx, sb, env, (64, 64), (3.0, 1.5, 6.0), P["target"])
        ctx.synchronize()
        ctx.reset_counters()
        for _ in range(5):            # the unit order is rebuilt after frames 1, 2 and 4
            a.dispatch()
        snaps.append(a.output_texture())
        for _ in range(4):            # two targets, interleaved
            a.dispatch()
            b.dispatch()
        counts.append(ctx.counters())
        ctx.reset_counters()          # a counter reset between pipelined frames
        a.dispatch()
        a.dispatch(weight=0.25)
        b.dispatch(weight=1.0)
        b.dispatch(weight=0.0)
        a.dispatch_frames(3)          # one launch on the context stream between pipelined frames
        a.dispatch()
        b.dispatch()
        a.update(nb.CameraController(world_position=(2.0, 2.5, 7.0), target=P["target"], aspect_ratio=128 / 96))
        for _ in range(3):            # another camera: the costs are halved before the first frame
            a.dispatch()
        rng = np.random.default_rng(5)
        a.output.write(rng.random((96, 128, 4), dtype=np.float32))
        a.dispatch(weight=0.5)        # mixes into the written image
        a.dispatch(weight=0.5)
        counts.append(ctx.counters())
        snaps += [a.output_texture(), b.output_texture()]
    return snaps, counts


def test_pipelined_equals_sequential_launches(nb, ctx, small_gpu):
    w, sb, env = small_gpu
    print("case: small scene, 128x96 and 64x64 targets, pipelined vs sequential nb_render", flush=True)
    seq_snaps, seq_counts = _sequence(nb, ctx, sb, env, pipeline=False)
    pipe_snaps, pipe_counts = _sequence(nb, ctx, sb, env, pipeline=True)
    for s, p in zip(seq_snaps, pipe_snaps):
        assert np.isfinite(s).all()
        np.testing.assert_array_equal(s.view(np.uint32), p.view(np.uint32))
    assert seq_counts == pipe_counts
    assert seq_counts[0]["samples"] == 9 * 128 * 96 + 4 * 64 * 64
    assert seq_counts[1]["samples"] == 11 * 128 * 96 + 3 * 64 * 64


def test_pipelined_weight_zero_gives_zero(nb, ctx, small_gpu):
    """weight 0: mix(color, s, 0) with color taken as 0 (the previous value is not read), so the
    image is 0 wherever the sample is finite."""
    w, sb, env = small_gpu
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    print("case: small scene 64x64, weight 0 after two frames", flush=True)
    with _Env(NB_PIPELINE="1"):
        pt, _ = _pathtracer(nb, ctx, sb, env, (64, 64), P["eye"], P["target"])
        pt.dispatch()
        pt.dispatch()
        assert pt.output_texture()[..., :3].max() > 0.0
        pt.dispatch(weight=0.0)
        img = pt.output_texture()
    assert (img[..., :3] == 0.0).all()
    assert (img[..., 3] == 1.0).all()


def test_pipelined_render_timing(nb, ctx, small_gpu):
    w, sb, env = small_gpu
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    print("case: small scene 64x64, per-render timing of pipelined frames", flush=True)
    with _Env(NB_PIPELINE="1"):
        pt, _ = _pathtracer(nb, ctx, sb, env, (64, 64), P["eye"], P["target"])
        ctx.enable_render_timing(True)
        try:
            pt.dispatch()
            pt.dispatch()
            ms = ctx.last_render_ms()
        finally:
            ctx.enable_render_timing(False)
    assert 0.0 < ms < 1000.0


def _moving(nb, ctx, sb, env, pipeline, frames=12, size=(128, 96)):
    """A camera that moves every second frame: the costs are halved before those frames and
    orders are rebuilt and switched to while the previous frame's paths are in flight.  The
    camera data changes without invalidating (the frames keep accumulating), so the final
    image depends on every frame; one move in the middle invalidates as CameraController
    updates do."""
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    out = []
    with _Env(NB_PIPELINE="1" if pipeline else "0"):
        pt, _ = _pathtracer(nb, ctx, sb, env, size, P["eye"], P["target"])
        ctx.synchronize()
        for k in range(frames):
            eye = (P["eye"][0] + 0.05 * k, P["eye"][1], P["eye"][2] - 0.03 * k)
            cam = nb.CameraController(world_position=eye, target=P["target"], aspect_ratio=size[0] / size[1])
            if k == frames // 2:
                pt.update(cam)
            else:
                pt.camera, pt._camera_data = cam, cam.data()
            pt.dispatch()
            pt.dispatch()
            if k % 4 == 3:
                out.append(pt.output_texture())
        out.append(pt.output_texture())
    return out


def test_pipelined_moving_camera_equals_sequential(nb, ctx, small_gpu):
    w, sb, env = small_gpu
    print("case: small scene 128x96, camera moving every second frame, pipelined vs sequential", flush=True)
    seq = _moving(nb, ctx, sb, env, pipeline=False)
    pipe = _moving(nb, ctx, sb, env, pipeline=True)
    for s, p in zip(seq, pipe):
        assert np.isfinite(s).all()
        np.testing.assert_array_equal(s.view(np.uint32), p.view(np.uint32))


def test_destroy_right_after_pipelined_frames(nb, ctx, small_gpu):
    """A scene and an environment map destroyed straight after pipelined nb_render calls, with no
    synchronisation in between: the destroy calls wait for the frames still in flight on the
    side streams, and the image is the one rendered with both kept alive."""
    w, sb, env = small_gpu
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    print("case: small scene 128x96, scene and environment destroyed right after 3 pipelined frames", flush=True)
    with _Env(NB_PIPELINE="1"):
        ref, _ = _pathtracer(nb, ctx, sb, env, (128, 96), P["eye"], P["target"])
        for _ in range(3):
            ref.dispatch()
        expected = ref.output_texture()
        sb2 = nb.SceneBuffers(ctx, w["blas"], w["tlas"], w["instances"], w["vertices"], w["indices"])
        env2 = nb.EnvMap(ctx, scenegen.sky_cube(128))
        pt, _ = _pathtracer(nb, ctx, sb2, env2, (128, 96), P["eye"], P["target"])
        for _ in range(3):
            pt.dispatch()
        env2.close()
        sb2.close()
        got = pt.output_texture()
    np.testing.assert_array_equal(expected.view(np.uint32), got.view(np.uint32))
