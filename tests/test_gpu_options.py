"""GPU parity of the native options (no counterpart in the reference) against the oracle:
the lat-long environment form and Russian roulette from a later bounce, through nb_render
(pipelined) and the sample-parallel nb_render_frames (serving kernels)."""
import numpy as np
import pytest

import oracle
from tests import _workloads as W

pytestmark = pytest.mark.gpu

REL_TOL = 1e-3


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
def small(nb, ctx):
    w = W.small()
    sb = nb.SceneBuffers(ctx, w["blas"], w["tlas"], w["instances"], w["vertices"], w["indices"])
    return w, sb


def _tracer(nb, ctx, sb, env, size=(64, 64), bounces=8):
    from nbounce_amd import scenegen
    P = scenegen.SMALL
    cam = nb.CameraController(world_position=P["eye"], target=P["target"], aspect_ratio=size[0] / size[1])
    pt = nb.Pathtracer(ctx, sb, cam, env, size[0], size[1], resolution_factor=1.0, table=W.table(), sobol_stride=17)
    pt.globals.bounces = bounces
    return pt, cam


def _oracle_frames(w, cam, env, frames, size=(64, 64), bounces=8):
    img = None
    for k in range(1, frames + 1):
        g = oracle.globals_(k, np.float32(1) / np.float32(k), bounces, 4.0)
        img, _ = oracle.render(W.oracle_scene(w), cam.data(), g, W.table(), 17, env, size[0], size[1], rgba=img,
                               threads=8)
    return img


def _compare(got, ref):
    same = (got.view(np.uint32) == ref.view(np.uint32)).all(-1).mean()
    err = W.rel_err(got, ref)
    print(f"bit-identical pixels {same:.5f}, max rel err {err.max():.3e}")
    assert err.max() <= REL_TOL


def test_latlong_environment_matches_oracle(nb, ctx, small):
    from nbounce_amd import scenegen
    w, sb = small
    sky = scenegen.sky_latlong(512, 256)
    env = nb.EnvMap.latlong(ctx, sky)
    print("case: small scene 64x64, 512x256 lat-long sky, 2 pipelined frames and 3 frames in one call", flush=True)
    pt, cam = _tracer(nb, ctx, sb, env)
    pt.dispatch()
    pt.dispatch()
    got = pt.output_texture()
    ref = _oracle_frames(w, cam, oracle.Env.latlong(sky), 2)
    assert (got[..., :3].sum(-1) > 0).mean() > 0.5
    _compare(got, ref)
    # the sample-parallel form (serving kernels) equals three nb_render calls bit for bit
    a, _ = _tracer(nb, ctx, sb, env)
    for _ in range(3):
        a.dispatch()
    b, _ = _tracer(nb, ctx, sb, env)
    assert b.dispatch_frames(3) == 3
    np.testing.assert_array_equal(a.output_texture().view(np.uint32), b.output_texture().view(np.uint32))


def test_roulette_from_bounce_three_matches_oracle(nb, ctx, small):
    from nbounce_amd import scenegen
    w, sb = small
    faces = scenegen.sky_cube(128)
    env = nb.EnvMap(ctx, faces)
    print("case: small scene 64x64, Russian roulette from bounce 3, 2 frames", flush=True)
    counts = {}
    try:
        for first in (0, 3):
            ctx.set_render_options(roulette_first_bounce=first)
            oracle.set_roulette_first(first)
            ctx.reset_counters()
            pt, cam = _tracer(nb, ctx, sb, env)
            pt.dispatch()
            pt.dispatch()
            got = pt.output_texture()
            counts[first] = ctx.counters()
            _compare(got, _oracle_frames(w, cam, oracle.Env(faces), 2))
        # the sample-parallel form carries the option too
        a, _ = _tracer(nb, ctx, sb, env)
        for _ in range(3):
            a.dispatch()
        b, _ = _tracer(nb, ctx, sb, env)
        assert b.dispatch_frames(3) == 3
        np.testing.assert_array_equal(a.output_texture().view(np.uint32), b.output_texture().view(np.uint32))
    finally:
        ctx.set_render_options(roulette_first_bounce=0)
        oracle.set_roulette_first(0)
    r0 = counts[0]["rays"] / counts[0]["samples"]
    r3 = counts[3]["rays"] / counts[3]["samples"]
    print(f"rays per sample: roulette from bounce 0 {r0:.3f}, from bounce 3 {r3:.3f}")
    assert r3 > r0


def test_bounce_limit_and_roulette_start_bounds(nb, ctx, small):
    """Bounce limits above 65535 are refused (the kernels keep the limit in 16 bits beside the
    first roulette bounce); a first roulette bounce beyond the limit means no roulette at all,
    and a huge one is clamped to the same."""
    from nbounce_amd import scenegen
    w, sb = small
    env = nb.EnvMap(ctx, scenegen.sky_cube(32))
    print("case: small scene 64x64, bounces 70000 refused; roulette start 9 vs 2^20", flush=True)
    pt, _ = _tracer(nb, ctx, sb, env)
    pt.globals.bounces = 70000
    with pytest.raises(nb.NbError):
        pt.dispatch()
    images = []
    try:
        for first in (9, 1 << 20):
            ctx.set_render_options(roulette_first_bounce=first)
            pt, _ = _tracer(nb, ctx, sb, env)
            pt.dispatch()
            images.append(pt.output_texture())
    finally:
        ctx.set_render_options(roulette_first_bounce=0)
    np.testing.assert_array_equal(images[0].view(np.uint32), images[1].view(np.uint32))


def test_small_configuration_as_written_matches_oracle(nb, ctx, small):
    """The small configuration as specified, at its full size: 256x256, 16 spp, 8 bounces, Russian
    roulette from bounce 3, a 512x256 lat-long sky.  All 65,536 pixels of the 16-spp image, rendered
    in one nb_render_frames call (sample-parallel kernels) and by 16 pipelined nb_render calls,
    against the oracle's 16 frames."""
    from nbounce_amd import scenegen
    w, sb = small
    sky = scenegen.sky_latlong(512, 256)
    env = nb.EnvMap.latlong(ctx, sky)
    size = (256, 256)
    print("case: small scene 256x256, 16 spp, roulette from bounce 3, 512x256 lat-long sky: one call "
          "and 16 pipelined frames vs the oracle", flush=True)
    try:
        ctx.set_render_options(roulette_first_bounce=3)
        oracle.set_roulette_first(3)
        a, cam = _tracer(nb, ctx, sb, env, size)
        assert a.dispatch_frames(16) == 16
        b, _ = _tracer(nb, ctx, sb, env, size)
        for _ in range(16):
            b.dispatch()
        got_a, got_b = a.output_texture(), b.output_texture()
        ref = _oracle_frames(w, cam, oracle.Env.latlong(sky), 16, size=size)
    finally:
        ctx.set_render_options(roulette_first_bounce=0)
        oracle.set_roulette_first(0)
    np.testing.assert_array_equal(got_a.view(np.uint32), got_b.view(np.uint32))
    assert (got_a[..., :3].sum(-1) > 0).mean() > 0.5
    _compare(got_a, ref)
