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
