"""Randomised parity: seeded random scenes, cameras and path settings, each image against the
oracle's.  A case draws its scene (spheres in a box of random size, or a random number of
instanced meshes over a random extent), a camera inside or outside the scene (some looking
exactly along an axis, so that primary rays have zero direction components and infinite slab
slopes), the image size, the bounce limit (0..13), the contribution factor, the first roulette
bounce and the sky (a cube or a lat-long map).  The scene arrays come from the oracle's builder
(the reference's layout, one BLAS per primitive).  Two pipelined nb_render frames must equal the
oracle's two frames (per-pixel relative error <= 1e-3, bit-identical fraction printed), and one
nb_render_frames call of two frames must equal the pipelined frames bit for bit."""
import numpy as np
import pytest

import oracle
from nbounce_amd import scenegen
from tests import _workloads as W

pytestmark = pytest.mark.gpu

REL_TOL = 1e-3
CASES = W.RANDOM_CASES


@pytest.fixture(scope="module")
def nb():
    import nbounce_amd
    return nbounce_amd


@pytest.fixture(scope="module")
def ctx(nb):
    c = nb.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("case", range(CASES))
def test_random_scene_matches_oracle(nb, ctx, case):
    c = W.random_case(case)
    v, i, pr = c["vit"]
    blas, tlas, inst, idx = oracle.build_scene(v, i, pr)
    w, h = c["size"]
    print(f"case {case}: {c['what']}, eye {np.round(c['eye'], 2)}, {w}x{h}, {c['bounces']} bounces, "
          f"cf {c['cf']:.2f}, roulette from {c['rr_first']}, {'lat-long' if c['latlong'] else 'cube'} sky", flush=True)
    if c["latlong"]:
        sky = scenegen.sky_latlong(64, 32)
        env, oenv = nb.EnvMap.latlong(ctx, sky), oracle.Env.latlong(sky)
    else:
        faces = scenegen.sky_cube(32)
        env, oenv = nb.EnvMap(ctx, faces), oracle.Env(faces)
    sb = nb.SceneBuffers(ctx, blas, tlas, inst, v, idx)
    osc = oracle.Scene(blas, tlas, inst, v, idx)
    cam = nb.CameraController(world_position=c["eye"], target=c["target"], aspect_ratio=w / h)

    def tracer():
        pt = nb.Pathtracer(ctx, sb, cam, env, w, h, resolution_factor=1.0, table=W.table(), sobol_stride=17)
        pt.globals.bounces = c["bounces"]
        pt.globals.contribution_factor = c["cf"]
        return pt

    try:
        ctx.set_render_options(roulette_first_bounce=c["rr_first"])
        oracle.set_roulette_first(c["rr_first"])
        a = tracer()
        a.dispatch()
        a.dispatch()
        got = a.output_texture()
        b = tracer()
        assert b.dispatch_frames(2) == 2
        got_b = b.output_texture()
        ref = None
        for k in (1, 2):
            g = oracle.globals_(k, np.float32(1) / np.float32(k), c["bounces"], c["cf"])
            ref, _ = oracle.render(osc, cam.data(), g, W.table(), 17, oenv, w, h, rgba=ref, threads=8)
    finally:
        ctx.set_render_options(roulette_first_bounce=0)
        oracle.set_roulette_first(0)
        sb.close()
        env.close()
    np.testing.assert_array_equal(got.view(np.uint32), got_b.view(np.uint32))
    same = (got.view(np.uint32) == ref.view(np.uint32)).all(-1).mean()
    err = W.rel_err(got, ref)
    print(f"bit-identical pixels {same:.5f}, max rel err {err.max():.3e}")
    assert err.max() <= REL_TOL
