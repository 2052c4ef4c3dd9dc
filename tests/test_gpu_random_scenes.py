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
CASES = 24


@pytest.fixture(scope="module")
def nb():
    import nbounce_amd
    return nbounce_amd


@pytest.fixture(scope="module")
def ctx(nb):
    c = nb.Context(0)
    yield c
    c.close()


def _case(case):
    """The scene, camera and settings of one case (deterministic in `case`)."""
    rng = np.random.default_rng(7000 + case)
    if case % 2 == 0:
        box = float(rng.uniform(2.0, 30.0))
        n = int(rng.integers(1, 41))
        v, i, pr = scenegen.spheres_in_box(seed=case, n_spheres=n, n_emissive=int(rng.integers(0, n + 1)),
                                           subdivisions=int(rng.integers(1, 4)), box=box)
        lo, hi = np.zeros(3), np.full(3, box)
        what = f"{n} spheres in a {box:.1f} box"
    else:
        extent = float(rng.uniform(5.0, 300.0))
        n = int(rng.integers(1, 49))
        v, i, pr = scenegen.instanced_meshes(seed=case, n_instances=n, extent=extent,
                                             emissive_every=int(rng.integers(1, 9)))
        lo, hi = np.zeros(3), np.full(3, extent)
        what = f"{n} instanced meshes over {extent:.0f}"
    span = hi - lo
    eye = lo + span * rng.uniform(-0.4, 1.4, 3) if rng.random() < 0.5 else lo + span * rng.uniform(0.1, 0.9, 3)
    if rng.random() < 0.3:
        # exactly along the x or z axis (the up vector is +y)
        axis = np.zeros(3)
        axis[int(rng.choice([0, 2]))] = float(rng.choice([-1.0, 1.0]))
        target = eye + axis * float(span.max())
        what += ", axis-aligned view"
    else:
        target = lo + span * rng.uniform(0.2, 0.8, 3)
        if np.linalg.norm(target - eye) < 1e-3 * span.max():
            target = target + span * 0.25
    size = [(8, 8), (48, 32), (32, 64), (64, 48)][int(rng.integers(0, 4))]
    bounces = int(rng.integers(0, 14))
    cf = float(rng.uniform(0.25, 8.0))
    rr_first = int(rng.integers(1, 6)) if case % 3 == 2 else 0
    latlong = rng.random() < 0.4
    return dict(vit=(v, i, pr), eye=tuple(float(x) for x in eye), target=tuple(float(x) for x in target),
                size=size, bounces=bounces, cf=cf, rr_first=rr_first, latlong=latlong, what=what)


@pytest.mark.parametrize("case", range(CASES))
def test_random_scene_matches_oracle(nb, ctx, case):
    c = _case(case)
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
