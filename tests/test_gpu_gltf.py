"""GPU parity on the reference's own scene files (spheres.glb, testscene.glb) read by the
library's glTF reader: primary hits and one frame against the CPU oracle, and the shared-geometry
reading (one copy of testscene's two Suzanne meshes, NB_BUILD_SHARE_BLAS) rendering the same
image as the reference's duplicated reading."""
import os

import numpy as np
import pytest

import oracle
from tests import _workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REL_TOL = 1e-3
VIEWS = {"spheres.glb": ((0.0, 1.0, 0.9), (0.0, 0.0, -0.5)),
         "testscene.glb": ((0.0, 1.0, 5.0), (0.0, 0.3, 0.0))}


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
def env(nb, ctx):
    from nbounce_amd import scenegen
    return nb.EnvMap(ctx, scenegen.sky_cube(64))


def _load(nb, name, share=False):
    scene = nb.Scene().parse_gltf(os.path.join(GOLDEN, name), share_geometry=share)
    blas, tlas, inst = nb.build_scene(scene, share_blas=share)
    return scene, blas, tlas, inst


def _tracer(nb, ctx, env, name, scene, blas, tlas, inst, size=(96, 64)):
    sb = nb.SceneBuffers(ctx, blas, tlas, inst, scene.vertices, scene.indices)
    eye, target = VIEWS[name]
    cam = nb.CameraController(world_position=eye, target=target, aspect_ratio=size[0] / size[1])
    pt = nb.Pathtracer(ctx, sb, cam, env, size[0], size[1], resolution_factor=1.0, table=W.table(), sobol_stride=17)
    return pt, cam, sb


@pytest.mark.parametrize("name", ["spheres.glb", "testscene.glb"])
def test_reference_asset_frame_matches_oracle(nb, ctx, env, name):
    print(f"case: {name} primary hits + one frame 96x64, 8 bounces", flush=True)
    scene, blas, tlas, inst = _load(nb, name)
    pt, cam, _ = _tracer(nb, ctx, env, name, scene, blas, tlas, inst)
    osc = oracle.Scene(blas, tlas, inst, scene.vertices, scene.indices)
    hits = pt.primary_hits(sample=1)
    ref_hits = oracle.primary_hits(osc, cam.data(), 96, 64, W.table(), 17, 1)
    assert (hits[..., 0] != 0xFFFFFFFF).mean() > 0.2
    np.testing.assert_array_equal(hits, ref_hits)
    pt.dispatch()
    got = pt.output_texture()
    ref, _ = oracle.render(osc, cam.data(), oracle.globals_(1, 1.0, 8, 4.0), W.table(), 17, oracle.Env(env.faces),
                           96, 64, threads=8)
    err = W.rel_err(got, ref)
    print(f"max rel err {err.max():.3e}")
    assert err.max() <= REL_TOL


def test_shared_geometry_renders_the_same_image(nb, ctx, env):
    print("case: testscene.glb shared vs duplicated geometry, 4 frames 96x64", flush=True)
    dup = _load(nb, "testscene.glb", share=False)
    shr = _load(nb, "testscene.glb", share=True)
    assert len(shr[0].vertices) < len(dup[0].vertices)
    imgs = []
    for scene, blas, tlas, inst in (dup, shr):
        pt, _, _ = _tracer(nb, ctx, env, "testscene.glb", scene, blas, tlas, inst)
        pt.dispatch_frames(4)
        imgs.append(pt.output_texture())
    assert imgs[0].tobytes() == imgs[1].tobytes()
