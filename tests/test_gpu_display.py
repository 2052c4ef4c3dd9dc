"""GPU display blit (nb_display_blit) against the restatement in oracle/display.py: bit-exact
RGBA16F surfaces for equal, magnified, minified and odd sizes, and for a rendered frame."""
import numpy as np
import pytest

from oracle import display

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


def _target(nb, ctx, img):
    from nbounce_amd.pathtracer import _Target
    t = _Target(ctx, img.shape[1], img.shape[0])
    t.write(img)
    return t


class _Holder:
    def __init__(self, target):
        self.output = target


@pytest.mark.parametrize("src,dst", [((64, 48), (64, 48)), ((64, 48), (200, 130)), ((256, 128), (100, 37)),
                                     ((8, 8), (1, 1)), ((1920, 1080), (2560, 1440))])
def test_blit_matches_oracle(nb, ctx, src, dst):
    print(f"case: blit {src[0]}x{src[1]} -> {dst[0]}x{dst[1]}", flush=True)
    rng = np.random.default_rng(src[0] * 7 + dst[1])
    img = rng.uniform(0.0, 8.0, size=(src[1], src[0], 4)).astype(np.float32)
    t = _target(nb, ctx, img)
    d = nb.BlitRenderer(ctx, *dst)
    d.render(_Holder(t))
    got = d.surface()
    want = display.blit(img, *dst)
    np.testing.assert_array_equal(got.view(np.uint16), want.view(np.uint16))
    d.close()
    t.close()


def test_blit_of_a_rendered_frame(nb, ctx):
    print("case: blit of a rendered 64x64 frame onto a 96x96 surface", flush=True)
    from nbounce_amd import scenegen
    from tests import _workloads as W
    wl = W.small()
    sb = nb.SceneBuffers(ctx, wl["blas"], wl["tlas"], wl["instances"], wl["vertices"], wl["indices"])
    env = nb.EnvMap(ctx, scenegen.sky_cube(32))
    P = scenegen.SMALL
    cam = nb.CameraController(world_position=P["eye"], target=P["target"], aspect_ratio=1.0)
    pt = nb.Pathtracer(ctx, sb, cam, env, 64, 64, resolution_factor=1.0)
    pt.dispatch_frames(4)
    d = nb.BlitRenderer(ctx, 96, 96)
    d.render(pt)
    got = d.surface()
    want = display.blit(pt.output_texture(), 96, 96)
    np.testing.assert_array_equal(got.view(np.uint16), want.view(np.uint16))
    d.close()
