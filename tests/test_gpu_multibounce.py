"""GPU parity on the multi-bounce cases that pin the integrator (tests/test_oracle_multibounce.py):
closed mirror and diffuse boxes rendered with marker and scanned Sobol tables, through the C ABI,
bit for bit against the CPU oracle, ray counts included (pathtracing.wgsl:151-159,164-239)."""
import math

import numpy as np
import pytest

import oracle
from tests import test_oracle_multibounce as M

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


def _render_both(nb, ctx, oscene, table, sample, bounces, cf):
    """One frame (sample `sample`, weight 1) of the 8x8 box view on the GPU and with the oracle;
    returns (gpu image, oracle image, gpu rays, oracle rays)."""
    from nbounce_amd import scenegen
    from nbounce_amd._lib import check, lib, ptr
    blas, tlas, inst, v, idx = oscene.arrays
    sb = nb.SceneBuffers(ctx, blas, tlas, inst, v, idx)
    env = nb.EnvMap(ctx, scenegen.constant_cube(4, 1.0))
    cam = nb.CameraController(world_position=(5.0, 5.0, 5.0), target=(5.83, 6.0, 7.21), aspect_ratio=1.0)
    assert np.array_equal(cam.data().view(np.uint32), M.CAM.view(np.uint32))
    pt = nb.Pathtracer(ctx, sb, cam, env, M.W, M.H, resolution_factor=1.0, table=table, sobol_stride=M.STRIDE)
    ctx.enable_counters(True)
    ctx.reset_counters()
    g = nb.Globals(sample, 1.0, bounces, cf).array()
    check(lib.nb_render(ctx.handle, sb.handle, pt.output.handle, ptr(pt._camera_data), ptr(g), env.handle))
    got = pt.output_texture()
    gpu_rays = ctx.counters()["rays"]
    og = oracle.globals_(sample, 1.0, bounces, cf)
    ref, cnt = oracle.render(oscene, cam.data(), og, table, M.STRIDE, M.ENV, M.W, M.H, threads=4)
    return got, ref, gpu_rays, int(cnt["rays"])


def _assert_same(got, ref, gpu_rays, ref_rays):
    same = (got.view(np.uint32) == ref.view(np.uint32)).all(axis=-1).mean()
    print(f"bit-identical pixels {same:.4f}, rays gpu {gpu_rays} oracle {ref_rays}", flush=True)
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert gpu_rays == ref_rays


@pytest.mark.parametrize("bounces,b0", [(8, 0), (8, 7), (12, 5), (12, 9), (12, 11)])
def test_roulette_marker_table_matches_oracle(nb, ctx, bounces, b0):
    """Grey mirror box with the roulette marker in entry 2*17 + 1 + 2*b0 (the next row's entries
    for b0 >= 8): pixel (4, 4) makes b0 + 1 rays on both sides, every pixel is bit-identical."""
    print(f"case: grey mirror box, marker bounce {b0} of {bounces}", flush=True)
    scene = M.closed_box(color=(0.2, 0.2, 0.2))
    shift = oracle.pixel_shift(4, 4)
    row = 2
    t = M.base_table(5, shift, x=0.5, y=0.5, z=0.02)
    t.reshape(-1, 4)[row * M.STRIDE + 1 + 2 * b0, 2] = M.entry_for(0.98, shift[2])[0]
    _, rays = M.rays_of(scene, t, row, bounces, 1.0)
    assert rays == b0 + 1
    _assert_same(*_render_both(nb, ctx, scene, t, row, bounces, 1.0))


@pytest.mark.parametrize("bounces,u", [(8, 0.3), (8, 0.7), (16, 0.5), (16, 0.95)])
def test_length_term_scan_matches_oracle(nb, ctx, bounces, u):
    print(f"case: white mirror box, roulette numbers {u}, {bounces} bounces", flush=True)
    scene = M.closed_box(color=(1.0, 1.0, 1.0))
    shift = oracle.pixel_shift(4, 4)
    t = M.base_table(4, shift, x=0.5, y=0.5, z=u)
    _assert_same(*_render_both(nb, ctx, scene, t, 1, bounces, 4.0))


@pytest.mark.parametrize("row", [1, 2, 3])
def test_diffuse_box_with_emitter_matches_oracle(nb, ctx, row):
    """Diffuse box with an emitter over an i.i.d. table: long paths through every row's sets."""
    print(f"case: diffuse box with emitter, i.i.d. table, row {row}, 12 bounces", flush=True)
    scene = M.closed_box(color=(0.7, 0.6, 0.5), roughness=1.0, metallic=0.0, emitter=3.0)
    t = np.random.default_rng(row).random((5, M.STRIDE, 4), dtype=np.float32)
    got, ref, gr, orr = _render_both(nb, ctx, scene, t, row, 12, 1e6)
    assert got[..., :3].max() > 0
    _assert_same(got, ref, gr, orr)
    assert math.isfinite(float(got.sum()))
