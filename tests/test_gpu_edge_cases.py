"""GPU parity on the path's edge cases, through the C ABI against the CPU oracle: zero bounces
(only the primary ray's emitter or sky term, pathtracing.wgsl:164-175), a zero contribution
factor (Russian roulette ends every path after its first scatter, pathtracing.wgsl:227-236),
16 bounces (the fixed 17-entry Sobol row stride reads two rows further, pathtracing.wgsl:4,151-159),
the smallest 8x8 target and non-square targets (pathtracer.rs:192-202), and a two-triangle scene
with one emitter (one instance per triangle, single-leaf BLASes, back-face culling); and the
render entry points' argument checks, after which a valid render is unaffected."""
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
    from nbounce_amd import scenegen
    w = W.small()
    sb = nb.SceneBuffers(ctx, w["blas"], w["tlas"], w["instances"], w["vertices"], w["indices"])
    env = nb.EnvMap(ctx, scenegen.sky_cube(64))
    return w, sb, env


def _frames_vs_oracle(nb, ctx, sb, osc, env, size, eye, target, frames, bounces=8, cf=4.0):
    """Renders `frames` progressive samples on the GPU and with the oracle; returns both images."""
    cam = nb.CameraController(world_position=eye, target=target, aspect_ratio=size[0] / size[1])
    pt = nb.Pathtracer(ctx, sb, cam, env, size[0], size[1], resolution_factor=1.0, table=W.table(), sobol_stride=17)
    pt.globals.bounces = bounces
    pt.globals.contribution_factor = cf
    for _ in range(frames):
        pt.dispatch()
    got = pt.output_texture()
    assert got.shape == (size[1], size[0], 4)
    ref = np.zeros((size[1], size[0], 4), dtype=np.float32)
    oenv = oracle.Env(env.faces)
    for k in range(1, frames + 1):
        g = oracle.globals_(k, np.float32(1.0) / np.float32(k), bounces, cf)
        ref, _ = oracle.render(osc, cam.data(), g, W.table(), 17, oenv, size[0], size[1], rgba=ref, threads=8)
    return got, ref


def _check(got, ref):
    err = W.rel_err(got, ref)
    same = (got.view(np.uint32) == ref.view(np.uint32)).all(axis=-1).mean()
    print(f"bit-identical pixels {same:.5f}, max rel err {err.max():.3e}", flush=True)
    assert np.isfinite(got).all()
    assert err.max() <= REL_TOL
    assert (got[..., 3] == 1.0).all()   # alpha is forced to 1 (pathtracing.wgsl:258)


def test_zero_bounces_match_oracle(nb, ctx, small_gpu):
    from nbounce_amd import scenegen
    w, sb, env = small_gpu
    P = scenegen.SMALL
    print("case: small scene 64x64, 0 bounces, 2 frames", flush=True)
    got, ref = _frames_vs_oracle(nb, ctx, sb, W.oracle_scene(w), env, (64, 64), P["eye"], P["target"], 2, bounces=0)
    _check(got, ref)
    # only emitters and the sky contribute: some pixels are black (diffuse or specular hits)
    assert (got[..., :3] == 0).all(axis=-1).any()


def test_zero_contribution_factor_matches_oracle(nb, ctx, small_gpu):
    from nbounce_amd import scenegen
    w, sb, env = small_gpu
    P = scenegen.SMALL
    print("case: small scene 64x64, contribution factor 0, 2 frames", flush=True)
    got, ref = _frames_vs_oracle(nb, ctx, sb, W.oracle_scene(w), env, (64, 64), P["eye"], P["target"], 2, cf=0.0)
    _check(got, ref)
    ctx.reset_counters()
    cam = nb.CameraController(world_position=P["eye"], target=P["target"], aspect_ratio=1.0)
    pt = nb.Pathtracer(ctx, sb, cam, env, 64, 64, resolution_factor=1.0, table=W.table(), sobol_stride=17)
    pt.globals.contribution_factor = 0.0
    pt.dispatch()
    pt.output_texture()
    # with p_continue = 0 every path ends at its first scatter: one ray per sample
    assert ctx.counters()["rays"] == 64 * 64


def test_sixteen_bounces_match_oracle(nb, ctx, small_gpu):
    from nbounce_amd import scenegen
    w, sb, env = small_gpu
    P = scenegen.SMALL
    print("case: small scene 48x32, 16 bounces, 3 frames", flush=True)
    got, ref = _frames_vs_oracle(nb, ctx, sb, W.oracle_scene(w), env, (48, 32), P["eye"], P["target"], 3, bounces=16)
    _check(got, ref)


@pytest.mark.parametrize("size", [(8, 8), (24, 16), (8, 40), (200, 8)])
def test_small_and_non_square_targets_match_oracle(nb, ctx, small_gpu, size):
    from nbounce_amd import scenegen
    w, sb, env = small_gpu
    P = scenegen.SMALL
    print(f"case: small scene {size[0]}x{size[1]}, 8 bounces, 2 frames", flush=True)
    got, ref = _frames_vs_oracle(nb, ctx, sb, W.oracle_scene(w), env, size, P["eye"], P["target"], 2)
    _check(got, ref)


def _two_triangles():
    """A front-facing emissive triangle in front of a larger diffuse one, both facing +z."""
    v = np.zeros(6, dtype=oracle.VERTEX)
    v["position"] = [(-1, -1, 0), (1, -1, 0), (0, 1, 0), (-4, -4, -2), (4, -4, -2), (0, 4, -2)]
    v["normal"] = (0, 0, 1)
    v["tangent"] = (1, 0, 0, 1)
    idx = np.arange(6, dtype=np.uint32)
    pr = np.zeros(2, dtype=oracle.PRIMITIVE)
    pr["local_to_world"] = np.eye(4, dtype=np.float32).reshape(-1)
    pr[0]["color"], pr[0]["emissive"] = (3.0, 2.0, 1.0, 1.0), 1.0
    pr[1]["color"], pr[1]["roughness"], pr[1]["metallic"] = (0.8, 0.7, 0.6, 1.0), 0.5, 0.0
    pr["index_start"], pr["index_end"] = (0, 3), (3, 6)
    return v, idx, pr


def test_two_triangle_scene_matches_oracle(nb, ctx):
    from nbounce_amd import scenegen
    v, i, pr = _two_triangles()
    blas, tlas, inst, idx = oracle.build_scene(v, i, pr)
    # the library's host builder makes the same arrays
    scene = nb.Scene(v, i.copy(), pr)
    b2, t2, i2 = nb.build_scene(scene)
    np.testing.assert_array_equal(b2.view(np.uint8), blas.view(np.uint8))
    np.testing.assert_array_equal(t2.view(np.uint8), tlas.view(np.uint8))
    np.testing.assert_array_equal(scene.indices, idx)
    sb = nb.SceneBuffers(ctx, blas, tlas, inst, v, idx)
    env = nb.EnvMap(ctx, scenegen.sky_cube(32))
    osc = oracle.Scene(blas, tlas, inst, v, idx)
    print("case: two-triangle scene 32x32, 8 bounces, 4 frames (front view)", flush=True)
    got, ref = _frames_vs_oracle(nb, ctx, sb, osc, env, (32, 32), (0.0, 0.0, 4.0), (0.0, 0.0, 0.0), 4)
    _check(got, ref)
    # the emitter's colour is seen directly at the image centre
    np.testing.assert_allclose(got[16, 16, :3], [3.0, 2.0, 1.0], rtol=1e-6)
    print("case: two-triangle scene 32x32 from behind (both culled: sky only)", flush=True)
    got, ref = _frames_vs_oracle(nb, ctx, sb, osc, env, (32, 32), (0.0, 0.0, -6.0), (0.0, 0.0, 0.0), 2)
    _check(got, ref)


def test_render_arguments_are_validated(nb, ctx, small_gpu):
    """nb_render / nb_render_frames refuse (NB_ERR_INVALID_ARGUMENT, nothing launched) a null
    scene, objects of another context, a context without a Sobol table and, for
    nb_render_frames, sample 0; n_frames = 0 renders nothing.  A valid render after all of them
    gives the image an untouched tracer gives."""
    import ctypes
    from nbounce_amd import scenegen
    from nbounce_amd._lib import lib, ptr
    w, sb, env = small_gpu
    P = scenegen.SMALL
    print("case: invalid render arguments, then a valid render", flush=True)
    cam = nb.CameraController(world_position=P["eye"], target=P["target"], aspect_ratio=1.0)
    pt = nb.Pathtracer(ctx, sb, cam, env, 32, 32, resolution_factor=1.0, table=W.table(), sobol_stride=17)
    camd = cam.data()
    g = nb.Globals(1, 1.0, 8, 4.0).array()
    t = pt.output.handle
    assert lib.nb_render(ctx.handle, None, t, ptr(camd), ptr(g), env.handle) == 1
    other = nb.Context(0)
    try:
        t2 = ctypes.c_void_p()
        assert lib.nb_target_create(other.handle, 32, 32, ctypes.byref(t2)) == 0
        # a target of another context, and a context that holds no Sobol table
        assert lib.nb_render(ctx.handle, sb.handle, t2, ptr(camd), ptr(g), env.handle) == 1
        assert b"belong to the context" in lib.nb_last_error()
        sb2 = nb.SceneBuffers(other, w["blas"], w["tlas"], w["instances"], w["vertices"], w["indices"])
        env2 = nb.EnvMap(other, scenegen.sky_cube(8))
        assert lib.nb_render(other.handle, sb2.handle, t2, ptr(camd), ptr(g), env2.handle) == 1
        assert b"Sobol" in lib.nb_last_error()
        sb2.close()
        env2.close()
        lib.nb_target_destroy(t2)
    finally:
        other.close()
    g0 = nb.Globals(0, 0.0, 8, 4.0).array()
    assert lib.nb_render_frames(ctx.handle, sb.handle, t, ptr(camd), ptr(g0), 2, env.handle) == 1
    assert lib.nb_render_frames(ctx.handle, sb.handle, t, ptr(camd), ptr(g), 0, env.handle) == 0
    assert (pt.output_texture() == 0.0).all()
    pt.dispatch()
    pt.dispatch()
    fresh = nb.Pathtracer(ctx, sb, cam, env, 32, 32, resolution_factor=1.0, table=W.table(), sobol_stride=17)
    fresh.dispatch()
    fresh.dispatch()
    np.testing.assert_array_equal(pt.output_texture().view(np.uint32), fresh.output_texture().view(np.uint32))


def _chain_blas(depth):
    """A BLAS whose left spine is `depth` inner nodes deep: inner node k has children (2k+1, 2k+2),
    the left one the next inner node, the right one a leaf; the last inner node has two leaves.
    Every leaf holds triangle 0."""
    from nbounce_amd import layout
    n = 2 * depth + 1
    nodes = np.zeros(n, dtype=layout.NODE)
    nodes["min"], nodes["max"] = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    for k in range(depth):
        i = 0 if k == 0 else 2 * k - 1
        nodes[i]["start"], nodes[i]["end"] = 2 * k + 1, 0
        nodes[2 * k + 2]["start"], nodes[2 * k + 2]["end"] = 0, 1
    nodes[2 * depth - 1]["start"], nodes[2 * depth - 1]["end"] = 0, 1
    return nodes


def test_malformed_scene_layouts_are_rejected(nb, ctx, small_gpu):
    """nb_scene_create validates the caller's reference-layout arrays on the host before anything
    reaches the device, since the kernels follow the node references they are given: an index
    past the vertices, an inner node whose children are not after it (including a child index of
    2^32 - 1, where the second child's index wraps to 0), a second child past the array, a leaf
    range outside the primitives or empty, a tree deeper than 32 levels, an instance whose BLAS
    root is out of range and a TLAS without instances each fail with NB_ERR_BAD_LAYOUT.  A tree
    exactly 32 levels deep is accepted, and a render on the context afterwards is unaffected."""
    from nbounce_amd import scenegen
    w, sb, env = small_gpu
    blas, tlas, inst, v, idx = w["blas"], w["tlas"], w["instances"], w["vertices"], w["indices"]
    inner_b = int(np.nonzero(blas["end"] == 0)[0][0])
    leaf_t = int(np.nonzero(tlas["end"] > 0)[0][0])

    def rejected(b=blas, t=tlas, i=inst, vv=v, ix=idx, what=""):
        print("case: malformed layout:", what, flush=True)
        with pytest.raises(nb.NbError) as e:
            nb.SceneBuffers(ctx, b, t, i, vv, ix)
        assert e.value.status == 4, (what, str(e.value))   # NB_ERR_BAD_LAYOUT

    ix = idx.copy(); ix[5] = len(v)
    rejected(ix=ix, what="index past the vertices")
    b = blas.copy(); b[inner_b]["start"] = inner_b
    rejected(b=b, what="children not after their parent")
    b = blas.copy(); b[inner_b]["start"] = 0xFFFFFFFF
    rejected(b=b, what="child index 2^32 - 1")
    b = blas.copy(); b[inner_b]["start"] = len(blas) - 1
    rejected(b=b, what="second child past the array")
    leaf_b = int(np.nonzero(blas["end"] > 0)[0][0])
    b = blas.copy(); b[leaf_b]["end"] = len(idx) // 3 + 1
    rejected(b=b, what="leaf range past the triangles")
    b = blas.copy(); b[leaf_b]["start"] = b[leaf_b]["end"]
    rejected(b=b, what="empty leaf range")
    t = tlas.copy(); t[leaf_t]["end"] = len(inst) + 1
    rejected(t=t, what="TLAS leaf range past the instances")
    i2 = inst.copy(); i2[3]["node"] = len(blas)
    rejected(i=i2, what="instance BLAS root out of range")
    rejected(i=inst[:0], what="TLAS without instances")
    # a one-instance scene over a chain BLAS: 33 inner levels are refused, 32 accepted
    i1 = inst[:1].copy(); i1[0]["node"] = 0
    t1 = np.zeros(1, dtype=tlas.dtype); t1[0]["min"], t1[0]["max"] = (-1e3,) * 3, (1e3,) * 3
    t1[0]["start"], t1[0]["end"] = 0, 1
    rejected(b=_chain_blas(33), t=t1, i=i1, what="BLAS 33 levels deep")
    print("case: BLAS exactly 32 levels deep is accepted", flush=True)
    nb.SceneBuffers(ctx, _chain_blas(32), t1, i1, v, idx).close()
    # the context still renders the valid scene exactly as the oracle does
    print("case: valid render after the rejections", flush=True)
    got, ref = _frames_vs_oracle(nb, ctx, sb, W.oracle_scene(w), env, (32, 32), scenegen.SMALL["eye"],
                                 scenegen.SMALL["target"], 2)
    _check(got, ref)


def test_non_finite_inputs_render_without_fault(nb, ctx, small_gpu):
    """Non-finite values reach the kernels as data, not as addresses: the slab test misses a NaN
    box or ray, the triangle test rejects a NaN triangle, and the environment lookup clamps a NaN
    direction to its border texels.  A camera of NaNs renders (every pixel NaN, alpha 1); a scene
    whose diffuse triangle has NaN vertices renders as the oracle renders it (that triangle is
    never hit); a render of the valid scene afterwards is unaffected."""
    from nbounce_amd import scenegen
    w, sb, env = small_gpu
    print("case: NaN camera, 32x32, 2 frames", flush=True)
    cam = nb.CameraController(world_position=scenegen.SMALL["eye"], target=scenegen.SMALL["target"], aspect_ratio=1.0)
    pt = nb.Pathtracer(ctx, sb, cam, env, 32, 32, resolution_factor=1.0, table=W.table(), sobol_stride=17)
    pt._camera_data = pt._camera_data.copy()
    pt._camera_data.view(np.float32)[:] = np.nan
    pt.dispatch()
    pt.dispatch()
    img = pt.output_texture()
    assert np.isnan(img[..., :3]).all() and (img[..., 3] == 1.0).all()
    print("case: a triangle of NaN vertices, 32x32, 4 frames", flush=True)
    v, i, pr = _two_triangles()
    v["position"][3:6] = np.nan
    blas, tlas, inst, idx = oracle.build_scene(v, i, pr)
    sb2 = nb.SceneBuffers(ctx, blas, tlas, inst, v, idx)
    env2 = nb.EnvMap(ctx, scenegen.sky_cube(32))
    got, ref = _frames_vs_oracle(nb, ctx, sb2, oracle.Scene(blas, tlas, inst, v, idx), env2, (32, 32),
                                 (0.0, 0.0, 4.0), (0.0, 0.0, 0.0), 4)
    _check(got, ref)
    sb2.close()
    print("case: valid render after the non-finite inputs", flush=True)
    got, ref = _frames_vs_oracle(nb, ctx, sb, W.oracle_scene(w), env, (32, 32), scenegen.SMALL["eye"],
                                 scenegen.SMALL["target"], 2)
    _check(got, ref)
