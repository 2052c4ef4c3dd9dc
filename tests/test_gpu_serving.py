"""GPU tests of the serving path (nb_server_*): each request's image equals the plain
renderer's and does not depend on the other requests in its steps; admission follows the HBM
budget; requests that arrive while others run join the running steps."""
import numpy as np
import pytest

import oracle
from tests import _workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(nb_mod=None):
    import nbounce_amd as nb
    from nbounce_amd import scenegen
    ctx = nb.Context(0)
    t = W.table()
    import ctypes
    from nbounce_amd._lib import lib, ptr, check
    check(lib.nb_sobol_set(ctx.handle, ptr(np.ascontiguousarray(t)), t.shape[0], t.shape[1], 17))
    small = W.small()
    sb_small = nb.SceneBuffers(ctx, small["blas"], small["tlas"], small["instances"], small["vertices"], small["indices"])
    w = W.large(0, 512)
    sc = nb.Scene(w["vertices"], w["indices_in"].copy(), w["primitives"])
    b, tl, inst = nb.build_scene(sc, share_blas=True)
    sb_large = nb.SceneBuffers(ctx, b, tl, inst, sc.vertices, sc.indices)
    faces = scenegen.sky_cube(32)
    e = nb.EnvMap(ctx, faces)
    # the reference-layout arrays of both scenes, for the CPU oracle
    _ORACLE.update(scenes=[oracle.Scene(small["blas"], small["tlas"], small["instances"], small["vertices"],
                                        small["indices"]),
                           oracle.Scene(b, tl, inst, sc.vertices, sc.indices)],
                   env=oracle.Env(faces), table=t)
    yield nb, ctx, [sb_small, sb_large], e
    ctx.close()


_ORACLE = {}


def _oracle_image(cam, scene, w, h, spp, bounces, cf=4.0):
    """The CPU oracle's progressive image of samples 1..spp (weights 1/k), pathtracing.wgsl:243-260."""
    img = np.zeros((h, w, 4), dtype=np.float32)
    for k in range(1, spp + 1):
        g = oracle.globals_(k, np.float32(1.0) / np.float32(k), bounces, cf)
        img, _ = oracle.render(_ORACLE["scenes"][scene], cam, g, _ORACLE["table"], 17, _ORACLE["env"], w, h, rgba=img,
                               threads=16)
    return img


def _matches_oracle(tag, got, ref, tol=1e-3):
    err = W.rel_err(got, ref)
    same = (got.view(np.uint32) == ref.view(np.uint32)).all(axis=-1).mean()
    print(f"{tag}: bit-identical pixels {same:.5f}, max rel err {err.max():.3e}", flush=True)
    assert np.isfinite(got).all() and err.max() <= tol


def _cam(nb, eye, target, w, h):
    return nb.CameraController(world_position=eye, target=target, aspect_ratio=w / h).data()


def _plain(nb, ctx, sb, e, cam, w, h, spp, bounces):
    from nbounce_amd._lib import lib, ptr, check
    import ctypes
    t = ctypes.c_void_p()
    check(lib.nb_target_create(ctx.handle, w, h, ctypes.byref(t)))
    g = nb.Globals(1, 0.0, bounces, 4.0).array()
    check(lib.nb_render_frames(ctx.handle, sb.handle, t, ptr(cam), ptr(g), spp, e.handle))
    img = np.zeros((h, w, 4), dtype=np.float32)
    check(lib.nb_target_read(t, ptr(img)))
    lib.nb_target_destroy(t)
    return img


def _run_to_completion(srv, max_steps=1000):
    done = {}
    for _ in range(max_steps):
        a, q = srv.step()
        for c in srv.poll():
            done[int(c["id"])] = c
        if a == 0 and q == 0:
            break
    return done


REQ = dict(scene=1, eye=(-20.0, 110.0, -20.0), target=(100.0, 95.0, 100.0), w=96, h=64, spp=5, bounces=8)


def test_request_equals_plain_render(env):
    from nbounce_amd.serving import Server
    nb, ctx, scenes, e = env
    print("case: one served request vs nb_render_frames", flush=True)
    srv = Server(ctx, max_paths=1 << 16, max_bounces=12)
    for s in scenes:
        srv.add_scene(s, e)
    cam = _cam(nb, REQ["eye"], REQ["target"], REQ["w"], REQ["h"])
    rid = srv.submit(cam, REQ["scene"], REQ["w"], REQ["h"], REQ["spp"], REQ["bounces"])
    done = _run_to_completion(srv)
    assert rid in done
    got = srv.read(rid, REQ["w"], REQ["h"])
    want = _plain(nb, ctx, scenes[1], e, cam, REQ["w"], REQ["h"], REQ["spp"], REQ["bounces"])
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    st = srv.stats()
    assert st["samples"] == REQ["w"] * REQ["h"] * REQ["spp"]
    srv.close()


def test_request_is_the_same_alone_and_in_a_mix(env):
    from nbounce_amd.serving import Server
    nb, ctx, scenes, e = env
    print("case: request alone vs in a mix of 8 requests over 2 scenes", flush=True)
    cam = _cam(nb, REQ["eye"], REQ["target"], REQ["w"], REQ["h"])
    alone = Server(ctx, max_paths=1 << 16, max_bounces=12)
    for s in scenes:
        alone.add_scene(s, e)
    rid = alone.submit(cam, REQ["scene"], REQ["w"], REQ["h"], REQ["spp"], REQ["bounces"])
    _run_to_completion(alone)
    img_alone = alone.read(rid, REQ["w"], REQ["h"])
    alone.close()

    mix = Server(ctx, max_paths=1 << 15, max_bounces=12)
    for s in scenes:
        mix.add_scene(s, e)
    rng = np.random.default_rng(3)
    others = []
    for k in range(7):
        w, h = [(64, 64), (128, 96), (32, 32)][k % 3]
        sc = int(k % 2)
        eye = (5.0, 5.0, 9.5) if sc == 0 else tuple(rng.uniform(0, 200, 3))
        tgt = (5.0, 4.0, 0.0) if sc == 0 else tuple(rng.uniform(0, 200, 3))
        c_k = _cam(nb, eye, tgt, w, h)
        spp, bounces = int(rng.integers(1, 9)), int(rng.integers(1, 13))
        others.append((mix.submit(c_k, sc, w, h, spp, bounces), c_k, sc, w, h, spp, bounces))
        if k == 3:
            rid2 = mix.submit(cam, REQ["scene"], REQ["w"], REQ["h"], REQ["spp"], REQ["bounces"])
            mix.step()     # the request arrives while earlier ones are being rendered
    done = _run_to_completion(mix)
    assert rid2 in done and all(o[0] in done for o in others)
    img_mix = mix.read(rid2, REQ["w"], REQ["h"])
    assert mix.stats()["steps"] > 2
    # every request of the mix (both scenes share steps, so waves move between scenes) equals the
    # plain renderer's image of it
    img_mixed = {}
    for oid, c_k, sc, w, h, spp, bounces in others:
        img_mixed[oid] = mix.read(oid, w, h)
        np.testing.assert_array_equal(img_mixed[oid].view(np.uint32),
                                      _plain(nb, ctx, scenes[sc], e, c_k, w, h, spp, bounces).view(np.uint32))
    mix.close()
    np.testing.assert_array_equal(img_alone.view(np.uint32), img_mix.view(np.uint32))
    # anchored to the CPU oracle: the request alone (= in the mix) and two of the mixed requests,
    # one per scene, against the oracle's rendering of the same samples
    _matches_oracle("request alone / in the mix vs oracle", img_alone,
                    _oracle_image(cam, REQ["scene"], REQ["w"], REQ["h"], REQ["spp"], REQ["bounces"]))
    for sc_want in (0, 1):
        oid, c_k, sc, w, h, spp, bounces = next(o for o in others if o[2] == sc_want)
        _matches_oracle(f"mixed request on scene {sc} ({w}x{h}, {spp} spp, {bounces} bounces) vs oracle",
                        img_mixed[oid], _oracle_image(c_k, sc, w, h, spp, bounces))


def _resident_bytes(ctx, scenes, e, max_paths):
    """Bytes a server with these resident scenes charges before any request."""
    from nbounce_amd.serving import Server
    probe = Server(ctx, max_paths=max_paths, max_bounces=8)
    for s in scenes:
        probe.add_scene(s, e)
    r = probe.stats()["bytes_in_use"]
    probe.close()
    return r


def test_admission_waits_for_hbm_budget(env):
    from nbounce_amd.serving import Server
    nb, ctx, scenes, e = env
    print("case: admission under a budget of the resident data plus 1.5 images", flush=True)
    w = h = 64
    resident = _resident_bytes(ctx, scenes[:1], e, 1 << 16)
    srv = Server(ctx, hbm_budget_bytes=resident + int(1.5 * w * h * 16), max_paths=1 << 16, max_bounces=8)
    srv.add_scene(scenes[0], e)
    cam = _cam(nb, (5.0, 5.0, 9.5), (5.0, 4.0, 0.0), w, h)
    ids = [srv.submit(cam, 0, w, h, 1, 8) for _ in range(3)]
    a, q = srv.step()
    assert (a, q) == (0, 2)                     # first admitted and finished, two wait
    first = srv.poll()
    assert [int(c["id"]) for c in first] == ids[:1]
    a, q = srv.step()
    assert (a, q) == (0, 2) and len(srv.poll()) == 0   # its image still holds the budget
    srv.release(ids[0])
    srv.step()
    assert [int(c["id"]) for c in srv.poll()] == ids[1:2]
    with pytest.raises(nb.NbError):
        srv.submit(_cam(nb, (5, 5, 9.5), (5, 4, 0), 128, 128), 0, 128, 128, 1, 8)   # larger than the budget
    srv.close()


def test_admission_counts_resident_scenes_and_step_buffers(env):
    """The budget covers the server's step buffers, the resident scenes and environments and the
    request images: a scene that does not fit is refused, a request that cannot fit beside the
    resident data is refused, and the step buffers alone can exceed a budget."""
    from nbounce_amd.serving import Server
    nb, ctx, scenes, e = env
    print("case: admission charges resident scenes", flush=True)
    own = _resident_bytes(ctx, [], e, 1 << 16)
    with_scene = _resident_bytes(ctx, scenes[1:], e, 1 << 16)
    assert with_scene - own >= scenes[1].device_bytes() > 0
    with pytest.raises(nb.NbError) as ex:
        Server(ctx, hbm_budget_bytes=own // 2, max_paths=1 << 16, max_bounces=8)
    assert ex.value.status == 6   # NB_ERR_BUDGET
    srv = Server(ctx, hbm_budget_bytes=(own + with_scene) // 2, max_paths=1 << 16, max_bounces=8)
    with pytest.raises(nb.NbError) as ex:
        srv.add_scene(scenes[1], e)
    assert ex.value.status == 6
    srv.close()
    w = h = 64
    srv = Server(ctx, hbm_budget_bytes=with_scene + w * h * 16 // 2, max_paths=1 << 16, max_bounces=8)
    srv.add_scene(scenes[1], e)
    cam = _cam(nb, REQ["eye"], REQ["target"], w, h)
    with pytest.raises(nb.NbError) as ex:
        srv.submit(cam, 0, w, h, 1, 8)    # half an image is left beside the scene
    assert ex.value.status == 6
    rid = srv.submit(cam, 0, 32, 32, 1, 8)
    done = _run_to_completion(srv)
    assert rid in done and srv.stats()["bytes_in_use"] == with_scene + 32 * 32 * 16
    srv.close()


def test_small_requests_overtake_a_long_one(env):
    from nbounce_amd.serving import Server
    nb, ctx, scenes, e = env
    print("case: continuous batching, small requests join a long one", flush=True)
    srv = Server(ctx, max_paths=1 << 17, max_bounces=8)
    srv.add_scene(scenes[0], e)
    big = srv.submit(_cam(nb, (5.0, 5.0, 9.5), (5.0, 4.0, 0.0), 256, 256), 0, 256, 256, 32, 8)
    srv.step()
    small = [srv.submit(_cam(nb, (5.0, 5.0, 9.5), (5.0, 4.0, 0.0), 32, 32), 0, 32, 32, 2, 8) for _ in range(3)]
    order = []
    for _ in range(200):
        a, q = srv.step()
        order += [int(c["id"]) for c in srv.poll()]
        if a == 0 and q == 0:
            break
    assert set(order) == set(small) | {big}
    assert order.index(big) == len(order) - 1
    srv.close()


def _cached_reference(nb, ctx, sb, e, cache, cam, w, h, spp, bounces):
    """nb_render_cached for samples 1..spp, weight 1/sample, on a fresh target."""
    import ctypes
    from nbounce_amd._lib import check, lib, ptr
    t = ctypes.c_void_p()
    check(lib.nb_target_create(ctx.handle, w, h, ctypes.byref(t)))
    for k in range(1, spp + 1):
        g = nb.Globals(k, float(np.float32(1.0) / np.float32(k)), bounces, 4.0).array()
        check(lib.nb_render_cached(ctx.handle, sb.handle, t, ptr(cam), ptr(g), e.handle, cache.handle, 0, 0.0, None))
    img = np.zeros((h, w, 4), dtype=np.float32)
    check(lib.nb_target_read(t, ptr(img)))
    lib.nb_target_destroy(t)
    return img


@pytest.mark.parametrize("fp16", [False, True])
def test_cached_request_equals_render_cached_and_plain_ones_are_unchanged(env, fp16):
    """A cache-assisted request (inference only) equals nb_render_cached of its samples on the
    scene's cache, and plain requests served in the same steps still equal nb_render_frames."""
    from nbounce_amd.cache import RadianceCache
    from nbounce_amd.serving import Server
    nb, ctx, scenes, e = env
    print(f"case: cached + plain requests in one server, {'fp16' if fp16 else 'bf16'} cache", flush=True)
    cache = RadianceCache(ctx, [0, 0, 0], [10, 10, 10], seed=4, fp16=fp16)
    srv = Server(ctx, max_paths=1 << 16, max_bounces=12)
    for s in scenes:
        srv.add_scene(s, e)
    srv.set_cache(0, cache)
    w, h = 64, 48
    cam0 = _cam(nb, (5.0, 5.0, 9.5), (5.0, 4.0, 0.0), w, h)
    cam1 = _cam(nb, REQ["eye"], REQ["target"], REQ["w"], REQ["h"])
    rc = srv.submit(cam0, 0, w, h, 3, 8, cached=True)
    rp = srv.submit(cam1, REQ["scene"], REQ["w"], REQ["h"], REQ["spp"], REQ["bounces"])
    with pytest.raises(nb.NbError):
        srv.submit(cam1, 1, 32, 32, 1, 8, cached=True)      # scene 1 has no cache
    done = _run_to_completion(srv)
    assert rc in done and rp in done
    got_c = srv.read(rc, w, h)
    got_p = srv.read(rp, REQ["w"], REQ["h"])
    st = srv.stats()
    assert st["cached_samples"] == 3 * w * h and st["cached_busy_s"] > 0
    srv.close()
    want_c = _cached_reference(nb, ctx, scenes[0], e, cache, cam0, w, h, 3, 8)
    want_p = _plain(nb, ctx, scenes[1], e, cam1, REQ["w"], REQ["h"], REQ["spp"], REQ["bounces"])
    assert np.isfinite(got_c).all() and got_c[..., :3].max() > 0
    np.testing.assert_array_equal(got_c.view(np.uint32), want_c.view(np.uint32))
    np.testing.assert_array_equal(got_p.view(np.uint32), want_p.view(np.uint32))
    cache.close()
