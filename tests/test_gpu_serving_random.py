"""Randomised serving parity: seeded random requests (sizes, samples, bounce limits, contribution
factors, cameras) over four random resident scenes, each with its own sky, served with the
smallest path budget a server takes (4096 paths per step), so that requests split their samples
over several steps and share each step with requests on other scenes.  Every served image
equals the CPU oracle's progressive image of the same request (samples 1..spp, weights 1/k)."""
import numpy as np
import pytest

import oracle
from nbounce_amd import scenegen
from tests import _workloads as W

pytestmark = pytest.mark.gpu

REL_TOL = 1e-3
SCENE_CASES = (0, 1, 6, 9)     # spheres and instanced meshes, cube and lat-long skies
REQUESTS = 14
MAX_PATHS = 4096


@pytest.fixture(scope="module")
def world():
    import nbounce_amd as nb
    from nbounce_amd._lib import lib, ptr, check
    ctx = nb.Context(0)
    t = W.table()
    check(lib.nb_sobol_set(ctx.handle, ptr(np.ascontiguousarray(t)), t.shape[0], t.shape[1], 17))
    scenes = []
    for case in SCENE_CASES:
        c = W.random_case(case)
        v, i, pr = c["vit"]
        blas, tlas, inst, idx = oracle.build_scene(v, i, pr)
        if c["latlong"]:
            sky = scenegen.sky_latlong(64, 32)
            env, oenv = nb.EnvMap.latlong(ctx, sky), oracle.Env.latlong(sky)
        else:
            faces = scenegen.sky_cube(32)
            env, oenv = nb.EnvMap(ctx, faces), oracle.Env(faces)
        sb = nb.SceneBuffers(ctx, blas, tlas, inst, v, idx)
        scenes.append(dict(case=c, sb=sb, env=env, osc=oracle.Scene(blas, tlas, inst, v, idx), oenv=oenv))
    yield nb, ctx, scenes, t
    ctx.close()


def test_random_requests_match_oracle(world):
    from nbounce_amd.serving import Server
    nb, ctx, scenes, table = world
    rng = np.random.default_rng(4242)
    srv = Server(ctx, max_paths=MAX_PATHS, max_bounces=12)
    for s in scenes:
        srv.add_scene(s["sb"], s["env"])
    reqs = []
    for r in range(REQUESTS):
        k = int(rng.integers(0, len(scenes)))
        c = scenes[k]["case"]
        w, h = [(8, 8), (16, 16), (24, 16), (32, 32), (48, 32)][int(rng.integers(0, 5))]
        spp, bounces, cf = int(rng.integers(1, 5)), int(rng.integers(0, 13)), float(rng.uniform(0.5, 8.0))
        eye = np.asarray(c["eye"]) + rng.normal(0.0, 0.5, 3)
        cam = nb.CameraController(world_position=tuple(float(x) for x in eye), target=c["target"],
                                  aspect_ratio=w / h).data()
        rid = srv.submit(cam, k, w, h, spp, bounces, contribution_factor=cf)
        reqs.append(dict(id=rid, scene=k, w=w, h=h, spp=spp, bounces=bounces, cf=cf, cam=cam))
    total = sum(r["w"] * r["h"] * r["spp"] for r in reqs)
    print(f"case: {REQUESTS} random requests over 4 random scenes, {total} paths, {MAX_PATHS} paths per step",
          flush=True)
    done = {}
    for _ in range(5000):
        a, q = srv.step()
        for comp in srv.poll():
            done[int(comp["id"])] = comp
        if a == 0 and q == 0:
            break
    assert all(r["id"] in done for r in reqs)
    st = srv.stats()
    assert st["steps"] >= -(-total // MAX_PATHS) and st["samples"] == total
    worst = 0.0
    for r in reqs:
        got = srv.read(r["id"], r["w"], r["h"])
        s = scenes[r["scene"]]
        ref = None
        for k in range(1, r["spp"] + 1):
            g = oracle.globals_(k, np.float32(1.0) / np.float32(k), r["bounces"], r["cf"])
            ref, _ = oracle.render(s["osc"], r["cam"], g, table, 17, s["oenv"], r["w"], r["h"], rgba=ref, threads=8)
        err = W.rel_err(got, ref)
        same = (got.view(np.uint32) == ref.view(np.uint32)).all(-1).mean()
        print(f"request {r['id']}: scene {r['scene']}, {r['w']}x{r['h']}, {r['spp']} spp, {r['bounces']} bounces: "
              f"bit-identical {same:.5f}, max rel err {err.max():.3e}", flush=True)
        assert np.isfinite(got).all()
        worst = max(worst, float(err.max()))
    srv.close()
    assert worst <= REL_TOL
