"""An environment map loaded from a BC6H cube DDS (nb_envmap_load, EnvMap::load in
src/pathtracing/envmap.rs:12) renders exactly like the same decoded faces uploaded with
nb_envmap_create_cube, and like the CPU oracle given those faces (pathtracing.wgsl:166-169)."""
import numpy as np
import pytest

import oracle
from tests import _workloads as W

pytestmark = pytest.mark.gpu


def test_bc6h_envmap_renders_like_its_decoded_faces(tmp_path):
    import nbounce_amd as nb
    from nbounce_amd import scenegen
    print("case: BC6H_UF16 sky cube from a DDS file, small scene 64x64, 3 frames", flush=True)
    path = str(tmp_path / "sky_bc6h.dds")
    nb.write_dds_cube(path, scenegen.sky_cube(64), nb.DDS_BC6H_UF16)
    ctx = nb.Context(0)
    w = W.small()
    sb = nb.SceneBuffers(ctx, w["blas"], w["tlas"], w["instances"], w["vertices"], w["indices"])
    loaded = nb.EnvMap.load(ctx, path)
    faces = nb.read_dds_cube(path)
    assert np.array_equal(loaded.faces, faces) and faces.shape == (6, 64, 64, 4)
    direct = nb.EnvMap(ctx, faces)
    P = scenegen.SMALL
    cam = nb.CameraController(world_position=P["eye"], target=P["target"], aspect_ratio=1.0)
    imgs = []
    for env in (loaded, direct):
        pt = nb.Pathtracer(ctx, sb, cam, env, 64, 64, resolution_factor=1.0, table=W.table(), sobol_stride=17)
        pt.dispatch_frames(3)
        imgs.append(pt.output_texture())
    np.testing.assert_array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
    ref = np.zeros((64, 64, 4), dtype=np.float32)
    for k in range(1, 4):
        g = oracle.globals_(k, np.float32(1.0) / np.float32(k), 8, 4.0)
        ref, _ = oracle.render(W.oracle_scene(w), cam.data(), g, W.table(), 17, oracle.Env(faces), 64, 64, rgba=ref,
                               threads=8)
    err = W.rel_err(imgs[0], ref)
    same = (imgs[0].view(np.uint32) == ref.view(np.uint32)).all(-1).mean()
    print(f"vs oracle: bit-identical pixels {same:.4f}, max rel err {err.max():.3e}", flush=True)
    assert err.max() <= 1e-3
    ctx.close()
