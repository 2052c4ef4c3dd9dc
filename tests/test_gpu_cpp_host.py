"""The C++ host mirror (include/nbounce.hpp) rendering on the GPU through its check program: the
reference's spheres.glb (Scene::parse_gltf, SceneBuffers::from_scene with the reference build),
a BC6H cube DDS sky (EnvMap::load) and Pathtracer's dispatch / dispatch_frames over 4 frames at
96x64 give the oracle's image; the program also checks the sample count, the output size and
that dispatch() does nothing beyond max_sample_count.  BlitRenderer's surface of that image
equals oracle/display.py's blit bit for bit, and write_pfm / write_exr write the image."""
import os
import subprocess

import numpy as np
import pytest

import nbounce_amd as nb
import oracle
from nbounce_amd import scenegen
from oracle import display
from nbounce_amd.pathtracer import DDS_BC6H_UF16, read_dds_cube, write_dds_cube
from tests import _workloads as W
from tests.test_image_io import read_exr, read_pfm

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "check_pathtracer")
GLB = os.path.join(ROOT, "tests", "golden", "spheres.glb")
EYE, TARGET = (0.0, 1.0, 0.9), (0.0, 0.0, -0.5)
WIDTH, HEIGHT, FRAMES = 96, 64, 4


def test_cpp_pathtracer_matches_oracle(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/check_pathtracer is not built: run __graft_entry__.build()")
    dds = str(tmp_path / "sky.dds")
    write_dds_cube(dds, scenegen.sky_cube(32), DDS_BC6H_UF16)
    out = str(tmp_path / "image.raw")
    args = [EXE, "render", GLB, dds, str(WIDTH), str(HEIGHT), str(FRAMES)] + [repr(x) for x in EYE + TARGET] + [out]
    print("case: spheres.glb through the C++ Pathtracer, BC6H sky, 4 frames 96x64", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "FAIL" not in r.stdout
    got = np.fromfile(out, dtype=np.float32).reshape(HEIGHT, WIDTH, 4)
    sc = nb.Scene().parse_gltf(GLB)
    blas, tlas, inst = nb.build_scene(sc)
    osc = oracle.Scene(blas, tlas, inst, sc.vertices, sc.indices)
    cam = nb.CameraController(world_position=EYE, target=TARGET, aspect_ratio=WIDTH / HEIGHT).data()
    oenv = oracle.Env(read_dds_cube(dds))
    ref = None
    for k in range(1, FRAMES + 1):
        g = oracle.globals_(k, np.float32(1) / np.float32(k), 8, 4.0)
        ref, _ = oracle.render(osc, cam, g, W.table(), 17, oenv, WIDTH, HEIGHT, rgba=ref, threads=8)
    assert (ref[..., :3].sum(-1) > 0).mean() > 0.5
    same = (got.view(np.uint32) == ref.view(np.uint32)).all(-1).mean()
    err = W.rel_err(got, ref)
    print(f"bit-identical pixels {same:.5f}, max rel err {err.max():.3e}")
    assert err.max() <= 1e-3
    # BlitRenderer's surface of the image (oracle/display.py's blit, bit for bit) and the files
    surface = np.fromfile(out + ".surface", dtype=np.uint16).reshape(HEIGHT + 8, WIDTH + 24, 4)
    want = display.blit(got, WIDTH + 24, HEIGHT + 8)
    np.testing.assert_array_equal(surface, want.view(np.uint16))
    np.testing.assert_array_equal(read_pfm(out + ".pfm").view(np.uint32), got[..., :3].view(np.uint32))
    chans, _ = read_exr(out + ".exr")
    for k, c in enumerate("RGBA"):
        np.testing.assert_array_equal(chans[c].view(np.uint32), np.ascontiguousarray(got[::-1, :, k]).view(np.uint32))
