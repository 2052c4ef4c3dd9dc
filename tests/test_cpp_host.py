"""The C++ host mirror (include/nbounce.hpp) through its check program, without a GPU: the
reference's spheres.glb read twice (the second copy appended), the reference build and a camera
buffer equal the Python mirror's byte for byte, and the build equals the oracle's builder; a
missing file and a file that is not glTF throw NB_ERR_IO and NB_ERR_FORMAT and leave the scene
unchanged (checked inside the program)."""
import os
import subprocess

import numpy as np
import pytest

import nbounce_amd as nb
import oracle
from nbounce_amd import layout

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "check_pathtracer")
GLB = os.path.join(ROOT, "tests", "golden", "spheres.glb")


def _read(path, dtype):
    return np.fromfile(path, dtype=dtype)


def test_cpp_host_mirror_equals_python_mirror_and_oracle(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/check_pathtracer is not built: run __graft_entry__.build()")
    r = subprocess.run([EXE, "host", GLB, str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "FAIL" not in r.stdout
    sc = nb.Scene().parse_gltf(GLB).parse_gltf(GLB)
    v0, i0, p0 = sc.vertices.copy(), sc.indices.copy(), sc.primitives.copy()
    assert _read(tmp_path / "vertices.bin", layout.VERTEX).tobytes() == v0.tobytes()
    assert _read(tmp_path / "primitives.bin", layout.PRIMITIVE).tobytes() == p0.tobytes()
    blas, tlas, inst = nb.build_scene(sc)
    assert _read(tmp_path / "blas.bin", layout.NODE).tobytes() == blas.tobytes()
    assert _read(tmp_path / "tlas.bin", layout.NODE).tobytes() == tlas.tobytes()
    assert _read(tmp_path / "instances.bin", layout.INSTANCE).tobytes() == inst.tobytes()
    assert np.array_equal(_read(tmp_path / "indices.bin", np.uint32), sc.indices)
    ob, ot, oi, oidx = oracle.build_scene(v0, i0, p0)
    assert ob.tobytes() == blas.tobytes() and ot.tobytes() == tlas.tobytes() and oi.tobytes() == inst.tobytes()
    assert np.array_equal(oidx, sc.indices)
    cam = nb.CameraController(world_position=(0.0, 1.0, 0.9), target=(0.0, 0.0, -0.5), aspect_ratio=1.5).data()
    assert _read(tmp_path / "camera.bin", layout.CAMERA).tobytes() == cam.tobytes()
