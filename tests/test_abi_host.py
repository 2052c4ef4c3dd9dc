"""Host-side checks of libnbounce through its C ABI (no GPU needed): the exported symbols,
the bit-exact scene builder, the Sobol table and the camera against the oracle, and the
argument validation of the entry points."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import oracle
from nbounce_amd import _lib, layout, scenegen
import nbounce_amd as nb
from tests import _workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "nbounce.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nb_[a-z0-9_]+)\s*\(", text)))


def test_package_refuses_to_import_without_the_library(tmp_path):
    """The library is the only compute path: with no libnbounce.so to load, importing the
    package fails (there is no CPU fallback)."""
    import subprocess
    import sys
    env = dict(os.environ, NB_LIBRARY=str(tmp_path / "absent" / "libnbounce.so"))
    r = subprocess.run([sys.executable, "-c", "import nbounce_amd"], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "libnbounce.so" in r.stderr


def test_library_exports_every_declared_symbol():
    names = _declared_symbols()
    assert len(names) >= 25
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(_lib.SIGNATURES) | {"nb_last_error", "nb_version", "nb_cache_param_count"}


def test_struct_layouts_match_the_reference_sizes():
    assert layout.NODE.itemsize == 32        # bvh.rs:15-24
    assert layout.VERTEX.itemsize == 48      # scene.rs:13-21
    assert layout.INSTANCE.itemsize == 160   # scene.rs:220-230
    assert layout.CAMERA.itemsize == 128     # camera.rs:7-12
    assert layout.GLOBALS.itemsize == 16     # pathtracer.rs:26-33
    assert layout.NODE == oracle.NODE and layout.VERTEX == oracle.VERTEX and layout.INSTANCE == oracle.INSTANCE


@pytest.mark.parametrize("seed", [0, 1])
def test_builder_is_bit_exact_with_the_oracle_small(seed):
    v, i, pr = scenegen.spheres_in_box(seed)
    sc = nb.Scene(v, i.copy(), pr)
    blas, tlas, inst = nb.build_scene(sc)
    ob, ot, oi, oidx = oracle.build_scene(v, i, pr)
    assert blas.tobytes() == ob.tobytes()
    assert tlas.tobytes() == ot.tobytes()
    assert inst.tobytes() == oi.tobytes()
    assert np.array_equal(sc.indices, oidx)


def test_builder_is_bit_exact_with_the_oracle_rotated_instances():
    v, i, pr = scenegen.instanced_meshes(3, n_instances=96, extent=40.0)
    sc = nb.Scene(v, i.copy(), pr)
    blas, tlas, inst = nb.build_scene(sc)
    ob, ot, oi, oidx = oracle.build_scene(v, i, pr)
    assert blas.tobytes() == ob.tobytes() and tlas.tobytes() == ot.tobytes()
    assert inst.tobytes() == oi.tobytes() and np.array_equal(sc.indices, oidx)


@pytest.mark.parametrize("case", range(W.RANDOM_CASES))
def test_builder_is_bit_exact_with_the_oracle_random(case):
    """The seeded random scenes of the randomised GPU parity cases (spheres in boxes of random
    size, random numbers of rotated and scaled instances): the library's builder and the
    oracle's produce the same bytes."""
    v, i, pr = W.random_case(case)["vit"]
    sc = nb.Scene(v, i.copy(), pr)
    blas, tlas, inst = nb.build_scene(sc)
    ob, ot, oi, oidx = oracle.build_scene(v, i, pr)
    assert blas.tobytes() == ob.tobytes() and tlas.tobytes() == ot.tobytes()
    assert inst.tobytes() == oi.tobytes() and np.array_equal(sc.indices, oidx)


def test_shared_blas_build_is_the_first_reference_build():
    """NB_BUILD_SHARE_BLAS builds each index range once; those trees equal the first tree the
    reference builds over that range (later reference rebuilds may differ in a few leaves)."""
    v, i, pr = scenegen.spheres_in_box(0)
    shared = nb.Scene(v, i.copy(), pr)
    sb, st, si = nb.build_scene(shared, share_blas=True)
    ref = nb.Scene(v, i.copy(), pr)
    rb, rt, ri = nb.build_scene(ref)
    assert len(set(si["node"].tolist())) == 2                       # box + one sphere tree
    sphere_root = max(si["node"])
    n = len(sb) - sphere_root
    ref_roots = sorted(set(ri["node"].tolist()))
    first = [r for r in ref_roots if r != 0][0]
    a, b = sb[sphere_root:].copy(), rb[first:first + n].copy()
    a["start"] = np.where(a["end"] == 0, a["start"] - sphere_root, a["start"])
    b["start"] = np.where(b["end"] == 0, b["start"] - first, b["start"])
    assert a.tobytes() == b.tobytes()
    assert np.array_equal(st, rt)


def test_sobol_table_equals_the_oracle():
    assert np.array_equal(nb.sobol_table(1024, 17, 0), oracle.sobol_table(1024, 17, 0))
    assert np.array_equal(nb.sobol_table(64, 25, 7), oracle.sobol_table(64, 25, 7))


def test_camera_equals_the_oracle():
    for eye, target, aspect in [((5, 0, 0), (0, 0, 0), 1.0), ((5, 5, 9.5), (5, 4, 0), 1.0),
                                ((-20, 110, -20), (100, 95, 100), 1920 / 1080)]:
        cam = nb.CameraController(world_position=eye, target=target, aspect_ratio=aspect).data()
        ref = oracle.camera(eye, target, (0, 1, 0), math.pi / 3, aspect, 0.1)
        assert cam.tobytes() == ref.tobytes()


def test_builder_rejects_bad_input():
    v, i, pr = scenegen.spheres_in_box(0)
    bad = pr.copy()
    bad["index_end"][3] = len(i) + 3
    with pytest.raises(nb.NbError):
        nb.build_scene(nb.Scene(v, i.copy(), bad))
    bad_idx = i.copy()
    bad_idx[5] = len(v) + 10
    with pytest.raises(nb.NbError):
        nb.build_scene(nb.Scene(v, bad_idx, pr))


def test_error_message_is_thread_local_text():
    bcap, tcap = ctypes.c_uint32(), ctypes.c_uint32()
    st = _lib.lib.nb_build_scene_capacity(0, None, 0, 0, ctypes.byref(bcap), ctypes.byref(tcap))
    assert st != 0
    assert b"null" in _lib.lib.nb_last_error()
