"""glTF ingestion (Scene::parse_gltf, src/pathtracing/scene.rs:108-217) and the .glb writer:
the library's native reader against the Python oracle (oracle/gltf.py) on the reference's own
assets and on generated documents, the oracle against what the assets and the mathematics fix,
and the reference's error behaviour."""
import base64
import json
import math
import os
import struct

import numpy as np
import pytest

import nbounce_amd as nb
from nbounce_amd import gltf, layout, scenegen
from oracle import gltf as og

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ASSETS = ["spheres.glb", "testscene.glb"]


def _read(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return f.read()


def _parse_bytes(data, tmp_path, share=False, name="scene.glb"):
    path = tmp_path / name
    path.write_bytes(data)
    return nb.Scene().parse_gltf(path, share_geometry=share)


def _split(data):
    jlen = struct.unpack_from("<I", data, 12)[0]
    doc = json.loads(data[20:20 + jlen])
    off = 20 + jlen
    blen = struct.unpack_from("<I", data, off)[0]
    return doc, bytearray(data[off + 8:off + 8 + blen])


def _join(doc, binary):
    js = json.dumps(doc).encode()
    js += b" " * ((-len(js)) % 4)
    binary = bytes(binary) + b"\0" * ((-len(binary)) % 4)
    doc_total = 12 + 8 + len(js) + 8 + len(binary)
    return (struct.pack("<4sII", b"glTF", 2, doc_total) + struct.pack("<II", len(js), 0x4E4F534A) + js
            + struct.pack("<II", len(binary), 0x004E4942) + binary)


def _same(scene, ref):
    v, i, p = ref
    assert scene.vertices.tobytes() == v.tobytes()
    assert scene.indices.tobytes() == i.tobytes()
    assert scene.primitives.tobytes() == p.tobytes()


# ---- the oracle against what the assets and the mathematics fix ---------------------------------

def test_oracle_spheres_asset_shape():
    v, i, p = og.parse_gltf(_read("spheres.glb"))
    assert len(p) == 15 and len(i) == 3 * 19200
    doc, _ = _split(_read("spheres.glb"))
    for node, prim in zip(doc["nodes"], p):
        m = prim["local_to_world"].reshape(4, 4)          # columns
        t = np.array(node.get("translation", [0, 0, 0]), dtype=np.float32)
        assert np.array_equal(m[:3, :3], np.eye(3, dtype=np.float32))
        assert np.array_equal(m[3, :3], t) and m[3, 3] == 1.0
    assert np.all(p["emissive"] == 0.0)
    # every primitive's range indexes its own 729-vertex block (icospheres, one mesh each)
    for prim in p:
        tri = i[prim["index_start"]:prim["index_end"]]
        assert tri.max() - tri.min() < 729


def test_oracle_testscene_asset_shape():
    v, i, p = og.parse_gltf(_read("testscene.glb"))
    assert len(p) == 4 and len(i) == 3 * 32460
    doc, _ = _split(_read("testscene.glb"))
    mats = doc["materials"]
    for node, prim in zip(doc["nodes"], p):
        mesh = doc["meshes"][node["mesh"]]
        mat = mats[mesh["primitives"][0]["material"]]["pbrMetallicRoughness"]
        assert prim["roughness"] == np.float32(mat.get("roughnessFactor", 1.0))
        assert prim["metallic"] == np.float32(mat.get("metallicFactor", 1.0))   # glTF default 1


def _rotation64(q):
    x, y, z, w = np.asarray(q, dtype=np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_oracle_trs_matrix_closed_forms():
    rng = np.random.default_rng(3)
    for _ in range(50):
        q = rng.normal(size=4)
        q = (q / np.linalg.norm(q)).astype(np.float32)
        t = rng.uniform(-5, 5, 3).astype(np.float32)
        s = rng.uniform(0.2, 3, 3).astype(np.float32)
        m = og.trs_matrix(t, q, s).reshape(4, 4).T.astype(np.float64)     # row-major
        want = _rotation64(q.astype(np.float64)) * s.astype(np.float64)[None, :]
        assert np.allclose(m[:3, :3], want, atol=2e-6 * s.max())
        assert np.array_equal(m[:3, 3], t.astype(np.float64)) and np.array_equal(m[3], [0, 0, 0, 1])
    # exact cases: half turns about the axes and the identity
    assert np.array_equal(og.trs_matrix([0, 0, 0], [0, 1, 0, 0], [1, 1, 1]).reshape(4, 4).T,
                          np.diag([-1, 1, -1, 1]).astype(np.float32))
    assert np.array_equal(og.trs_matrix([0, 0, 0], [1, 0, 0, 0], [2, 2, 2]).reshape(4, 4).T,
                          np.diag([2, -2, -2, 1]).astype(np.float32))
    assert np.array_equal(og.trs_matrix([1, 2, 3], [0, 0, 0, 1], [1, 1, 1]).reshape(4, 4).T[:3, 3], [1, 2, 3])


def test_oracle_testscene_matrices_are_its_trs():
    doc, _ = _split(_read("testscene.glb"))
    _, _, p = og.parse_gltf(_read("testscene.glb"))
    for node, prim in zip(doc["nodes"], p):
        m = prim["local_to_world"].reshape(4, 4).T.astype(np.float64)
        want = np.eye(4)
        want[:3, :3] = _rotation64(node.get("rotation", [0, 0, 0, 1])) * np.array(node.get("scale", [1, 1, 1]))
        want[:3, 3] = node.get("translation", [0, 0, 0])
        assert np.allclose(m, want, atol=1e-6)


# ---- the native reader against the oracle --------------------------------------------------------

@pytest.mark.parametrize("name", ASSETS)
def test_reader_equals_oracle_on_reference_assets(name):
    _same(nb.Scene().parse_gltf(os.path.join(GOLDEN, name)), og.parse_gltf(_read(name)))


@pytest.mark.parametrize("name", ASSETS)
def test_reader_shared_geometry_on_reference_assets(name):
    """spheres.glb has one accessor set per mesh, so sharing changes nothing; testscene.glb's two
    Suzanne meshes name the same accessors with different materials, which the reference reads
    as two copies and sharing reads as one range used by both primitives."""
    shared = nb.Scene().parse_gltf(os.path.join(GOLDEN, name), share_geometry=True)
    _same(shared, og.parse_gltf(_read(name), share_geometry=True))
    v, i, p = og.parse_gltf(_read(name))
    ranges = {(int(a), int(b)) for a, b in zip(shared.primitives["index_start"], shared.primitives["index_end"])}
    assert len(ranges) == (15 if name == "spheres.glb" else 3)
    if name == "spheres.glb":
        _same(shared, (v, i, p))
    for a, b in zip(shared.primitives, p):
        ta = shared.vertices[shared.indices[a["index_start"]:a["index_end"]]]
        tb = v[i[b["index_start"]:b["index_end"]]]
        assert ta.tobytes() == tb.tobytes()
        assert a["local_to_world"].tobytes() == b["local_to_world"].tobytes()
        assert a["color"].tobytes() == b["color"].tobytes()


def test_parses_append_like_the_reference():
    """A second parse continues the vertex and index numbering (scene.rs:135-136, 172)."""
    s = nb.Scene().parse_gltf(os.path.join(GOLDEN, "spheres.glb")).parse_gltf(os.path.join(GOLDEN, "testscene.glb"))
    v1, i1, p1 = og.parse_gltf(_read("spheres.glb"))
    v2, i2, p2 = og.parse_gltf(_read("testscene.glb"))
    p2 = p2.copy()
    p2["index_start"] += len(i1)
    p2["index_end"] += len(i1)
    _same(s, (np.concatenate([v1, v2]), np.concatenate([i1, i2 + np.uint32(len(v1))]), np.concatenate([p1, p2])))


@pytest.mark.parametrize("make, exact", [(lambda: scenegen.spheres_in_box(2, n_spheres=12), True),
                                         (lambda: scenegen.instanced_meshes(1, n_instances=200, extent=30.0), True),
                                         (lambda: scenegen.instanced_meshes(1, n_instances=5, extent=30.0), False)])
def test_writer_round_trip(make, exact, tmp_path):
    """Every index range in use: the arrays come back exactly.  Unused base meshes are not
    written, so then each primitive's triangles come back instead."""
    v, i, p = make()
    scene = nb.Scene(v, i, p)
    used = np.zeros(len(i), dtype=bool)
    for r in p:
        used[r["index_start"]:r["index_end"]] = True
    assert bool(used.all()) == exact
    data = gltf.glb_bytes(scene)
    shared = _parse_bytes(data, tmp_path, share=True)
    _same(shared, og.parse_gltf(data, share_geometry=True))
    if exact:
        _same(shared, (scene.vertices, scene.indices, scene.primitives))
    for a, b in zip(shared.primitives, scene.primitives):
        ta = shared.vertices[shared.indices[a["index_start"]:a["index_end"]]]
        tb = scene.vertices[scene.indices[b["index_start"]:b["index_end"]]]
        assert ta.tobytes() == tb.tobytes()
        assert a.tobytes()[:-8] == b.tobytes()[:-8]           # all but the index range
    # without sharing, the reference appends one copy of the geometry per primitive
    full = _parse_bytes(data, tmp_path)
    _same(full, og.parse_gltf(data))
    assert len(full.indices) == int(np.sum(p["index_end"] - p["index_start"]))
    for a, b in zip(full.primitives, scene.primitives):
        ta = full.vertices[full.indices[a["index_start"]:a["index_end"]]]
        tb = scene.vertices[scene.indices[b["index_start"]:b["index_end"]]]
        assert ta.tobytes() == tb.tobytes()


def test_generated_scene_builds_identically_after_round_trip(tmp_path):
    v, i, p = scenegen.spheres_in_box(1, n_spheres=10)
    ref = nb.Scene(v, i.copy(), p)
    back = _parse_bytes(gltf.glb_bytes(nb.Scene(v, i, p)), tmp_path, share=True)
    a = nb.build_scene(ref)
    b = nb.build_scene(back)
    for x, y in zip(a, b):
        assert x.tobytes() == y.tobytes()
    assert np.array_equal(ref.indices, back.indices)


# ---- accessor forms the reader must handle -------------------------------------------------------

def _one_triangle_doc():
    v = scenegen.make_vertices(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], dtype=np.float32),
                               np.tile(np.array([[0, 0, 1]], dtype=np.float32), (4, 1)))
    v["u"] = [0.0, 1.0, 0.0, 1.0]
    v["v"] = [0.0, 0.0, 1.0, 1.0]
    p = np.zeros(1, dtype=layout.PRIMITIVE)
    p["local_to_world"] = np.eye(4, dtype=np.float32).reshape(16)
    p["color"] = [0.5, 0.25, 0.125, 1.0]
    p["roughness"], p["metallic"] = 0.5, 0.0
    p["index_start"], p["index_end"] = 0, 6
    return _split(gltf.glb_bytes(nb.Scene(v, np.array([0, 1, 2, 2, 1, 3], dtype=np.uint32), p)))


def interleaved_sparse_doc():
    """The one-triangle document with POSITION and NORMAL interleaved in one strided view,
    TEXCOORD_0 as normalized u16 and a sparse override of tangent 2."""
    doc, binary = _one_triangle_doc()
    # interleave POSITION and NORMAL into one strided view
    acc = doc["accessors"]
    pos_view = doc["bufferViews"][acc[0]["bufferView"]]
    nrm_view = doc["bufferViews"][acc[1]["bufferView"]]
    pos = np.frombuffer(bytes(binary), np.float32, 12, pos_view["byteOffset"]).reshape(4, 3)
    nrm = np.frombuffer(bytes(binary), np.float32, 12, nrm_view["byteOffset"]).reshape(4, 3)
    inter = np.concatenate([pos, nrm], 1).astype(np.float32).tobytes()
    off = len(binary)
    binary += inter
    doc["bufferViews"].append({"buffer": 0, "byteOffset": off, "byteLength": len(inter), "byteStride": 24})
    acc[0]["bufferView"] = acc[1]["bufferView"] = len(doc["bufferViews"]) - 1
    acc[1]["byteOffset"] = 12
    # TEXCOORD_0 as normalized u16
    uv = (np.array([[0, 0], [1, 0], [0, 1], [1, 1]]) * np.array([65535, 32768])).astype(np.uint16).tobytes()
    off = len(binary)
    binary += uv
    doc["bufferViews"].append({"buffer": 0, "byteOffset": off, "byteLength": len(uv)})
    acc[2].update({"bufferView": len(doc["bufferViews"]) - 1, "componentType": 5123, "normalized": True})
    # a sparse override of tangent 2
    idx = np.array([2], dtype=np.uint8).tobytes() + b"\0\0\0"
    val = np.array([0, 1, 0, -1], dtype=np.float32).tobytes()
    off = len(binary)
    binary += idx + val
    doc["bufferViews"] += [{"buffer": 0, "byteOffset": off, "byteLength": 1},
                           {"buffer": 0, "byteOffset": off + 4, "byteLength": 16}]
    nv = len(doc["bufferViews"])
    acc[3]["sparse"] = {"count": 1, "indices": {"bufferView": nv - 2, "componentType": 5121},
                        "values": {"bufferView": nv - 1}}
    doc["buffers"][0]["byteLength"] = len(binary)
    return doc, binary


def test_interleaved_sparse_and_normalized_accessors(tmp_path):
    data = _join(*interleaved_sparse_doc())
    s = _parse_bytes(data, tmp_path)
    _same(s, og.parse_gltf(data))
    assert np.array_equal(s.vertices["tangent"][2], [0, 1, 0, -1])
    assert s.vertices["v"][2] == np.float32(32768) / np.float32(65535)
    assert np.array_equal(s.vertices["normal"], np.tile([0, 0, 1], (4, 1)))


def test_gltf_json_with_data_uri_and_external_buffer(tmp_path):
    doc, binary = _one_triangle_doc()
    glb = _join(doc, binary)
    want = _parse_bytes(glb, tmp_path)
    doc["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(bytes(binary)).decode()
    (tmp_path / "a.gltf").write_text(json.dumps(doc))
    _same(nb.Scene().parse_gltf(tmp_path / "a.gltf"), (want.vertices, want.indices, want.primitives))
    doc["buffers"][0]["uri"] = "tri%20data.bin"
    (tmp_path / "tri data.bin").write_bytes(bytes(binary))
    (tmp_path / "b.gltf").write_text(json.dumps(doc))
    _same(nb.Scene().parse_gltf(tmp_path / "b.gltf"), (want.vertices, want.indices, want.primitives))


def test_memory_parse_reads_files_only_beside_a_given_directory(tmp_path, monkeypatch):
    """nb_host_scene_parse_gltf_memory with no base directory refuses a file URI (even one that
    names a file in the working directory); with the directory it reads the file beside it."""
    import ctypes
    from nbounce_amd._lib import lib
    doc, binary = _one_triangle_doc()
    (tmp_path / "tri.bin").write_bytes(bytes(binary))
    doc["buffers"][0]["uri"] = "tri.bin"
    text = json.dumps(doc).encode()
    monkeypatch.chdir(tmp_path)
    h = ctypes.c_void_p()
    assert lib.nb_host_scene_create(ctypes.byref(h)) == 0
    try:
        assert lib.nb_host_scene_parse_gltf_memory(h, text, len(text), None, 0) == 7   # NB_ERR_FORMAT
        assert b"needs the document's directory" in lib.nb_last_error()
        assert lib.nb_host_scene_parse_gltf_memory(h, text, len(text), str(tmp_path).encode(), 0) == 0
        nv, ni, npr = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        assert lib.nb_host_scene_counts(h, ctypes.byref(nv), ctypes.byref(ni), ctypes.byref(npr)) == 0
        assert (nv.value, ni.value, npr.value) == (4, 6, 1)
    finally:
        lib.nb_host_scene_destroy(h)


def test_sparse_accessor_without_a_view_is_capped(tmp_path):
    """A sparse accessor with no bufferView costs memory that no buffer backs: one of more than
    256 MiB is refused before anything is allocated; one within the cap reads as zeros plus its
    overrides (here the whole TANGENT accessor)."""
    doc, binary = interleaved_sparse_doc()
    acc = doc["accessors"][3]
    del acc["bufferView"]
    s = _parse_bytes(_join(doc, binary), tmp_path)
    want = np.zeros((4, 4), np.float32)
    want[2] = [0, 1, 0, -1]
    assert np.array_equal(s.vertices["tangent"], want)
    acc["count"] = (1 << 24) + 1          # VEC4 f32: 16 B each, just over 256 MiB
    with pytest.raises(nb.NbError) as e:
        _parse_bytes(_join(doc, binary), tmp_path)
    assert e.value.status == 7 and "256 MiB" in str(e.value)


def test_default_and_emissive_materials(tmp_path):
    doc, binary = _one_triangle_doc()
    del doc["meshes"][0]["primitives"][0]["material"]
    s = _parse_bytes(_join(doc, binary), tmp_path)
    p = s.primitives[0]
    assert np.array_equal(p["color"], [1, 1, 1, 1]) and p["roughness"] == 1 and p["metallic"] == 1 and p["emissive"] == 0
    doc["meshes"][0]["primitives"][0]["material"] = 0
    doc["materials"][0]["emissiveFactor"] = [4.0, 0.0, 2.5]
    s = _parse_bytes(_join(doc, binary), tmp_path)
    assert np.array_equal(s.primitives[0]["color"], [4.0, 0.0, 2.5, 1.0]) and s.primitives[0]["emissive"] == 1.0
    _same(s, og.parse_gltf(_join(doc, binary)))


def test_node_trs_and_hierarchy(tmp_path):
    doc, binary = _one_triangle_doc()
    node = doc["nodes"][0]
    del node["matrix"]
    node.update({"translation": [1.5, -2, 0.25], "rotation": [0.2, -0.4, 0.1, math.sqrt(1 - 0.21)],
                 "scale": [2, 0.5, 1.25]})
    # a parent whose transform the reference ignores (node-local matrices only)
    doc["nodes"].append({"children": [0], "translation": [100, 0, 0]})
    doc["scenes"][0]["nodes"] = [1]
    data = _join(doc, binary)
    s = _parse_bytes(data, tmp_path)
    _same(s, og.parse_gltf(data))
    assert len(s.primitives) == 1
    assert s.primitives[0]["local_to_world"][12] == np.float32(1.5)


@pytest.mark.parametrize("edit, message", [
    (lambda d: d["meshes"][0]["primitives"][0].update(mode=1), "Not a triangle list"),
    (lambda d: d["meshes"][0]["primitives"][0]["attributes"].pop("POSITION"), "Missing positions"),
    (lambda d: d["meshes"][0]["primitives"][0]["attributes"].pop("NORMAL"), "Missing normals"),
    (lambda d: d["meshes"][0]["primitives"][0]["attributes"].pop("TANGENT"), "Tangent generation is not yet supported"),
    (lambda d: d["meshes"][0]["primitives"][0].pop("indices"), "Missing indices"),
])
def test_reference_mesh_errors(edit, message, tmp_path):
    doc, binary = _one_triangle_doc()
    edit(doc)
    data = _join(doc, binary)
    with pytest.raises(og.MeshError, match=message):
        og.parse_gltf(data)
    scene = nb.Scene().parse_gltf(os.path.join(GOLDEN, "spheres.glb"))
    before = (scene.vertices.copy(), scene.indices.copy(), scene.primitives.copy())
    with pytest.raises(nb.NbError, match=message):
        scene.parse_gltf(_write(tmp_path, data))
    _same(scene, before)


def _write(tmp_path, data, name="bad.glb"):
    p = tmp_path / name
    p.write_bytes(data)
    return p


@pytest.mark.parametrize("corrupt", [
    lambda d: d[:30],                                        # truncated JSON chunk
    lambda d: d[:12] + struct.pack("<II", 10 ** 6, 0x4E4F534A) + d[20:],   # JSON length past the end
    lambda d: b"glTF" + struct.pack("<II", 1, len(d)) + d[12:],           # version 1
    lambda d: d.replace(b'"byteLength"', b'"byteLengthX"', 1),            # a view without length
])
def test_malformed_files_are_rejected(corrupt, tmp_path):
    doc, binary = _one_triangle_doc()
    with pytest.raises(nb.NbError, match="Gltf error"):
        nb.Scene().parse_gltf(_write(tmp_path, corrupt(_join(doc, binary))))


def test_out_of_range_accessor_is_rejected(tmp_path):
    doc, binary = _one_triangle_doc()
    doc["accessors"][0]["count"] = 1000
    with pytest.raises(nb.NbError, match="exceeds"):
        nb.Scene().parse_gltf(_write(tmp_path, _join(doc, binary)))


def test_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(nb.NbError) as e:
        nb.Scene().parse_gltf(tmp_path / "nope.glb")
    assert e.value.status == 8
