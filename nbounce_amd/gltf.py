"""glTF 2.0 binary (.glb) writer for in-memory scenes, the inverse of Scene.parse_gltf.

The reference only reads glTF (src/pathtracing/scene.rs:108-217); this writer is how generated
scenes reach it.  Each primitive becomes one node (its local_to_world as a column-major
`matrix`), one mesh and one material; primitives over the same index range share one set of
accessors (POSITION, NORMAL, TEXCOORD_0, TANGENT, indices), so the file holds each geometry
once.  Emissive primitives write their colour as `emissiveFactor` (the reader sets alpha to 1).
Reading the file back with share_geometry=True reproduces the arrays exactly when the index
ranges tile the index array in order and each range's vertices are a contiguous block in the
same order; without it the reader appends one copy of the geometry per primitive, as the
reference does.
"""
import json
import struct

import numpy as np

from . import layout

_F32, _U16, _U32 = 5126, 5123, 5125
_ARRAY_BUFFER, _ELEMENT_ARRAY_BUFFER = 34962, 34963


def _num(x):
    x = float(np.float32(x))
    if not np.isfinite(x):
        raise ValueError("glTF JSON cannot hold non-finite numbers")
    return x


def glb_bytes(scene, generator="nbounce_amd"):
    """The scene (vertices, indices, primitives as in Scene) as .glb bytes."""
    verts = np.asarray(scene.vertices, dtype=layout.VERTEX)
    idx = np.asarray(scene.indices, dtype=np.uint32)
    prims = np.asarray(scene.primitives, dtype=layout.PRIMITIVE)
    bin_chunks, views, accessors = [], [], []
    offset = 0

    def add_view(data, target):
        nonlocal offset
        raw = data.tobytes()
        views.append({"buffer": 0, "byteOffset": offset, "byteLength": len(raw), "target": target})
        pad = (-len(raw)) % 4
        bin_chunks.append(raw + b"\0" * pad)
        offset += len(raw) + pad
        return len(views) - 1

    def add_accessor(data, comp, typ, target, minmax=False):
        acc = {"bufferView": add_view(data, target), "componentType": comp, "count": int(len(data)), "type": typ}
        if minmax:
            acc["min"] = [_num(v) for v in data.min(axis=0)]
            acc["max"] = [_num(v) for v in data.max(axis=0)]
        accessors.append(acc)
        return len(accessors) - 1

    geometry = {}
    for p in sorted(prims, key=lambda q: int(q["index_start"])):
        key = (int(p["index_start"]), int(p["index_end"]))
        if key in geometry:
            continue
        s, e = key
        if e <= s or e > len(idx) or (e - s) % 3:
            raise ValueError(f"primitive index range [{s}, {e}) is not a triangle list inside the index array")
        tri = idx[s:e]
        lo, hi = int(tri.min()), int(tri.max())
        v = verts[lo:hi + 1]
        local = (tri - np.uint32(lo)).astype(np.uint16 if hi - lo < 65536 else np.uint32)
        geometry[key] = {
            "attributes": {
                "POSITION": add_accessor(np.ascontiguousarray(v["position"]), _F32, "VEC3", _ARRAY_BUFFER, True),
                "NORMAL": add_accessor(np.ascontiguousarray(v["normal"]), _F32, "VEC3", _ARRAY_BUFFER),
                "TEXCOORD_0": add_accessor(np.ascontiguousarray(np.stack([v["u"], v["v"]], 1)), _F32, "VEC2",
                                           _ARRAY_BUFFER),
                "TANGENT": add_accessor(np.ascontiguousarray(v["tangent"]), _F32, "VEC4", _ARRAY_BUFFER),
            },
            "indices": add_accessor(local, _U16 if local.dtype == np.uint16 else _U32, "SCALAR",
                                    _ELEMENT_ARRAY_BUFFER),
        }

    # Meshes are ordered by index range so that a reader, which appends geometry in mesh order,
    # meets the ranges in index order; nodes keep the primitive order.
    order = sorted(range(len(prims)), key=lambda k: (int(prims[k]["index_start"]), k))
    mesh_of = {k: m for m, k in enumerate(order)}
    nodes, meshes, materials = [], [None] * len(prims), [None] * len(prims)
    for k, p in enumerate(prims):
        nodes.append({"mesh": mesh_of[k], "matrix": [_num(c) for c in p["local_to_world"]]})
    for k in order:
        p = prims[k]
        m_id = mesh_of[k]
        g = geometry[(int(p["index_start"]), int(p["index_end"]))]
        mat = {"pbrMetallicRoughness": {"metallicFactor": _num(p["metallic"]),
                                        "roughnessFactor": _num(p["roughness"])}}
        if p["emissive"] != 0.0:
            mat["emissiveFactor"] = [_num(c) for c in p["color"][:3]]
            mat["pbrMetallicRoughness"]["baseColorFactor"] = [1.0, 1.0, 1.0, 1.0]
        else:
            mat["pbrMetallicRoughness"]["baseColorFactor"] = [_num(c) for c in p["color"]]
        materials[m_id] = mat
        meshes[m_id] = {"primitives": [{"attributes": dict(g["attributes"]), "indices": g["indices"],
                                        "material": m_id, "mode": 4}]}

    doc = {"asset": {"version": "2.0", "generator": generator}, "scene": 0,
           "scenes": [{"nodes": list(range(len(nodes)))}], "nodes": nodes, "meshes": meshes,
           "materials": materials, "accessors": accessors, "bufferViews": views,
           "buffers": [{"byteLength": offset}]}
    js = json.dumps(doc, separators=(",", ":")).encode()
    js += b" " * ((-len(js)) % 4)
    binary = b"".join(bin_chunks)
    total = 12 + 8 + len(js) + 8 + len(binary)
    return (struct.pack("<4sII", b"glTF", 2, total) + struct.pack("<II", len(js), 0x4E4F534A) + js
            + struct.pack("<II", len(binary), 0x004E4942) + binary)


def write_glb(path, scene, generator="nbounce_amd"):
    """Writes glb_bytes(scene) to `path`."""
    data = glb_bytes(scene, generator)
    with open(path, "wb") as f:
        f.write(data)
    return len(data)
