"""Oracle (test infrastructure only): a Python restatement of Scene::parse_gltf
(src/pathtracing/scene.rs:108-217) over the `gltf` 1.4.1 crate's reader, used to check the
library's native glTF reader.  Nothing in the product imports this module.

Followed, line by line:
  * scene.rs:131-183  meshes in document order, primitives in order; mode must be TRIANGLES
    (:139-141); positions (:146), normals (:147) required; TEXCOORD_0 as f32 or repeat (0, 0)
    (:149-152); tangents required (:161, :176); vertices zipped to the shortest attribute; indices
    required (:171), offset by the first vertex of the primitive (:172);
  * scene.rs:185-212  every node with a mesh, document order, local matrix only (:187); colour =
    (emissive, 1) if emissive_factor != 0 else base colour (:190-197); roughness, metallic,
    emissive 0/1 (:199-205).  glTF defaults: base colour 1, metallic 1, roughness 1, emissive 0.
  * gltf crate Transform::matrix: T * R * S, quaternion (x, y, z, w) to matrix with
    x2 = x + x ..., xx2 = x2 * x ..., sx2 = x2 * w ..., columns
    (1 - yy2 - zz2, xy2 + sz2, xz2 - sy2), (xy2 - sz2, 1 - xx2 - zz2, yz2 + sx2),
    (xz2 + sy2, yz2 - sx2, 1 - xx2 - yy2); a product's column j is
    a0 * bj.x + a1 * bj.y + a2 * bj.z + a3 * bj.w summed left to right, all in float32.

Pins: the two reference assets' published shape (spheres.glb: 15 icosphere nodes with
translation-only transforms, 19,200 triangles; testscene.glb: 4 TRS nodes, 32,460 triangles);
translation-only matrices are exact; TRS matrices against a float64 rotation-matrix closed form
and exact 90-degree rotations; default materials.  Parity unpinned: the crate's TRS expression
order at the ulp level for general rotations (restated from its published source, not run).
"""
import json
import struct

import numpy as np

VERTEX = np.dtype([("position", "<f4", 3), ("u", "<f4"), ("normal", "<f4", 3), ("v", "<f4"),
                   ("tangent", "<f4", 4)])
PRIMITIVE = np.dtype([("local_to_world", "<f4", 16), ("color", "<f4", 4), ("roughness", "<f4"),
                      ("metallic", "<f4"), ("emissive", "<f4"), ("index_start", "<u4"), ("index_end", "<u4")])

_DTYPES = {5120: "<i1", 5121: "<u1", 5122: "<i2", 5123: "<u2", 5125: "<u4", 5126: "<f4"}
_COMPS = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}


class MeshError(Exception):
    pass


def _load(data):
    if data[:4] == b"glTF":
        _, version, length = struct.unpack_from("<4sII", data, 0)
        jlen, jtype = struct.unpack_from("<II", data, 12)
        assert version == 2 and jtype == 0x4E4F534A
        doc = json.loads(data[20:20 + jlen])
        off = 20 + jlen
        bin_chunk = None
        if off + 8 <= length:
            blen, btype = struct.unpack_from("<II", data, off)
            if btype == 0x004E4942:
                bin_chunk = data[off + 8:off + 8 + blen]
        buffers = [bin_chunk if "uri" not in b else None for b in doc.get("buffers", [])]
        return doc, buffers
    raise ValueError("the oracle reads .glb containers only")


def _accessor(doc, buffers, index):
    a = doc["accessors"][index]
    if "bufferView" not in a and "sparse" not in a:
        return None
    dt = np.dtype(_DTYPES[a["componentType"]])
    comps = _COMPS[a["type"]]
    count = a["count"]
    out = np.zeros((count, comps), dtype=dt)
    if "bufferView" in a:
        v = doc["bufferViews"][a["bufferView"]]
        buf = buffers[v["buffer"]]
        start = v.get("byteOffset", 0) + a.get("byteOffset", 0)
        elem = dt.itemsize * comps
        stride = v.get("byteStride", elem)
        for i in range(count) if stride != elem else []:
            out[i] = np.frombuffer(buf, dtype=dt, count=comps, offset=start + i * stride)
        if stride == elem:
            out[:] = np.frombuffer(buf, dtype=dt, count=count * comps, offset=start).reshape(count, comps)
    if "sparse" in a:
        sp = a["sparse"]
        iv, vv = sp["indices"], sp["values"]
        idt = np.dtype(_DTYPES[iv["componentType"]])
        ivw, vvw = doc["bufferViews"][iv["bufferView"]], doc["bufferViews"][vv["bufferView"]]
        ids = np.frombuffer(buffers[ivw["buffer"]], dtype=idt, count=sp["count"],
                            offset=ivw.get("byteOffset", 0) + iv.get("byteOffset", 0))
        vals = np.frombuffer(buffers[vvw["buffer"]], dtype=dt, count=sp["count"] * comps,
                             offset=vvw.get("byteOffset", 0) + vv.get("byteOffset", 0)).reshape(-1, comps)
        out[ids.astype(np.int64)] = vals
    return a, out


def _f32_attr(doc, buffers, attrs, name):
    if name not in attrs:
        return None
    r = _accessor(doc, buffers, attrs[name])
    return None if r is None else r[1].astype(np.float32)


def trs_matrix(t, q, s):
    """gltf Transform::matrix for a decomposed transform: T * R * S, column-major, float32."""
    f = np.float32
    t = [f(c) for c in t]
    x, y, z, w = [f(c) for c in q]
    s = [f(c) for c in s]
    x2, y2, z2 = x + x, y + y, z + z
    xx2, xy2, xz2 = x2 * x, x2 * y, x2 * z
    yy2, yz2, zz2 = y2 * y, y2 * z, z2 * z
    sy2, sz2, sx2 = y2 * w, z2 * w, x2 * w
    one, zero = f(1), f(0)
    T = [[one, zero, zero, zero], [zero, one, zero, zero], [zero, zero, one, zero], [t[0], t[1], t[2], one]]
    R = [[one - yy2 - zz2, xy2 + sz2, xz2 - sy2, zero],
         [xy2 - sz2, one - xx2 - zz2, yz2 + sx2, zero],
         [xz2 + sy2, yz2 - sx2, one - xx2 - yy2, zero],
         [zero, zero, zero, one]]
    S = [[s[0], zero, zero, zero], [zero, s[1], zero, zero], [zero, zero, s[2], zero], [zero, zero, zero, one]]

    def mul(a, b):   # a, b as lists of columns
        out = []
        for j in range(4):
            col = []
            for i in range(4):
                acc = a[0][i] * b[j][0]
                acc = acc + a[1][i] * b[j][1]
                acc = acc + a[2][i] * b[j][2]
                acc = acc + a[3][i] * b[j][3]
                col.append(acc)
            out.append(col)
        return out

    m = mul(mul(T, R), S)
    return np.array([m[j][i] for j in range(4) for i in range(4)], dtype=np.float32)


def node_matrix(node):
    if "matrix" in node:
        return np.array(node["matrix"], dtype=np.float64).astype(np.float32)
    return trs_matrix(node.get("translation", [0, 0, 0]), node.get("rotation", [0, 0, 0, 1]),
                      node.get("scale", [1, 1, 1]))


def parse_gltf(data, share_geometry=False):
    """(vertices, indices, primitives) of one .glb document, as the reference's Scene arrays."""
    doc, buffers = _load(data)
    verts, idx = [], []
    n_verts = n_idx = 0
    ranges, shared = [], {}
    for mesh in doc.get("meshes", []):
        mr = []
        for p in mesh["primitives"]:
            if p.get("mode", 4) != 4:
                raise MeshError("Not a triangle list")
            attrs = p["attributes"]
            key = tuple(attrs.get(n, -1) for n in ("POSITION", "NORMAL", "TEXCOORD_0", "TANGENT")) + (p.get("indices", -1),)
            if share_geometry and key in shared:
                mr.append(shared[key])
                continue
            pos = _f32_attr(doc, buffers, attrs, "POSITION")
            if pos is None:
                raise MeshError("Missing positions")
            nrm = _f32_attr(doc, buffers, attrs, "NORMAL")
            if nrm is None:
                raise MeshError("Missing normals")
            tc = None
            if "TEXCOORD_0" in attrs:
                r = _accessor(doc, buffers, attrs["TEXCOORD_0"])
                if r is not None:
                    a, raw = r
                    if a["componentType"] == 5126:
                        tc = raw.astype(np.float32)
                    elif a["componentType"] == 5121:
                        tc = raw.astype(np.float32) / np.float32(255)
                    else:
                        tc = raw.astype(np.float32) / np.float32(65535)
            tan = _f32_attr(doc, buffers, attrs, "TANGENT")
            if tan is None:
                raise MeshError("Tangent generation is not yet supported")
            n = min(len(pos), len(nrm), len(tan), len(tc) if tc is not None else len(pos))
            v = np.zeros(n, dtype=VERTEX)
            v["position"], v["normal"], v["tangent"] = pos[:n], nrm[:n], tan[:n]
            if tc is not None:
                v["u"], v["v"] = tc[:n, 0], tc[:n, 1]
            verts.append(v)
            start_vertex = n_verts
            n_verts += n
            r = _accessor(doc, buffers, p["indices"]) if "indices" in p else None
            if r is None:
                raise MeshError("Missing indices")
            ii = (r[1].reshape(-1).astype(np.uint64) + start_vertex).astype(np.uint32)
            idx.append(ii)
            rng = (n_idx, n_idx + len(ii))
            n_idx += len(ii)
            mr.append(rng)
            shared[key] = rng
        ranges.append(mr)

    prims = []
    for node in doc.get("nodes", []):
        if "mesh" not in node:
            continue
        m = node_matrix(node)
        for k, p in enumerate(doc["meshes"][node["mesh"]]["primitives"]):
            rec = np.zeros((), dtype=PRIMITIVE)
            mat = doc["materials"][p["material"]] if "material" in p else {}
            pbr = mat.get("pbrMetallicRoughness", {})
            emissive = np.array(mat.get("emissiveFactor", [0, 0, 0]), dtype=np.float64).astype(np.float32)
            is_emissive = bool(np.any(emissive != 0))
            rec["local_to_world"] = m
            if is_emissive:
                rec["color"] = np.append(emissive, np.float32(1))
            else:
                rec["color"] = np.array(pbr.get("baseColorFactor", [1, 1, 1, 1]), dtype=np.float64).astype(np.float32)
            rec["roughness"] = np.float32(np.float64(pbr.get("roughnessFactor", 1.0)))
            rec["metallic"] = np.float32(np.float64(pbr.get("metallicFactor", 1.0)))
            rec["emissive"] = 1.0 if is_emissive else 0.0
            rec["index_start"], rec["index_end"] = ranges[node["mesh"]][k]
            prims.append(rec)
    vertices = np.concatenate(verts) if verts else np.zeros(0, dtype=VERTEX)
    indices = np.concatenate(idx) if idx else np.zeros(0, dtype=np.uint32)
    return vertices, indices, np.array(prims, dtype=PRIMITIVE).reshape(-1)
