"""numpy dtypes with the reference's byte layouts (the bytemuck casts of the Rust structs).

They are the element types of the arrays that cross the C ABI (include/nbounce.h).
"""
import numpy as np

# BVHNode, src/pathtracing/bvh.rs:15-24 (32 B)
NODE = np.dtype([("min", "<f4", 3), ("start", "<u4"), ("max", "<f4", 3), ("end", "<u4")])
# Vertex, src/pathtracing/scene.rs:13-21 (48 B)
VERTEX = np.dtype([("position", "<f4", 3), ("u", "<f4"), ("normal", "<f4", 3), ("v", "<f4"),
                   ("tangent", "<f4", 4)])
# Instance, src/pathtracing/scene.rs:220-230 (160 B), column-major matrices
INSTANCE = np.dtype([("world_to_local", "<f4", 16), ("local_to_world", "<f4", 16), ("color", "<f4", 4),
                     ("roughness", "<f4"), ("metallic", "<f4"), ("emissive", "<f4"), ("node", "<u4")])
# Primitive, src/pathtracing/scene.rs:59-67 (index_range as two u32)
PRIMITIVE = np.dtype([("local_to_world", "<f4", 16), ("color", "<f4", 4), ("roughness", "<f4"),
                      ("metallic", "<f4"), ("emissive", "<f4"), ("index_start", "<u4"), ("index_end", "<u4")])
# CameraBuffer, src/common/camera.rs:7-12 (128 B)
CAMERA = np.dtype([("world_to_clip", "<f4", 16), ("clip_to_world", "<f4", 16)])
# Globals, src/pathtracing/pathtracer.rs:26-33 (16 B)
GLOBALS = np.dtype([("sample", "<u4"), ("weight", "<f4"), ("bounces", "<u4"), ("contribution_factor", "<f4")])
COUNTERS = np.dtype([("rays", "<u8"), ("n_aabb", "<u8"), ("n_tri", "<u8"), ("n_inst", "<u8"), ("samples", "<u8"), ("hits", "<u8")])

# nb_request / nb_completion / nb_server_stats (include/nbounce.h, serving)
REQUEST = np.dtype([("camera", CAMERA), ("scene", "<u4"), ("width", "<u4"), ("height", "<u4"), ("spp", "<u4"),
                    ("bounces", "<u4"), ("contribution_factor", "<f4"), ("arrival_s", "<f8"), ("mode", "<u4"),
                    ("train_paths", "<u4"), ("lr", "<f4"), ("pad", "<u4")], align=True)
COMPLETION = np.dtype([("id", "<u8"), ("arrival_s", "<f8"), ("admit_s", "<f8"), ("done_s", "<f8")])
SERVER_STATS = np.dtype([("steps", "<u8"), ("rays", "<u8"), ("samples", "<u8"), ("n_aabb", "<u8"), ("n_tri", "<u8"),
                         ("n_inst", "<u8"), ("hits", "<u8"), ("bytes_in_use", "<u8"), ("bytes_budget", "<u8"),
                         ("active", "<u4"), ("queued", "<u4"), ("completed", "<u4"), ("pad", "<u4"),
                         ("gpu_busy_s", "<f8"), ("cached_samples", "<u8"), ("cached_busy_s", "<f8")])

assert NODE.itemsize == 32 and VERTEX.itemsize == 48 and INSTANCE.itemsize == 160
assert PRIMITIVE.itemsize == 100 and CAMERA.itemsize == 128 and GLOBALS.itemsize == 16
assert REQUEST.itemsize == 176 and COMPLETION.itemsize == 32 and SERVER_STATS.itemsize == 112


def mat4_cols(m):
    """A 4x4 row-major numpy matrix as the 16 column-major floats glam stores."""
    return np.asarray(m, dtype=np.float32).T.reshape(16)
