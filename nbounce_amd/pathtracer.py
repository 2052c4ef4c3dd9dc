"""Host-side mirror of the reference's `pathtracing` renderer surface over libnbounce.

Names, arguments and behaviour follow src/pathtracing/{scene,pathtracer,envmap}.rs and
src/common/camera.rs; the window / wgpu context is replaced by `Context` (one HIP device) and an
explicit output size.
"""
import ctypes
import os
import math

import numpy as np

from . import layout
from ._lib import NbError, check, lib, ptr

__all__ = ["Context", "Scene", "SceneBuffers", "CameraController", "EnvMap", "Globals", "Pathtracer",
           "BlitRenderer", "build_scene", "sobol_table", "write_exr", "write_pfm", "NbError"]


class Context:
    """One HIP device and stream (replaces WGPUContext, src/common/wgpu_context.rs:13-79)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib.nb_context_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device

    def synchronize(self):
        check(lib.nb_context_synchronize(self.handle))

    def counters(self):
        c = np.zeros(1, dtype=layout.COUNTERS)
        check(lib.nb_counters_get(self.handle, ptr(c)))
        return {k: int(c[k][0]) for k in layout.COUNTERS.names}

    def enable_counters(self, on=True):
        check(lib.nb_counters_enable(self.handle, 1 if on else 0))

    def reset_counters(self):
        check(lib.nb_counters_reset(self.handle))

    def set_render_options(self, roulette_first_bounce=0):
        """Native options of nb_render / nb_render_frames on this context (no counterpart in the
        reference): Russian roulette from bounce `roulette_first_bounce` on (0: the reference's
        rule, pathtracing.wgsl:227-236)."""
        opts = np.array([roulette_first_bounce], dtype=np.uint32)
        check(lib.nb_render_options_set(self.handle, ptr(opts)))

    def enable_render_timing(self, on=True):
        """Per-render HIP events for last_render_ms (off by default: they cost a frame stream
        about 0.5 %)."""
        check(lib.nb_render_timing(self.handle, 1 if on else 0))

    def last_render_ms(self):
        ms = ctypes.c_float()
        check(lib.nb_last_render_ms(self.handle, ctypes.byref(ms)))
        return ms.value

    def timer_start(self):
        check(lib.nb_stream_timer_start(self.handle))

    def timer_stop(self):
        ms = ctypes.c_float()
        check(lib.nb_stream_timer_stop(self.handle, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "handle", None):
            lib.nb_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """In-memory scene (src/pathtracing/scene.rs:100-105): vertices, indices, primitives."""

    def __init__(self, vertices=None, indices=None, primitives=None):
        def arr(a, dt):
            return np.zeros(0, dtype=dt) if a is None else np.ascontiguousarray(a, dtype=dt).reshape(-1)
        self.vertices = arr(vertices, layout.VERTEX)
        self.indices = arr(indices, np.uint32)
        self.primitives = arr(primitives, layout.PRIMITIVE)

    def parse_gltf(self, path, share_geometry=False):
        """Scene::parse_gltf (scene.rs:108-217): appends a glTF 2.0 file's mesh primitives and
        one primitive per mesh primitive of each mesh node.  Raises NbError (NB_ERR_FORMAT with
        the reference's MeshError text) and leaves the scene unchanged on a rejected file.

        share_geometry: mesh primitives that name the same accessors share one copy of the
        geometry and one index range (use with build_scene(share_blas=True))."""
        h = ctypes.c_void_p()
        check(lib.nb_host_scene_create(ctypes.byref(h)))
        try:
            check(lib.nb_host_scene_parse_gltf(h, str(path).encode(), 1 if share_geometry else 0))
            nv, ni, nprim = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
            check(lib.nb_host_scene_counts(h, ctypes.byref(nv), ctypes.byref(ni), ctypes.byref(nprim)))
            v = np.zeros(nv.value, dtype=layout.VERTEX)
            i = np.zeros(ni.value, dtype=np.uint32)
            pr = np.zeros(nprim.value, dtype=layout.PRIMITIVE)
            check(lib.nb_host_scene_copy(h, ptr(v), ptr(i), ptr(pr)))
        finally:
            lib.nb_host_scene_destroy(h)
        # scene.rs:135-136: indices and index ranges continue after what the scene already holds
        i += np.uint32(len(self.vertices))
        pr["index_start"] += np.uint32(len(self.indices))
        pr["index_end"] += np.uint32(len(self.indices))
        self.vertices = np.concatenate([self.vertices, v])
        self.indices = np.concatenate([self.indices, i])
        self.primitives = np.concatenate([self.primitives, pr])
        return self


def build_scene(scene, share_blas=False):
    """SceneBuffers::from_scene (scene.rs:280-307) on the host, without the upload.

    Returns (blas, tlas, instances) and permutes scene.indices in place, as the reference does.
    """
    flags = 1 if share_blas else 0
    bcap, tcap = ctypes.c_uint32(), ctypes.c_uint32()
    check(lib.nb_build_scene_capacity(len(scene.indices), ptr(scene.primitives), len(scene.primitives), flags,
                                      ctypes.byref(bcap), ctypes.byref(tcap)))
    blas = np.zeros(max(bcap.value, 1), dtype=layout.NODE)
    tlas = np.zeros(max(tcap.value, 1), dtype=layout.NODE)
    inst = np.zeros(len(scene.primitives), dtype=layout.INSTANCE)
    nb, nt = ctypes.c_uint32(), ctypes.c_uint32()
    check(lib.nb_build_scene(ptr(scene.vertices), len(scene.vertices), ptr(scene.indices), len(scene.indices),
                             ptr(scene.primitives), len(scene.primitives), flags, ptr(blas), bcap.value,
                             ctypes.byref(nb), ptr(tlas), tcap.value, ctypes.byref(nt), ptr(inst)))
    return blas[:nb.value].copy(), tlas[:nt.value].copy(), inst


class SceneBuffers:
    """Device-resident scene (SceneBuffers, scene.rs:271-277): the five group-1 arrays."""

    def __init__(self, ctx, blas, tlas, instances, vertices, indices):
        self.ctx = ctx
        blas = np.ascontiguousarray(blas, dtype=layout.NODE)
        tlas = np.ascontiguousarray(tlas, dtype=layout.NODE)
        instances = np.ascontiguousarray(instances, dtype=layout.INSTANCE)
        vertices = np.ascontiguousarray(vertices, dtype=layout.VERTEX)
        indices = np.ascontiguousarray(indices, dtype=np.uint32)
        h = ctypes.c_void_p()
        check(lib.nb_scene_create(ctx.handle, ptr(blas), len(blas), ptr(tlas), len(tlas), ptr(instances),
                                  len(instances), ptr(vertices), len(vertices), ptr(indices), len(indices),
                                  ctypes.byref(h)))
        self.handle = h
        self.n_instances = len(instances)
        self.n_triangles = len(indices) // 3

    @classmethod
    def from_scene(cls, ctx, scene, share_blas=False):
        """SceneBuffers::from_scene (scene.rs:280): build on the host, then upload."""
        blas, tlas, inst = build_scene(scene, share_blas=share_blas)
        sb = cls(ctx, blas, tlas, inst, scene.vertices, scene.indices)
        sb.blas, sb.tlas, sb.instances = blas, tlas, inst
        return sb

    def device_bytes(self):
        b = ctypes.c_uint64()
        check(lib.nb_scene_device_bytes(self.handle, ctypes.byref(b)))
        return b.value

    def close(self):
        if getattr(self, "handle", None):
            lib.nb_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sobol_table(rows, dims, seed=0):
    """pathtracer.rs:62-64: rows x dims float4 entries of the Owen-scrambled Sobol sequence."""
    out = np.zeros((rows, dims, 4), dtype=np.float32)
    check(lib.nb_sobol_generate(rows, dims, seed, ptr(out)))
    return out


class CameraController:
    """Camera state and CameraBuffer (src/common/camera.rs:14-52,121-142); no input handling."""

    def __init__(self, world_position=(5.0, 0.0, 0.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
                 fov=math.pi / 3.0, aspect_ratio=1.0, near=0.1):
        self.world_position = np.array(world_position, dtype=np.float32)
        self.target = np.array(target, dtype=np.float32)
        self.up = np.array(up, dtype=np.float32)
        self.fov, self.aspect_ratio, self.near = float(fov), float(aspect_ratio), float(near)

    def resize(self, aspect_ratio):
        self.aspect_ratio = float(aspect_ratio)

    def data(self):
        cam = np.zeros(1, dtype=layout.CAMERA)
        check(lib.nb_camera_look_at(ptr(self.world_position), ptr(self.target), ptr(self.up), self.fov,
                                    self.aspect_ratio, self.near, ptr(cam)))
        return cam


DDS_RGBA32F, DDS_RGBA16F, DDS_BC6H_UF16 = 0, 1, 2


def read_dds_cube(path):
    """Mip 0 of the six faces of a cube DDS file as (6, n, n, 4) RGBA32F (nb_dds_read_cube)."""
    n = ctypes.c_uint32()
    p = os.fsencode(path)
    check(lib.nb_dds_read_cube(p, None, 0, ctypes.byref(n)))
    faces = np.zeros((6, n.value, n.value, 4), dtype=np.float32)
    check(lib.nb_dds_read_cube(p, ptr(faces), faces.size, ctypes.byref(n)))
    return faces


def write_dds_cube(path, faces, fmt=DDS_RGBA32F):
    """Writes (6, n, n, 4) RGBA32F faces as a cube DDS (RGBA32F, RGBA16F or BC6H_UF16)."""
    faces = np.ascontiguousarray(faces, dtype=np.float32)
    check(lib.nb_dds_write_cube(os.fsencode(path), ptr(faces), faces.shape[1], fmt))


class EnvMap:
    """Cube environment map (EnvMap, src/pathtracing/envmap.rs:5-26) from RGBA32F faces, or from
    a cube DDS file with EnvMap.load (envmap.rs:12)."""

    def __init__(self, ctx, faces):
        faces = np.ascontiguousarray(faces, dtype=np.float32)
        if faces.ndim != 4 or faces.shape[0] != 6 or faces.shape[1] != faces.shape[2] or faces.shape[3] != 4:
            raise ValueError("faces must be (6, n, n, 4) RGBA32F in +X,-X,+Y,-Y,+Z,-Z order")
        h = ctypes.c_void_p()
        check(lib.nb_envmap_create_cube(ctx.handle, ptr(faces), faces.shape[1], ctypes.byref(h)))
        self.handle = h
        self.faces = faces

    @classmethod
    def latlong(cls, ctx, texels):
        """An equirectangular (height, width, 4) RGBA32F map, row 0 at the top (+y); no counterpart
        in the reference, which samples only the cube.  `faces` is None and `texels` holds it."""
        texels = np.ascontiguousarray(texels, dtype=np.float32)
        if texels.ndim != 3 or texels.shape[2] != 4:
            raise ValueError("texels must be (height, width, 4) RGBA32F")
        self = cls.__new__(cls)
        h = ctypes.c_void_p()
        check(lib.nb_envmap_create_latlong(ctx.handle, ptr(texels), texels.shape[1], texels.shape[0],
                                           ctypes.byref(h)))
        self.handle = h
        self.faces = None
        self.texels = texels
        return self

    @classmethod
    def load(cls, ctx, path):
        """EnvMap::load (envmap.rs:12): a BC6H (or RGBA16F / RGBA32F) cube DDS, decoded on the host
        and uploaded (nb_envmap_load); `faces` holds the decoded texels."""
        self = cls.__new__(cls)
        h = ctypes.c_void_p()
        check(lib.nb_envmap_load(ctx.handle, os.fsencode(path), ctypes.byref(h)))
        self.handle = h
        self.faces = read_dds_cube(path)
        return self

    def close(self):
        if getattr(self, "handle", None):
            lib.nb_envmap_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Globals:
    """Push constants (pathtracer.rs:26-44): sample, weight, bounces, contribution_factor."""

    def __init__(self, sample=0, weight=0.0, bounces=8, contribution_factor=4.0):
        self.sample, self.weight = sample, weight
        self.bounces, self.contribution_factor = bounces, contribution_factor

    def array(self):
        g = np.zeros(1, dtype=layout.GLOBALS)
        g["sample"], g["weight"] = self.sample, np.float32(self.weight)
        g["bounces"], g["contribution_factor"] = self.bounces, np.float32(self.contribution_factor)
        return g


class _Target:
    def __init__(self, ctx, width, height):
        h = ctypes.c_void_p()
        check(lib.nb_target_create(ctx.handle, width, height, ctypes.byref(h)))
        self.handle, self.width, self.height = h, width, height

    def read(self):
        out = np.zeros((self.height, self.width, 4), dtype=np.float32)
        check(lib.nb_target_read(self.handle, ptr(out)))
        return out

    def write(self, rgba):
        rgba = np.ascontiguousarray(rgba, dtype=np.float32)
        assert rgba.shape == (self.height, self.width, 4)
        check(lib.nb_target_write(self.handle, ptr(rgba)))

    def close(self):
        if getattr(self, "handle", None):
            lib.nb_target_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pathtracer:
    """Progressive path tracer (Pathtracer, src/pathtracing/pathtracer.rs:15-239).

    `dispatch()` adds one sample per pixel exactly like the reference's compute dispatch.
    """

    COMPUTE_SIZE = 8
    SOBOL_SETS_PER_BOUNCE = 2

    def __init__(self, ctx, scene, camera, envmap, width, height, resolution_factor=0.3, sobol_stride=None,
                 table=None):
        self.ctx = ctx
        self.scene = scene
        self.camera = camera
        self.envmap = envmap
        self.window = (int(width), int(height))
        self.resolution_factor = float(resolution_factor)
        self.globals = Globals()
        self.max_sample_count = 1024
        dims = self.globals.bounces * self.SOBOL_SETS_PER_BOUNCE + 1
        # The table's row stride is fixed when the table is made (pathtracing.wgsl:4 hard-codes 17).
        self.sobol_stride = dims if sobol_stride is None else sobol_stride
        self.table = sobol_table(self.max_sample_count, dims, 0) if table is None else table
        self._upload_table()
        self.output = self._create_output_texture()
        self._camera_data = camera.data()

    def _upload_table(self):
        t = np.ascontiguousarray(self.table, dtype=np.float32)
        rows, dims = t.shape[0], t.shape[1]
        check(lib.nb_sobol_set(self.ctx.handle, ptr(t), rows, dims, self.sobol_stride))

    def _create_output_texture(self):  # pathtracer.rs:192-202
        w = int(np.float32(self.window[0]) * np.float32(self.resolution_factor))
        h = int(np.float32(self.window[1]) * np.float32(self.resolution_factor))
        w, h = w // self.COMPUTE_SIZE * self.COMPUTE_SIZE, h // self.COMPUTE_SIZE * self.COMPUTE_SIZE
        if w == 0 or h == 0:
            raise ValueError(f"output size {w}x{h} is empty")
        return _Target(self.ctx, w, h)

    @property
    def size(self):
        return self.output.width, self.output.height

    def output_texture(self):
        """The RGBA32F accumulation image as a (height, width, 4) array (waits for the GPU)."""
        return self.output.read()

    def resize(self, width, height):
        self.window = (int(width), int(height))
        self.output.close()
        self.output = self._create_output_texture()

    def update(self, camera=None, envmap=None):
        if camera is not None:
            self.camera = camera
        if envmap is not None:
            self.envmap = envmap
        self._camera_data = self.camera.data()
        self.invalidate()

    def sample_count(self):
        return self.globals.sample

    def invalidate(self):
        self.globals.sample = 0

    def dispatch(self, weight=None):
        """pathtracer.rs:225-239: +1 sample per pixel; no-op once max_sample_count is reached.
        `weight` overrides the accumulation weight 1/sample (1.0 makes a fresh one-sample image)."""
        if self.globals.sample >= self.max_sample_count:
            return
        self.globals.sample += 1
        self.globals.weight = float(np.float32(1.0) / np.float32(self.globals.sample)) if weight is None else weight
        g = self.globals.array()
        check(lib.nb_render(self.ctx.handle, self.scene.handle, self.output.handle, ptr(self._camera_data), ptr(g),
                            self.envmap.handle))

    def dispatch_frames(self, n):
        """n dispatches in one launch (same image as n dispatch() calls)."""
        n = min(n, self.max_sample_count - self.globals.sample)
        if n <= 0:
            return 0
        g = Globals(self.globals.sample + 1, 0.0, self.globals.bounces, self.globals.contribution_factor).array()
        check(lib.nb_render_frames(self.ctx.handle, self.scene.handle, self.output.handle, ptr(self._camera_data),
                                   ptr(g), n, self.envmap.handle))
        self.globals.sample += n
        self.globals.weight = float(np.float32(1.0) / np.float32(self.globals.sample))
        return n

    def primary_hits(self, sample=1):
        w, h = self.size
        out = np.zeros((h, w, 5), dtype=np.uint32)
        check(lib.nb_primary_hits(self.ctx.handle, self.scene.handle, ptr(self._camera_data), w, h, sample, ptr(out)))
        return out


class BlitRenderer:
    """Draws the accumulation target onto an RGBA16F surface (BlitRenderer,
    src/pathtracing/blit_renderer.rs; blit.wgsl:27-38; surface format wgpu_context.rs:62):
    bilinear, clamp-to-edge, surface row 0 at the top.  `render(pathtracer)` enqueues the blit of
    its output texture; `surface()` waits and returns the (height, width, 4) float16 image."""

    def __init__(self, ctx, width, height):
        h = ctypes.c_void_p()
        check(lib.nb_display_create(ctx.handle, int(width), int(height), ctypes.byref(h)))
        self.handle, self.ctx, self.width, self.height = h, ctx, int(width), int(height)

    def render(self, pathtracer):
        check(lib.nb_display_blit(self.handle, pathtracer.output.handle))

    def surface(self):
        out = np.zeros((self.height, self.width, 4), dtype=np.float16)
        check(lib.nb_display_read(self.handle, ptr(out)))
        return out

    def close(self):
        if getattr(self, "handle", None):
            lib.nb_display_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _image_arg(rgba):
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    if rgba.ndim != 3 or rgba.shape[2] != 4:
        raise ValueError(f"expected a (height, width, 4) RGBA image, got shape {rgba.shape}")
    return rgba


def write_pfm(path, rgba):
    """Portable Float Map (RGB) of an image in target row order (output_texture())."""
    rgba = _image_arg(rgba)
    check(lib.nb_image_write_pfm(str(path).encode(), ptr(rgba), rgba.shape[1], rgba.shape[0]))


def write_exr(path, rgba, half=False):
    """Uncompressed scanline OpenEXR of an image in target row order (output_texture())."""
    rgba = _image_arg(rgba)
    check(lib.nb_image_write_exr(str(path).encode(), ptr(rgba), rgba.shape[1], rgba.shape[0], 1 if half else 0))
