"""nbounce_amd: MI355X (gfx950) path-tracing kernel behind the reference's `pathtracing` surface.

The compute path is libnbounce.so (HIP kernels + C ABI, include/nbounce.h); this package is its
Python host mirror.  Importing it requires the built library.
"""
from .pathtracer import (DDS_BC6H_UF16, DDS_RGBA16F, DDS_RGBA32F, BlitRenderer, CameraController, Context, EnvMap,
                         Globals, NbError, Pathtracer, Scene, SceneBuffers, build_scene, read_dds_cube, sobol_table,
                         write_dds_cube, write_exr, write_pfm)

__all__ = ["BlitRenderer", "CameraController", "Context", "DDS_BC6H_UF16", "DDS_RGBA16F", "DDS_RGBA32F", "EnvMap",
           "Globals", "NbError", "Pathtracer", "Scene", "SceneBuffers", "build_scene", "read_dds_cube", "sobol_table",
           "write_dds_cube", "write_exr", "write_pfm"]
