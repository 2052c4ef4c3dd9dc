"""ctypes binding of libnbounce.so (include/nbounce.h).

The library is the only compute path: importing this module without the built library raises,
there is no Python or CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import layout

# NB_LIBRARY points at another build of the same library (for comparing builds); default: in-tree
LIB_PATH = os.environ.get("NB_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "libnbounce.so")

_c = ctypes
_p = _c.c_void_p
_u32 = _c.c_uint32
_f32 = _c.c_float
_ptr = _c.POINTER

# name -> argtypes (every entry point returns nb_status, an int, except where noted)
SIGNATURES = {
    "nb_context_create": [_c.c_int, _c.POINTER(_p)],
    "nb_context_destroy": [_p],
    "nb_context_synchronize": [_p],
    "nb_build_scene_capacity": [_u32, _p, _u32, _u32, _ptr(_u32), _ptr(_u32)],
    "nb_build_scene": [_p, _u32, _p, _u32, _p, _u32, _u32, _p, _u32, _ptr(_u32), _p, _u32, _ptr(_u32), _p],
    "nb_host_scene_create": [_c.POINTER(_p)],
    "nb_host_scene_destroy": [_p],
    "nb_host_scene_parse_gltf": [_p, _c.c_char_p, _u32],
    "nb_host_scene_parse_gltf_memory": [_p, _p, _c.c_size_t, _c.c_char_p, _u32],
    "nb_host_scene_counts": [_p, _ptr(_u32), _ptr(_u32), _ptr(_u32)],
    "nb_host_scene_copy": [_p, _p, _p, _p],
    "nb_scene_create": [_p, _p, _u32, _p, _u32, _p, _u32, _p, _u32, _p, _u32, _c.POINTER(_p)],
    "nb_scene_destroy": [_p],
    "nb_scene_device_bytes": [_p, _ptr(_c.c_uint64)],
    "nb_sobol_generate": [_u32, _u32, _u32, _p],
    "nb_sobol_set": [_p, _p, _u32, _u32, _u32],
    "nb_envmap_create_cube": [_p, _p, _u32, _c.POINTER(_p)],
    "nb_envmap_create_latlong": [_p, _p, _u32, _u32, _c.POINTER(_p)],
    "nb_render_options_set": [_p, _p],
    "nb_envmap_destroy": [_p],
    "nb_envmap_load": [_p, _c.c_char_p, _c.POINTER(_p)],
    "nb_dds_read_cube": [_c.c_char_p, _p, _c.c_uint64, _ptr(_u32)],
    "nb_dds_read_cube_memory": [_p, _c.c_size_t, _p, _c.c_uint64, _ptr(_u32)],
    "nb_dds_write_cube": [_c.c_char_p, _p, _u32, _u32],
    "nb_bc6h_decode_block": [_p, _u32, _p],
    "nb_camera_look_at": [_p, _p, _p, _f32, _f32, _f32, _p],
    "nb_target_create": [_p, _u32, _u32, _c.POINTER(_p)],
    "nb_target_destroy": [_p],
    "nb_target_read": [_p, _p],
    "nb_target_write": [_p, _p],
    "nb_target_clear": [_p],
    "nb_display_create": [_p, _u32, _u32, _c.POINTER(_p)],
    "nb_display_destroy": [_p],
    "nb_display_blit": [_p, _p],
    "nb_display_read": [_p, _p],
    "nb_image_write_pfm": [_c.c_char_p, _p, _u32, _u32],
    "nb_image_write_exr": [_c.c_char_p, _p, _u32, _u32, _u32],
    "nb_render": [_p, _p, _p, _p, _p, _p],
    "nb_render_frames": [_p, _p, _p, _p, _p, _u32, _p],
    "nb_last_render_ms": [_p, _ptr(_f32)],
    "nb_render_timing": [_p, _u32],
    "nb_stream_timer_start": [_p],
    "nb_stream_timer_stop": [_p, _ptr(_f32)],
    "nb_counters_get": [_p, _p],
    "nb_counters_reset": [_p],
    "nb_counters_enable": [_p, _u32],
    "nb_primary_hits": [_p, _p, _p, _u32, _u32, _u32, _p],
    "nb_primary_hits_noalloc": [_p, _p, _p, _u32, _u32, _u32, _p],
    "nb_server_create": [_p, _c.c_uint64, _u32, _u32, _c.POINTER(_p)],
    "nb_server_destroy": [_p],
    "nb_server_add_scene": [_p, _p, _p, _ptr(_u32)],
    "nb_server_set_cache": [_p, _u32, _p],
    "nb_server_submit": [_p, _p, _ptr(_c.c_uint64)],
    "nb_server_step": [_p, _ptr(_u32), _ptr(_u32)],
    "nb_server_poll": [_p, _p, _u32, _ptr(_u32)],
    "nb_server_read": [_p, _c.c_uint64, _p],
    "nb_server_release": [_p, _c.c_uint64],
    "nb_server_now": [_p, _ptr(_c.c_double)],
    "nb_server_stats_get": [_p, _p],
    "nb_cache_create": [_p, _p, _p, _u32, _u32, _c.POINTER(_p)],
    "nb_cache_create_ex": [_p, _p, _p, _u32, _u32, _u32, _c.POINTER(_p)],
    "nb_cache_loss_scale": [_p, _ptr(_f32)],
    "nb_cache_destroy": [_p],
    "nb_cache_get_params": [_p, _p],
    "nb_cache_set_params": [_p, _p],
    "nb_cache_encode": [_p, _p, _u32, _p],
    "nb_cache_infer": [_p, _p, _u32, _p],
    "nb_cache_gradient": [_p, _p, _p, _u32, _ptr(_f32), _p],
    "nb_cache_train": [_p, _p, _p, _u32, _u32, _f32, _ptr(_f32)],
    "nb_render_cached": [_p, _p, _p, _p, _p, _p, _p, _u32, _f32, _ptr(_f32)],
    "nb_cache_training_records": [_p, _p, _u32, _u32, _p, _p, _p, _p, _u32, _p, _p, _u32, _ptr(_u32)],
    "nb_cache_query_records": [_p, _p, _u32, _u32, _p, _p, _p, _p, _p, _p, _p, _u32, _ptr(_u32)],
}


class NbError(RuntimeError):
    """A call into libnbounce returned a nonzero nb_status."""

    def __init__(self, status, message):
        super().__init__(f"nb_status {status}: {message}")
        self.status = status


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python nbounce_amd/build.py` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _c.c_int
    lib.nb_last_error.argtypes = []
    lib.nb_last_error.restype = _c.c_char_p
    lib.nb_version.argtypes = []
    lib.nb_version.restype = _u32
    lib.nb_cache_param_count.argtypes = []
    lib.nb_cache_param_count.restype = _u32
    return lib


lib = _load()


def check(status):
    if status != 0:
        raise NbError(status, lib.nb_last_error().decode(errors="replace"))


def ptr(a):
    """Address of a C-contiguous numpy array (or None for a null pointer).  The returned
    c_void_p holds a reference to the array, so `f(ptr(make_array()))` cannot pass the address
    of an array that was freed before the call."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays crossing the ABI must be C-contiguous"
    return a.ctypes.data_as(_c.c_void_p)


def struct_bytes(dtype, **fields):
    a = np.zeros(1, dtype=dtype)
    for k, v in fields.items():
        a[k] = v
    return a
