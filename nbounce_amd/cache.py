"""Learned radiance cache over libnbounce (nb_cache_* / nb_render_cached in include/nbounce.h)."""
import ctypes

import numpy as np

from ._lib import check, lib, ptr


class RadianceCache:
    """MLP radiance cache bound to a context; scene_min/scene_max normalise positions."""

    def __init__(self, ctx, scene_min, scene_max, seed=0, max_batch=1 << 16, fp16=False):
        lo = np.ascontiguousarray(scene_min, dtype=np.float32)
        hi = np.ascontiguousarray(scene_max, dtype=np.float32)
        h = ctypes.c_void_p()
        check(lib.nb_cache_create_ex(ctx.handle, ptr(lo), ptr(hi), seed, max_batch, 1 if fp16 else 0, ctypes.byref(h)))
        self.handle, self.ctx, self.fp16 = h, ctx, fp16

    def loss_scale(self):
        s = ctypes.c_float()
        check(lib.nb_cache_loss_scale(self.handle, ctypes.byref(s)))
        return s.value

    @staticmethod
    def param_count():
        return lib.nb_cache_param_count()

    def params(self):
        p = np.zeros(self.param_count(), dtype=np.float32)
        check(lib.nb_cache_get_params(self.handle, ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib.nb_cache_set_params(self.handle, ptr(p)))

    def infer(self, raw):
        raw = np.ascontiguousarray(raw, dtype=np.float32)
        out = np.zeros((len(raw), 3), dtype=np.float32)
        check(lib.nb_cache_infer(self.handle, ptr(raw), len(raw), ptr(out)))
        return out

    def render(self, pathtracer, train_paths=0, lr=1e-3, weight=None):
        """One cache-assisted dispatch of `pathtracer` (same sample counting as dispatch()),
        training on `train_paths` full paths first.  `weight` overrides the accumulation weight
        1/sample (1.0 makes a fresh one-sample image).  Returns the last training loss or None."""
        pt = pathtracer
        if pt.globals.sample >= pt.max_sample_count:
            return None
        pt.globals.sample += 1
        pt.globals.weight = float(np.float32(1.0) / np.float32(pt.globals.sample)) if weight is None else weight
        g = pt.globals.array()
        loss = ctypes.c_float(-1.0)
        check(lib.nb_render_cached(self.ctx.handle, pt.scene.handle, pt.output.handle, ptr(pt._camera_data), ptr(g),
                                   pt.envmap.handle, self.handle, train_paths, lr, ctypes.byref(loss)))
        return loss.value if train_paths else None

    def close(self):
        if getattr(self, "handle", None):
            lib.nb_cache_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rel_mse(img, ref, eps=1e-2):
    """Relative mean squared error: mean over pixels and channels of (img - ref)^2 / (ref^2 + eps)."""
    img = np.asarray(img, dtype=np.float64)[..., :3]
    ref = np.asarray(ref, dtype=np.float64)[..., :3]
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))
