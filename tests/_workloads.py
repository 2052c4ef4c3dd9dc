"""Shared test workloads: the same seeded scene built for the oracle and for libnbounce."""
import functools
import math

import numpy as np

import oracle
from nbounce_amd import scenegen


@functools.lru_cache(maxsize=None)
def small(seed=0):
    """Spheres-in-box scene built by the oracle's builder (reference arrays)."""
    v, i, pr = scenegen.spheres_in_box(seed)
    blas, tlas, inst, idx = oracle.build_scene(v, i, pr)
    return dict(vertices=v, indices_in=i, primitives=pr, blas=blas, tlas=tlas, instances=inst, indices=idx)


@functools.lru_cache(maxsize=None)
def large(seed=0, n_instances=4096):
    v, i, pr = scenegen.instanced_meshes(seed, n_instances=n_instances)
    return dict(vertices=v, indices_in=i, primitives=pr)


def oracle_scene(w):
    return oracle.Scene(w["blas"], w["tlas"], w["instances"], w["vertices"], w["indices"])


def camera(eye, target, aspect):
    return oracle.camera(eye, target, (0.0, 1.0, 0.0), math.pi / 3.0, aspect, 0.1)


@functools.lru_cache(maxsize=None)
def table(rows=1024, dims=17, seed=0):
    return oracle.sobol_table(rows, dims, seed)


def rel_err(a, b):
    """Per-pixel relative error |a - b| / max(|b|, 1e-6) over RGB."""
    a = np.asarray(a, dtype=np.float64)[..., :3]
    b = np.asarray(b, dtype=np.float64)[..., :3]
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-6)
