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


RANDOM_CASES = 24


def random_case(case):
    """One seeded random parity case (deterministic in `case`): a scene (spheres in a box of random
    size, or a random number of instanced meshes over a random extent), a camera inside or
    outside it (some looking exactly along the x or z axis), the image size, bounce limit,
    contribution factor, first roulette bounce and sky form."""
    rng = np.random.default_rng(7000 + case)
    if case % 2 == 0:
        box = float(rng.uniform(2.0, 30.0))
        n = int(rng.integers(1, 41))
        v, i, pr = scenegen.spheres_in_box(seed=case, n_spheres=n, n_emissive=int(rng.integers(0, n + 1)),
                                           subdivisions=int(rng.integers(1, 4)), box=box)
        lo, hi = np.zeros(3), np.full(3, box)
        what = f"{n} spheres in a {box:.1f} box"
    else:
        extent = float(rng.uniform(5.0, 300.0))
        n = int(rng.integers(1, 49))
        v, i, pr = scenegen.instanced_meshes(seed=case, n_instances=n, extent=extent,
                                             emissive_every=int(rng.integers(1, 9)))
        lo, hi = np.zeros(3), np.full(3, extent)
        what = f"{n} instanced meshes over {extent:.0f}"
    span = hi - lo
    eye = lo + span * rng.uniform(-0.4, 1.4, 3) if rng.random() < 0.5 else lo + span * rng.uniform(0.1, 0.9, 3)
    if rng.random() < 0.3:
        # exactly along the x or z axis (the up vector is +y)
        axis = np.zeros(3)
        axis[int(rng.choice([0, 2]))] = float(rng.choice([-1.0, 1.0]))
        target = eye + axis * float(span.max())
        what += ", axis-aligned view"
    else:
        target = lo + span * rng.uniform(0.2, 0.8, 3)
        if np.linalg.norm(target - eye) < 1e-3 * span.max():
            target = target + span * 0.25
    size = [(8, 8), (48, 32), (32, 64), (64, 48)][int(rng.integers(0, 4))]
    bounces = int(rng.integers(0, 14))
    cf = float(rng.uniform(0.25, 8.0))
    rr_first = int(rng.integers(1, 6)) if case % 3 == 2 else 0
    latlong = rng.random() < 0.4
    return dict(vit=(v, i, pr), eye=tuple(float(x) for x in eye), target=tuple(float(x) for x in target),
                size=size, bounces=bounces, cf=cf, rr_first=rr_first, latlong=latlong, what=what)
