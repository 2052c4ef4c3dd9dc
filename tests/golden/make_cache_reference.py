"""Writes cache_reference_spheres_1080p.npz: the CPU oracle's float64 mean of 4096 samples per
pixel (Sobol rows 1..4096 of a 4097-row, 17-set table, seed 0) at every 8th pixel of the
spheres-in-box scene (seed 0) seen from the learned-cache workload's evaluation pose, 1920x1080,
8 bounces, contribution factor 4, 128-texel sky cube.  Used as the ground truth for relative-MSE
measurements of the plain and cache-assisted renderers.

    python tests/golden/make_cache_reference.py [threads]
"""
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from nbounce_amd import scenegen  # noqa: E402


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else os.cpu_count()
    P = scenegen.CACHE
    v, i, pr = scenegen.spheres_in_box(0)
    blas, tlas, inst, idx = oracle.build_scene(v, i, pr)
    sc = oracle.Scene(blas, tlas, inst, v, idx)
    env = oracle.Env(scenegen.sky_cube(P["env_face"]))
    eye, target = scenegen.orbit_pose(P["frames"], P["frames"], P["centre"], P["radius"], P["height_y"])
    cam = oracle.camera(eye, target, (0, 1, 0), math.pi / 3, P["width"] / P["height"], 0.1)
    table = oracle.sobol_table(4097, 17, 0)
    xy = scenegen.eval_pixels(P["width"], P["height"], P["eval_stride"])
    g = oracle.globals_(1, 1.0, P["bounces"], P["contribution_factor"])
    t = time.time()
    ref = oracle.render_pixels_mean_f64(sc, cam, g, table, 17, env, P["width"], P["height"], xy, 1, 4096, threads)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cache_reference_spheres_1080p.npz")
    np.savez_compressed(out, xy=xy, mean=ref.astype(np.float32), eye=np.array(eye), target=np.array(target),
                        samples=4096)
    print(f"{len(xy)} pixels x 4096 samples in {time.time() - t:.0f} s -> {out}")


if __name__ == "__main__":
    main()
