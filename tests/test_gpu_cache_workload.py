"""The learned radiance cache at its full workload: the spheres-in-box scene at 1920x1080, 8
bounces, 240 frames of one sample while the camera orbits, each frame training the cache online
(Adam, learning rate 1e-3) on 2^16 full paths of the plain integrator (every record of their first
three shading points); then, at the orbit's final pose, the relative MSE of the plain and the
cache-assisted images against the committed float64 4096-spp CPU-oracle reference
(tests/golden/cache_reference_spheres_1080p.npz, every 8th pixel).

The cache-assisted frame must reach the plain tracer's relative MSE in at most half the plain
tracer's frame time, with the training of that frame excluded and included."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_cache_reaches_the_plain_error_in_half_the_time():
    import bench
    print("case: cache workload 1920x1080, 240-frame orbit, 2^16 training paths per frame, lr 1e-3", flush=True)
    r = bench.cache_equal_error(train_paths=1 << 16, lr=1e-3)
    plain = r["plain"]
    print(f"plain rel-MSE by spp: {[(p['spp'], round(p['rel_mse'], 4)) for p in plain]}", flush=True)
    print(f"cache 1 spp: rel-MSE {r['cache_1spp']['rel_mse']:.4f} in {r['cache_1spp']['ms']:.3f} ms; plain spp at "
          f"equal error {r['plain_spp_at_equal_error']} x {r['plain_ms_per_spp']} ms; ratio {r['ratio']}, "
          f"with training {r['with_training_ratio']} ({r['cache_frame_with_training_ms']} ms per trained frame)",
          flush=True)
    # the plain estimator converges towards the reference (error ~ 1/spp)
    errs = [p["rel_mse"] for p in plain]
    assert all(b < a for a, b in zip(errs, errs[1:]))
    assert errs[-1] < errs[0] / 30.0
    # training lowered the loss
    assert r["loss_first_last"][1] < r["loss_first_last"][0]
    assert r["ratio"] is not None and r["ratio"] <= 0.5
    assert r["with_training_ratio"] is not None and r["with_training_ratio"] <= 0.5
    # the cached image is no noisier than two plain samples' worth
    assert r["plain_spp_at_equal_error"] > 1.5
    assert np.isfinite(r["cache_1spp"]["rel_mse"])
