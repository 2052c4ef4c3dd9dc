"""nb_primary_hits_noalloc (the primary-hit debug view without device allocation) through the C
ABI: the same output as nb_primary_hits on two calls in a row, the second smaller, and sizes
beyond its scratch refused."""
import ctypes

import numpy as np
import pytest

from tests import _workloads as W

pytestmark = pytest.mark.gpu


def test_noalloc_primary_hits_equal_primary_hits():
    import nbounce_amd as nb
    from nbounce_amd import scenegen
    from nbounce_amd._lib import NbError, check, lib, ptr
    print("case: nb_primary_hits_noalloc vs nb_primary_hits, 256x256 then 64x48 (small scene)", flush=True)
    ctx = nb.Context(0)
    w = W.small()
    sb = nb.SceneBuffers(ctx, w["blas"], w["tlas"], w["instances"], w["vertices"], w["indices"])
    t = W.table()
    check(lib.nb_sobol_set(ctx.handle, ptr(np.ascontiguousarray(t)), t.shape[0], t.shape[1], 17))
    P = scenegen.SMALL
    for (width, height, sample) in ((256, 256, 3), (64, 48, 7)):
        cam = nb.CameraController(world_position=P["eye"], target=P["target"], aspect_ratio=width / height).data()
        want = np.zeros((height, width, 5), dtype=np.uint32)
        check(lib.nb_primary_hits(ctx.handle, sb.handle, ptr(cam), width, height, sample, ptr(want)))
        got = np.full((height, width, 5), 0xDEADBEEF, dtype=np.uint32)
        check(lib.nb_primary_hits_noalloc(ctx.handle, sb.handle, ptr(cam), width, height, sample, ptr(got)))
        print(f"{width}x{height}: hits {(want[..., 0] != 0xFFFFFFFF).mean():.3f}", flush=True)
        np.testing.assert_array_equal(got, want)
    # beyond the 2^21-pixel scratch: refused before any GPU work
    big = np.zeros(8, dtype=np.uint32)
    with pytest.raises(NbError) as ex:
        check(lib.nb_primary_hits_noalloc(ctx.handle, sb.handle, ptr(cam), 2048, 1032, 1, ptr(big)))
    assert ex.value.status == 1
    ctx.close()
