"""Pins of the oracle's multi-bounce integrator that do not restate it
(pathtracing.wgsl:151-159 Sobol fetch, :164-239 bounce loop, :227-236 Russian roulette).

Each test drives one path of orc_sample_pixel through a closed, inward-facing box and reads the
number of intersect_scene calls (rays) it made, with Sobol tables built so that the outcome
follows from the reference's rules by hand:
  * a mirror box (metallic 1, roughness 0, white) keeps the path throughput at exactly 1 and
    its direction independent of the table, so the path ends exactly at the first bounce b
    whose roulette number reaches 1 - (b / B)^8: scanning that number pins the length term;
  * a marker table whose only large roulette number sits in entry i*17 + 1 + 2b ends the path
    at bounce b: it pins the per-bounce entry of the first dimension set, row overlap included;
  * in a diffuse box, perturbing one table entry's second coordinate changes the path only if
    the integrator reads that entry as a cosine-lobe direction: the set of entries that matter
    is exactly {i*17} and i*17 + 2 + 2b for the bounces the path made;
  * over a real Sobol table, the mean ray count per sample in the white mirror box equals
    sum_b prod_{k<b} (1 - (k/B)^8).
Each test also checks that a plausible slip (exponent 4 or 5, set index 2 + 2b, a 0-based row)
would have failed it.
"""
import math

import numpy as np
import pytest

import oracle
from nbounce_amd import scenegen

W = H = 8          # target size; the tests follow single pixels
STRIDE = 17        # pathtracing.wgsl:4


def closed_box(L=10.0, color=(1.0, 1.0, 1.0), roughness=0.0, metallic=1.0, emitter=None):
    """A closed cube [0, L]^3 of 12 triangles wound inward (visible from inside only).  With
    `emitter` = radiance, an emissive square patch (a third of the wall) sits 0.01 in front of
    the -z wall."""
    faces = {
        "-x": ([[0, 0, 0], [0, 0, L], [0, L, L], [0, L, 0]], [1, 0, 0]),
        "+x": ([[L, 0, 0], [L, L, 0], [L, L, L], [L, 0, L]], [-1, 0, 0]),
        "-y": ([[0, 0, 0], [L, 0, 0], [L, 0, L], [0, 0, L]], [0, 1, 0]),
        "+y": ([[0, L, 0], [0, L, L], [L, L, L], [L, L, 0]], [0, -1, 0]),
        "-z": ([[0, 0, 0], [0, L, 0], [L, L, 0], [L, 0, 0]], [0, 0, 1]),
        "+z": ([[0, 0, L], [L, 0, L], [L, L, L], [0, L, L]], [0, 0, -1]),
    }
    quads = list(faces.values())
    if emitter is not None:
        a, b, z = 0.2 * L, 0.8 * L, 0.01
        quads.append(([[a, a, z], [a, b, z], [b, b, z], [b, a, z]], [0, 0, 1]))
    bp, bn, bf = [], [], []
    for corners, normal in quads:
        k = sum(len(x) for x in bp)
        bp.append(np.array(corners, dtype=np.float64))
        bn.append(np.tile(np.array(normal, dtype=np.float64), (4, 1)))
        bf.append(np.array([[0, 1, 2], [0, 2, 3]]) + k)
    bp, bn, bf = np.concatenate(bp), np.concatenate(bn), np.concatenate(bf)
    bf = scenegen.orient_outward(bp, bf, bn)
    v = scenegen.make_vertices(bp, bn)
    idx = bf.reshape(-1).astype(np.uint32)
    prims = [scenegen._prim(np.eye(4), (*color, 1.0), roughness, metallic, 0.0, 0, 36)]
    if emitter is not None:
        prims.append(scenegen._prim(np.eye(4), (emitter, emitter, emitter, 1.0), 1.0, 0.0, 1.0, 36, 42))
    prims = np.concatenate(prims)
    blas, tlas, inst, pidx = oracle.build_scene(v, idx, prims)
    return oracle.Scene(blas, tlas, inst, v, pidx)


ENV = oracle.Env(scenegen.constant_cube(4, 1.0))
# From the box centre towards a point off the diagonal: every wall is met at an incidence whose
# cosine lies in 0.52..0.64, never at a corner or an edge.
CAM = oracle.camera((5.0, 5.0, 5.0), (5.83, 6.0, 7.21), (0.0, 1.0, 0.0), math.pi / 3.0, 1.0, 0.1)


def fract32(a):
    a = np.float32(a)
    return np.float32(a - np.floor(a))


def entry_for(u, shift):
    """A table value s with fract(s + shift) = u (up to float32 rounding) and that fract."""
    s = fract32(np.float32(u) - np.float32(shift))
    return s, fract32(s + np.float32(shift))


def base_table(rows, shift, x=0.0, y=0.0, z=0.0):
    """Every entry shifted by the pixel's rotation to the values (x, y, z) (w = 0)."""
    t = np.zeros((rows, STRIDE, 4), dtype=np.float32)
    for c, u in enumerate((x, y, z)):
        t[..., c] = entry_for(u, shift[c])[0]
    return t


def rays_of(scene, table, row, bounces, cf, x=4, y=4):
    g = oracle.globals_(row, 1.0, bounces, cf)
    out, cnt = oracle.sample_pixel(scene, CAM, g, table, STRIDE, ENV, W, H, x, y)
    return out, int(cnt["rays"])


def first_death(u, bounces, exponent=8):
    """Bounce at which a path whose roulette numbers all equal u ends (p = 1 - (b/B)^e, u < p
    continues; bounce B has p = 0)."""
    for b in range(bounces):
        if not u < 1.0 - (b / bounces) ** exponent:
            return b
    return bounces


@pytest.fixture(scope="module")
def mirror_white():
    return closed_box(color=(1.0, 1.0, 1.0))


@pytest.mark.parametrize("bounces", [4, 8, 12, 16])
def test_length_term_of_the_roulette(mirror_white, bounces):
    """pathtracing.wgsl:227: p_continue = min(1 - (bounce / bounces)^8, 1) x min(lum(T) cf, 1).
    In the white mirror box T stays >= 1, so with cf = 4 only the length term remains and a path
    whose roulette numbers are all u makes first_death(u) + 1 rays."""
    row = 1
    shift = oracle.pixel_shift(4, 4)
    observed = []
    for c in np.linspace(0.0, 0.995, 200):
        t = base_table(4, shift, x=0.5, y=0.5, z=c)
        u = float(entry_for(c, shift[2])[1])
        _, rays = rays_of(mirror_white, t, row, bounces, 4.0)
        observed.append((u, rays))
    # geometry check: with u = 0 the path runs the whole loop, every ray hits a wall
    assert observed[0][1] == bounces + 1
    for u, rays in observed:
        if min(abs(u - (1.0 - (b / bounces) ** 8)) for b in range(bounces)) < 1e-5:
            continue   # too close to a threshold for a float64 closed form
        assert rays == first_death(u, bounces) + 1, (u, rays)
    # the scan tells the exponent apart from nearby ones (and from a linear length term)
    for e in (4, 5, 7, 9, 1):
        assert any(rays != first_death(u, bounces, e) + 1 for u, rays in observed), e


@pytest.mark.parametrize("bounces", [8, 12])
def test_roulette_reads_set_1_plus_2b_of_the_row(bounces):
    """pathtracing.wgsl:151-153,227-236: bounce b's roulette number is entry sample*17 + 1 + 2b,
    component z, of the table; for b >= 8 that is the next row's entries (fixed stride 17).
    A grey mirror box (throughput x Schlick ~0.21 per bounce, cf = 1) keeps p in 0.07..0.22: a
    marker entry of 0.98 ends the path there, every other entry (0.02) continues."""
    scene = closed_box(color=(0.2, 0.2, 0.2))
    shift = oracle.pixel_shift(4, 4)
    row = 2
    for b0 in range(bounces):
        t = base_table(5, shift, x=0.5, y=0.5, z=0.02)
        e = row * STRIDE + 1 + 2 * b0
        t.reshape(-1, 4)[e, 2] = entry_for(0.98, shift[2])[0]
        _, rays = rays_of(scene, t, row, bounces, 1.0)
        assert rays == b0 + 1, (b0, rays)
        # slips that read another entry never meet the marker: the path would run to the end
        for wrong in (row * STRIDE + 2 + 2 * b0, (row - 1) * STRIDE + 1 + 2 * b0, row * (2 * bounces + 1) + 1 + 2 * b0):
            if wrong != e:
                t2 = base_table(5, shift, x=0.5, y=0.5, z=0.02)
                t2.reshape(-1, 4)[wrong, 2] = entry_for(0.98, shift[2])[0]
                _, rays2 = rays_of(scene, t2, row, bounces, 1.0)
                assert rays2 != b0 + 1 or wrong == e, (b0, wrong, rays2)


@pytest.mark.parametrize("bounces", [8, 12])
def test_cosine_lobe_reads_set_2_plus_2b(bounces):
    """pathtracing.wgsl:151-153,214-222: the diffuse lobe's direction comes from entry
    sample*17 + 2 + 2b (components y, z), the jitter from entry sample*17 (x, y).  In a diffuse
    box with a small emitter, a path whose lobe numbers (x = 0.999) always pick the diffuse lobe
    and whose roulette numbers (z = 0.02, cf large) always continue ends on the emitter after
    some bounces; perturbing component y of an entry changes its radiance or length exactly when
    the entry is the jitter or one of those bounces' direction entries.  Paths of every length
    up to the bounce limit are probed, so for 12 bounces the entries of the next row are too."""
    scene = closed_box(color=(0.7, 0.6, 0.5), roughness=1.0, metallic=0.0, emitter=3.0)
    rows, row = 4, 1
    by_length = {}
    for seed in range(8):
        for px in range(W):
            for py in range(H):
                shift = oracle.pixel_shift(px, py)
                t = base_table(rows, shift, x=0.999, y=0.37, z=0.02)
                rng = np.random.default_rng(1000 * seed + 8 * px + py)
                flat = t.reshape(-1, 4)
                for e in range(rows * STRIDE):
                    flat[e, 1] = entry_for(rng.uniform(0.05, 0.95), shift[1])[0]
                    off = e - row * STRIDE      # direction entries sit at even offsets 2 + 2b
                    if off > 0 and off % 2 == 0:
                        flat[e, 2] = entry_for(rng.uniform(0.05, 0.9), shift[2])[0]
                out, rays = rays_of(scene, t, row, bounces, 1e6, px, py)
                if out.max() > 0 and rays >= 2:      # scattered, then ended on the emitter
                    by_length.setdefault(rays, (px, py, t, out))
    assert set(range(2, bounces + 2)) <= set(by_length)   # paths of every length up to the limit
    for rays, (px, py, t, out) in sorted(by_length.items()):
        n_bounce = rays - 1    # the last ray hit the emitter
        expect = {row * STRIDE} | {row * STRIDE + 2 + 2 * b for b in range(n_bounce)}
        changed = set()
        for e in range(rows * STRIDE):
            t2 = t.copy()
            flat = t2.reshape(-1, 4)
            flat[e, 1] = fract32(flat[e, 1] + np.float32(0.25))
            o2, r2 = rays_of(scene, t2, row, bounces, 1e6, px, py)
            if r2 != rays or not np.array_equal(o2.view(np.uint32), out.view(np.uint32)):
                changed.add(e)
        assert changed == expect, (px, py, sorted(changed), sorted(expect))


def expected_rays(bounces, exponent=8):
    """sum over b = 0..B of prod_{k<b} p_k with p_k = 1 - (k/B)^e (p_B = 0): the mean number of
    intersect_scene calls of a path whose roulette numbers are independent and uniform."""
    total, reach = 0.0, 1.0
    for b in range(bounces + 1):
        total += reach
        reach *= (1.0 - (b / bounces) ** exponent) if b < bounces else 0.0
    return total


@pytest.mark.parametrize("bounces", [8, 12])
def test_mean_path_length_over_the_sobol_table(bounces):
    """Over the real Sobol table (rows 1..256, 32x32 pixels) the mean rays per sample in a white
    mirror box equals expected_rays(B) within 0.5 %; an exponent of 4 or 5 is off by > 5 %.  The
    box is 1000 wide so that the 0.1 minimum hit distance lets almost no path leak out next to
    an edge (leaked paths see the sky: under 1 % of them)."""
    from tests import _workloads as Wk
    scene = closed_box(L=1000.0, color=(1.0, 1.0, 1.0))
    cam = oracle.camera((500.0, 500.0, 500.0), (583.0, 600.0, 721.0), (0.0, 1.0, 0.0), math.pi / 3.0, 1.0, 0.1)
    table = Wk.table()
    rays = samples = leaked = 0
    for row in range(1, 257):
        g = oracle.globals_(row, 1.0, bounces, 4.0)
        img, cnt = oracle.render(scene, cam, g, table, STRIDE, ENV, 32, 32, threads=8)
        rays += int(cnt["rays"])
        samples += int(cnt["samples"])
        leaked += int((img[..., 0] > 0).sum())
    mean = rays / samples
    assert leaked / samples < 0.01
    assert abs(mean / expected_rays(bounces) - 1.0) < 0.005, (mean, expected_rays(bounces))
    for e in (4, 5):
        assert abs(mean / expected_rays(bounces, e) - 1.0) > 0.05


def test_training_path_is_the_sample_path():
    """orc_train_path follows orc_sample_pixel's path (same radiance bits) and records its first
    three non-emissive shading points; the first arrives with throughput 1, so its target is the
    radiance itself, and the first record's position lies on the primary ray's closest hit."""
    from nbounce_amd import scenegen as sg
    from tests import _workloads as Wk
    wl = Wk.small()
    osc = Wk.oracle_scene(wl)
    env = oracle.Env(sg.sky_cube(16))
    cam = Wk.camera(sg.SMALL["eye"], sg.SMALL["target"], 1.0)
    table = Wk.table()
    n_paths = 0
    for k in range(64):
        x, y = (37 * k) % 64, (11 * k) % 64
        g = oracle.globals_(1 + k % 7, 1.0, 8, 4.0)
        raw, tgt, rad = oracle.train_path(osc, cam, g, table, STRIDE, env, 64, 64, x, y)
        out, _ = oracle.sample_pixel(osc, cam, g, table, STRIDE, env, 64, 64, x, y)
        assert np.array_equal(rad.view(np.uint32), out.view(np.uint32))
        assert len(raw) <= 3 and np.isfinite(tgt).all()
        if len(raw):
            n_paths += 1
            assert np.array_equal(tgt[0].view(np.uint32), rad.view(np.uint32))
            assert np.all(raw[:, 14:] == 0) and np.allclose(np.linalg.norm(raw[:, 3:6], axis=1), 1, atol=1e-5)
    assert n_paths > 32
