"""Parity on the benchmarked scene itself: the seeded terrain of 3,084,292 triangles at
1920x1080 (heightfield 512x512 + 2000 icospheres, GGX/Lambert, 2 quad lights), against the
CPU restatement of the reference (oracle/cpu_ref: Bvh build and queries, bvh.cpp:16-647;
GiTracer/DiTracer, gitracer.cpp:13-140, ditracer.cpp:57-239).

* Ray queries: 65,536 camera rays, 65,536 incoherent bounce rays from their hit points
  (uniform sphere directions) and 65,536 shadow rays towards points on the 4 light triangles,
  as closest-hit queries, plus the shadow rays as any-hit queries (262,144 queries). Primitive
  agreement >= 99.99 %, t and barycentrics within 1e-5 relative where they agree, and every
  disagreement adjudicated in FP64 as an edge, grazing or tie case.
The scene gets the default node layout, the quantised 4-wide one (64-B nodes, 8-bit child
boxes).

* GI: a band of 1920 x 192 pixels at 64 spp, depth 8, rendered with the default launch shape
  (2^23 paths per chunk: three chunks dealt over the two wavefront lanes, 3 trace workgroups per
  CU per lane); its 32 x 32 sub-crop matches GiTracer::GetLi sample for sample and per pixel
  within 3 standard errors.
"""
import numpy as np
import pytest

import oracle_lib
from adjudicate import classify

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oracle_lib.available("cpu"), reason="CPU restatement not built")]

W, H, GRID, DEPTH, SEED = 1920, 1080, 8, 8, 1


@pytest.fixture(scope="module")
def bench_scene(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import terrain_scene
    sc = terrain_scene(width=W, height=H)
    assert sc.ntris == 3_084_292
    s = hip.Scene(ctx, sc)
    info = s.info()
    assert info["width"] == 4 and info["quantized"] and info["node_bytes"] == 64
    return sc, s, oracle_lib.Oracle("cpu").scene(sc)


def _uniform_dirs(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def _check_closest(sc, rays, gp, gt, rp, rt, label):
    same = gp == rp
    n = len(rays)
    print(f"{label}: {n} closest-hit rays, primitive agreement {same.mean():.6f} ({(~same).sum()} differ), "
          f"{(rp >= 0).mean():.3f} hit")
    assert same.mean() >= 0.9999
    hit = same & (rp >= 0)
    np.testing.assert_allclose(gt[hit, 0], rt[hit, 0], rtol=1e-5)
    np.testing.assert_allclose(gt[hit, 1:], rt[hit, 1:], rtol=1e-5, atol=1e-6)
    verdicts = [classify(sc, rays[i], int(gp[i]), int(rp[i])) for i in np.nonzero(~same)[0]]
    assert "fail" not in verdicts, verdicts


def test_ray_queries_match_restatement(bench_scene):
    sc, gpu, cpu = bench_scene
    rng = np.random.Generator(np.random.PCG64(31))
    n = 1 << 16
    # camera rays through jittered image-plane points (PerscpectiveCamera::GenerateRay)
    cam = cpu.o.camera_rays(sc.camera, rng.uniform(0, 1, (n, 2)).astype(np.float32))
    gp, gt = gpu.intersect(cam)
    rp, rt = cpu.intersect(cam)
    _check_closest(sc, cam, gp, gt, rp, rt, "camera")
    # incoherent bounce rays from the camera rays' hit points (GiTracer's epsilon 0.01)
    hit = rp >= 0
    p = cam[hit, 0:3] + rt[hit, 0:1] * cam[hit, 4:7]
    m = int(hit.sum())
    idx = rng.integers(0, m, n)
    bounce = np.zeros((n, 8), np.float32)
    bounce[:, 0:3] = p[idx]
    bounce[:, 3] = 0.01
    bounce[:, 4:7] = _uniform_dirs(rng, n)
    bounce[:, 7] = 1e6
    gp, gt = gpu.intersect(bounce)
    rp, rt = cpu.intersect(bounce)
    _check_closest(sc, bounce, gp, gt, rp, rt, "bounce")
    # shadow rays to uniform points on the light triangles, tmax = dist - 0.01 (ditracer.cpp:118)
    lt = sc.lights
    b = sc.bundles
    tris = np.array([b[L[0], 0] + L[1] for L in lt])
    V = sc.verts[sc.tri_v[tris[rng.integers(0, len(tris), n)]]].astype(np.float64)
    u, v = rng.uniform(size=n), rng.uniform(size=n)
    a, bb = np.sqrt(u) * (1 - v), np.sqrt(u) * v
    q = (1 - a - bb)[:, None] * V[:, 0] + a[:, None] * V[:, 1] + bb[:, None] * V[:, 2]
    o = p[rng.integers(0, m, n)].astype(np.float64)
    dvec = q - o
    dist = np.linalg.norm(dvec, axis=1)
    shadow = np.zeros((n, 8), np.float32)
    shadow[:, 0:3] = o
    shadow[:, 3] = 0.01
    shadow[:, 4:7] = dvec / dist[:, None]
    shadow[:, 7] = dist - 0.01
    gp, gt = gpu.intersect(shadow)
    rp, rt = cpu.intersect(shadow)
    _check_closest(sc, shadow, gp, gt, rp, rt, "shadow (closest)")
    go, _ = gpu.intersect(shadow, any_hit=True)
    ro, _ = cpu.intersect(shadow, any_hit=True)
    agree = go == ro
    print(f"shadow (any hit): {n} rays, agreement {agree.mean():.6f}, {ro.mean():.3f} occluded")
    assert agree.mean() >= 0.9999
    # an any-hit disagreement is adjudicated on the closest hit inside (tmin, tmax]
    verdicts = [classify(sc, shadow[i], int(gp[i]), int(rp[i])) for i in np.nonzero(~agree)[0]]
    assert "fail" not in verdicts, verdicts
    # any hit <=> a closest hit inside the interval, on each side
    assert np.array_equal(go == 1, gp >= 0) and np.array_equal(ro == 1, rp >= 0)


def test_gi_band_matches_reference_per_sample(ctx, bench_scene):
    sc, gpu, cpu = bench_scene
    y0, bh = 480, 192
    sx, sy, cw = 944, 80, 32      # the 32 x 32 sub-crop inside the band
    old = ctx.tuning()
    try:
        ctx.set_tuning(0, 0, 0, 0, 0)   # the library's defaults: the benchmarked launch shape
        S, Q, N, PS = gpu.render(W, H, GRID, DEPTH, SEED, crop=(0, y0, W, bh), per_sample=True)
        st = gpu.stats()
    finally:
        ctx.set_tuning(**old)
    # 1920 x 192 x 64 = 23.6M paths: three chunks of at most 2^23 over two lanes
    assert st["trace_launches"] == 3 * (DEPTH + 1)
    assert np.all(N == GRID * GRID)
    got = PS[sy:sy + cw, sx:sx + cw]
    _, _, _, ref, _ = cpu.render_keyed(W, H, sx, y0 + sy, cw, cw, GRID, 1, DEPTH, SEED, per_sample=True)
    close = np.all(np.abs(got - ref) <= 1e-4 + 1e-3 * np.abs(ref), axis=-1)
    print(f"bench scene: {close.size} camera samples, {close.mean():.5f} agree per sample")
    assert close.mean() >= 0.97
    n = GRID * GRID
    mg, mr = got.mean(axis=2), ref.mean(axis=2)
    se = np.sqrt(2 * (ref.var(axis=2, ddof=1) + 1e-12) / n)
    within = np.abs(mg - mr) <= 3 * se + 1e-6
    print(f"bench scene: per-pixel 3SE agreement {within.mean():.5f}")
    assert within.mean() >= 0.997
    # the accumulated sums are the per-sample values' sums
    np.testing.assert_allclose(S[sy:sy + cw, sx:sx + cw], got.astype(np.float64).sum(axis=2), rtol=1e-5, atol=1e-5)
