"""The quirk-free estimator at depth > 1 against an independent estimator, and the RMSE the
guided comparison reports against a converged image.

* The GPU's quirk-free tracer (next-event estimation with light and BSDF strategies under MIS
  in solid angle, BSH_RENDER_NO_QUIRKS) and the CPU's BSDF-sampling-only tracer
  (oracle/cpu_ref: no light sampling, no MIS, no light pdfs; emission collected where a path
  lands) estimate the same integral. On the Cornell box at depth 5 their block means agree
  within the block standard errors.
* On the occluded room, the per-pixel variance estimate of the RMSE (what the time-to-RMSE
  comparison uses) equals the RMSE against a 1024-spp render, for the plain and for the trained
  guided tracer, and the guided image has no bias against that render.
"""
import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.gpu


def _stats(s, q, n):
    s, q = s.astype(np.float64), q.astype(np.float64)
    n = np.asarray(n, np.float64)[..., None]
    m = s / n
    var = np.maximum(q / n - m * m, 0) * n / np.maximum(n - 1, 1)
    return m, var / n


def _blocks(m, v, b):
    H, W = m.shape[:2]
    h, w = H // b, W // b
    bm = m[:h * b, :w * b].reshape(h, b, w, b, 3).mean(axis=(1, 3))
    bv = v[:h * b, :w * b].reshape(h, b, w, b, 3).sum(axis=(1, 3)) / (b * b) ** 2
    return bm, bv


@pytest.mark.skipif(not oracle_lib.available("cpu"), reason="CPU restatement not built")
def test_quirk_free_equals_bsdf_only_estimator(ctx, cornell):
    from banshee_amd import hip
    S = hip.Scene(ctx, cornell)
    W, depth = 256, 5
    g = S.render_guided(W, W, 8, depth, 71)                       # 64 spp, NEE + MIS
    c = oracle_lib.Oracle("cpu").scene(cornell).render_bsdf_only(W, W, 16, depth, 72)   # 256 spp, BSDF only
    mg, vg = _stats(*g[:3])
    mc, vc = _stats(*c)
    bg_, bvg = _blocks(mg, vg, 32)
    bc, bvc = _blocks(mc, vc, 32)
    z = (bg_ - bc) / np.sqrt(bvg + bvc)
    print(f"64 blocks x 3 channels: max |z| {np.abs(z).max():.2f}, mean z {z.mean():.3f}, mean z^2 {np.mean(z * z):.3f}; "
          f"image mean {mg.mean():.5f} (NEE+MIS) vs {mc.mean():.5f} (BSDF only)")
    assert np.abs(z).max() < 4.5
    assert abs(z.mean()) < 0.4
    assert 0.5 < np.mean(z * z) < 1.6
    # a dropped bounce would show: the same test against depth 4 fails
    g4 = S.render_guided(W, W, 8, depth - 1, 73)
    m4, v4 = _stats(*g4[:3])
    b4, bv4 = _blocks(m4, v4, 32)
    z4 = (b4 - bc) / np.sqrt(bv4 + bvc)
    assert np.abs(z4).max() > 4.5 or abs(z4.mean()) > 0.4


def test_rmse_estimate_equals_rmse_against_a_converged_render(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import occluded_room
    W, depth = 256, 8
    sc = occluded_room(target_tris=120_000, width=W, height=W)
    S = hip.Scene(ctx, sc)
    g = hip.Guide(ctx, seed=1, bsdf_fraction=1.0, lr=5e-3, beta2=0.99)
    cap = W * W * 4 * 4
    recs = ctx.alloc((cap, hip.RECORD_FLOATS), np.float32)
    for it in range(24):
        g.set_bsdf_fraction(max(0.5, 1.0 - 0.5 * it / 4))
        n = S.render_guided(W, W, 2, depth, 600 + it, guide=g, records=recs, download=False)
        if n:
            g.train(recs, min(n, cap), 10, 1 << 18)
    ctx.sync()
    ref = S.render_guided(W, W, 32, depth, 700)           # 1024 spp plain
    mr, vr = _stats(*ref[:3])
    out = {}
    for name, guide, seed in (("plain", None, 801), ("guided", g, 802)):
        r = S.render_guided(W, W, 4, depth, seed, guide=guide)   # 16 spp
        m, v = _stats(*r[:3])
        proxy = np.sqrt(np.mean(v))
        # MSE against the reference, less the reference's own variance of the mean
        true = np.sqrt(max(np.mean((m - mr) ** 2) - np.mean(vr), 0.0))
        # bias: the image mean per channel against the reference's (pixels are independent)
        npx = m.shape[0] * m.shape[1]
        bias_z = (m.mean(axis=(0, 1)) - mr.mean(axis=(0, 1))) / np.sqrt((v.sum(axis=(0, 1)) + vr.sum(axis=(0, 1))) / npx ** 2)
        out[name] = (proxy, true, bias_z)
        print(f"{name}: RMSE estimate {proxy:.5f}, RMSE vs 1024 spp {true:.5f}, image-mean z {np.round(bias_z, 2)}")
        assert abs(true / proxy - 1) < 0.15, (name, true, proxy)
        assert np.all(np.abs(bias_z) < 4.5)
    assert out["guided"][1] < out["plain"][1]
