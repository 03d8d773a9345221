"""Guided probe baking on the device: the quirk-free bake guided at every vertex of a probe path,
the probe itself included (its rays drawn from the uniform/guide mixture and weighted by
1 / (4 pi p_mix)).

* Without a guide, bsh_bake_sh_guided is the plain quirk-free bake (bsh_bake_sh with
  BSH_RENDER_NO_QUIRKS), bit for bit, and its rays are the stratified directions, weight 1.
* A guide whose outputs are NaN falls back to exactly that bake.
* With a guide trained on the probe paths' own records, the projected coefficients agree
  with the plain quirk-free bake within 5 standard errors per coefficient, and the mean of
  the z-scores is near 0. The standard error of c_i = (4 pi / N) sum_j w_j L_j Y_i(w_j) comes
  from every ray's radiance, direction and weight, treating the rays as independent (which
  only overstates it for the stratified part).
* The trained guide lowers the coefficients' variance against the plain bake.
"""
import numpy as np
import pytest

from test_gpu_sh import PROBES, _sh_fp64, probe_dirs

pytestmark = pytest.mark.gpu


def projection_and_se(rad, dirs):
    """rad [P, N, 3], dirs [P, N, 4] (direction, weight) -> mean and standard error [P, 9, 3]."""
    t = (_sh_fp64(dirs[..., :3].astype(np.float64))[..., None] * dirs[..., 3:4, None].astype(np.float64)
         * rad[:, :, None, :].astype(np.float64))
    n = rad.shape[1]
    return 4 * np.pi * t.mean(1), 4 * np.pi * t.std(1, ddof=1) / np.sqrt(n)


def train_on_probes(S, g, ctx, passes=20, seed0=100):
    from banshee_amd import hip
    rng = np.random.default_rng(17)
    train_probes = rng.uniform(0.1, 1.9, (64, 3)).astype(np.float32)
    Nt = 256
    cap = 64 * Nt * 4
    recs = ctx.alloc((cap, hip.RECORD_FLOATS), np.float32)
    n = 0
    for it in range(passes):
        g.set_bsdf_fraction(max(0.5, 1.0 - 0.5 * it / 6))
        n = S.bake_sh_guided(train_probes, Nt, 2, 5, seed0 + it, guide=g, records=recs, download=False)
        if n:
            g.train(recs, min(n, cap), 4, 1 << 14)
    ctx.sync()
    return recs.numpy()[:min(n, cap)]


def test_unguided_bake_is_the_quirk_free_bake(ctx, cornell):
    from banshee_amd import hip
    S = hip.Scene(ctx, cornell)
    c1, r1 = S.bake_sh(PROBES, 256, 2, 5, 9, radiance=True, quirk_free=True)
    c2, r2, d2, n = S.bake_sh_guided(PROBES, 256, 2, 5, 9, radiance=True, ray_dirs=True)
    assert n == 0
    assert np.array_equal(c1, c2) and np.array_equal(r1, r2)
    assert np.all(d2[..., 3] == 1)
    assert np.allclose(d2[..., :3], probe_dirs(9, PROBES.shape[0], 256), atol=2e-6)


def test_non_finite_guide_bakes_like_the_plain_bake(ctx, cornell):
    from banshee_amd import hip
    S = hip.Scene(ctx, cornell)
    g = hip.Guide(ctx, seed=1, bsdf_fraction=0.5)
    g.set_params(np.full(hip.Guide.num_params(), np.nan, np.float32))
    c1, r1, _ = S.bake_sh_guided(PROBES, 256, 2, 5, 13, radiance=True)
    c2, r2, _ = S.bake_sh_guided(PROBES, 256, 2, 5, 13, guide=g, radiance=True)
    assert np.array_equal(c1, c2) and np.array_equal(r1, r2)


def test_guided_bake_is_unbiased(ctx, cornell):
    from banshee_amd import hip
    S = hip.Scene(ctx, cornell)
    g = hip.Guide(ctx, seed=5, bsdf_fraction=1.0, lr=5e-3, beta2=0.99)
    r = train_on_probes(S, g, ctx)
    assert len(r) > 0
    assert np.all((r[:, 0:3] >= -1e-5) & (r[:, 0:3] <= 1 + 1e-5))
    assert np.allclose(np.linalg.norm(r[:, 8:11], axis=1), 1, atol=1e-4)
    assert np.all(np.isfinite(r[:, 11])) and np.all(r[:, 11] > 0)
    probe_rec = r[:, 7] == 0                    # the probe vertex's own records: normal 0, roughness 0
    assert probe_rec.any() and (~probe_rec).any()
    assert np.all(r[probe_rec, 3] == 0) and np.all(r[probe_rec, 4:7] == 0)
    assert np.all(np.isfinite(g.get_params()))

    N = 16384
    cp, rp, dp, _ = S.bake_sh_guided(PROBES, N, 2, 5, 1, radiance=True, ray_dirs=True)
    cg, rg, dg, _ = S.bake_sh_guided(PROBES, N, 2, 5, 2, guide=g, radiance=True, ray_dirs=True)
    assert np.any(dg[..., 3] != 1)              # the probe rays were guided
    mp, sp = projection_and_se(rp, dp)
    mg, sg = projection_and_se(rg, dg)
    # the returned rays reproduce the device projection
    assert np.allclose(mp, cp, rtol=1e-3, atol=1e-5) and np.allclose(mg, cg, rtol=1e-3, atol=1e-5)
    se = np.sqrt(sp ** 2 + sg ** 2)
    live = se > 1e-6 * np.abs(cp[:, :1, :])
    z = (cg - cp)[live] / se[live]
    assert np.max(np.abs(z)) < 5.0, np.max(np.abs(z))
    assert abs(np.mean(z)) < 0.5, np.mean(z)


def test_trained_probe_guide_lowers_variance(ctx, cornell):
    """Cornell box probes see the small bright light directly: a guide trained on the probe
    paths lowers the coefficients' variance (estimated from the rays) against the plain bake
    of the same ray count."""
    from banshee_amd import hip
    S = hip.Scene(ctx, cornell)
    g = hip.Guide(ctx, seed=5, bsdf_fraction=1.0, lr=5e-3, beta2=0.99)
    train_on_probes(S, g, ctx)
    N = 4096
    _, rp, dp, _ = S.bake_sh_guided(PROBES, N, 2, 5, 31, radiance=True, ray_dirs=True)
    _, rg, dg, _ = S.bake_sh_guided(PROBES, N, 2, 5, 32, guide=g, radiance=True, ray_dirs=True)
    _, sp = projection_and_se(rp, dp)
    _, sg = projection_and_se(rg, dg)
    ratio = float(np.sum(sg ** 2) / np.sum(sp ** 2))
    print(f"guided / plain coefficient variance: {ratio:.3f}")
    assert ratio < 0.8, ratio
