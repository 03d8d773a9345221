"""Guided probe baking on the device: the quirk-free bake whose probe paths continue from every
surface vertex with the one-sample MIS mixture of BSDF and learned guide.

* Without a guide, bsh_bake_sh_guided is the plain quirk-free bake (bsh_bake_sh with
  BSH_RENDER_NO_QUIRKS), bit for bit.
* With a guide trained on the probe paths' own records, the projected coefficients agree
  with the plain quirk-free bake within 5 standard errors per coefficient, and the mean of
  the z-scores is near 0. The standard error of c_i = (4 pi / N) sum_j L_j Y_i(w_j) is taken
  from the per-ray radiance and the probe directions restated from the keyed stream,
  treating the stratified rays as independent (which only overstates it).
"""
import numpy as np
import pytest

from test_gpu_sh import PROBES, _sh_fp64, probe_dirs

pytestmark = pytest.mark.gpu


def projection_and_se(rad, dirs):
    t = _sh_fp64(dirs)[..., None] * rad[:, :, None, :].astype(np.float64)   # [P, N, 9, 3]
    n = rad.shape[1]
    return 4 * np.pi * t.mean(1), 4 * np.pi * t.std(1, ddof=1) / np.sqrt(n)


def test_unguided_bake_is_the_quirk_free_bake(ctx, cornell):
    from banshee_amd import hip
    S = hip.Scene(ctx, cornell)
    c1, r1 = S.bake_sh(PROBES, 256, 2, 5, 9, radiance=True, quirk_free=True)
    c2, r2, n = S.bake_sh_guided(PROBES, 256, 2, 5, 9, radiance=True)
    assert n == 0
    assert np.array_equal(c1, c2) and np.array_equal(r1, r2)


def test_guided_bake_is_unbiased(ctx, cornell):
    from banshee_amd import hip
    S = hip.Scene(ctx, cornell)
    g = hip.Guide(ctx, seed=5, bsdf_fraction=0.3, lr=5e-3, beta2=0.99)
    rng = np.random.default_rng(17)
    train_probes = rng.uniform(0.1, 1.9, (64, 3)).astype(np.float32)
    Nt = 256
    cap = 64 * Nt * 4
    recs = ctx.alloc((cap, hip.RECORD_FLOATS), np.float32)
    n = 0
    for it in range(20):
        n = S.bake_sh_guided(train_probes, Nt, 2, 5, 100 + it, guide=g, records=recs, download=False)
        if n:
            g.train(recs, min(n, cap), 4, 1 << 14)
    ctx.sync()
    assert 0 < n <= cap
    r = recs.numpy()[:n]
    assert np.all((r[:, 0:3] >= -1e-5) & (r[:, 0:3] <= 1 + 1e-5))
    assert np.allclose(np.linalg.norm(r[:, 8:11], axis=1), 1, atol=1e-4)
    assert np.all(np.isfinite(r[:, 11])) and np.all(r[:, 11] > 0)
    assert np.all(np.isfinite(g.get_params()))

    N = 16384
    P = PROBES.shape[0]
    cp, rp, _ = S.bake_sh_guided(PROBES, N, 2, 5, 1, radiance=True)
    cg, rg, _ = S.bake_sh_guided(PROBES, N, 2, 5, 2, guide=g, radiance=True)
    mp, sp = projection_and_se(rp, probe_dirs(1, P, N))
    mg, sg = projection_and_se(rg, probe_dirs(2, P, N))
    # the restated directions reproduce the device projection of the same radiance
    assert np.allclose(mp, cp, rtol=1e-3, atol=1e-5) and np.allclose(mg, cg, rtol=1e-3, atol=1e-5)
    se = np.sqrt(sp ** 2 + sg ** 2)
    live = se > 1e-6 * np.abs(cp[:, :1, :])    # coefficients the scene's symmetry makes ~0 carry no noise
    z = (cg - cp)[live] / se[live]
    assert np.max(np.abs(z)) < 5.0, np.max(np.abs(z))
    assert abs(np.mean(z)) < 0.5, np.mean(z)
