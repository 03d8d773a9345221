"""PointLight, DirectionalLight, EnvironmentLight and EnvironmentLightIs on the GPU
(bsh_scene_set_other_lights, bsh_scene_set_env_map): the device lights against the compiled
reference's (tests/golden/lights.npz), whole frames of a scene holding every light type per
camera sample against the reference's GiTracer, and closed forms of the light reflected by a
Lambert floor (point, directional, constant environment)."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

FX = golden("lights")


def _close(a, b, rtol, atol):
    return np.abs(a - b) <= atol + rtol * np.abs(b)


@pytest.mark.parametrize("tag", ["env", "envis"])
def test_device_lights_match_reference(ctx, tag):
    from banshee_amd import hip
    from banshee_amd.scenes import ENV_LIGHT, ENV_LIGHT_IS, open_scene
    sc = open_scene(ENV_LIGHT if tag == "env" else ENV_LIGHT_IS)
    assert sc.digest() == str(FX[f"digest_{tag}"])
    s = hip.Scene(ctx, sc)
    ref = FX[f"probe_{tag}"]
    for k in range(ref.shape[0]):
        got = s.diag_light_full(k, FX["p"], FX["nrm"], FX["u"], FX["w"])
        r = ref[k]
        # le, pdf and GetLe through sin/cos/atan2/acos/pow of two libraries: a few ulp, more where
        # the map is steep; directions to 1e-5
        ok = np.all(_close(got[:, 0:4], r[:, 0:4], 1e-4, 1e-6), 1) & np.all(_close(got[:, 4:7], r[:, 4:7], 1e-5, 1e-5), 1)
        ok &= _close(got[:, 7], r[:, 7], 1e-4, 1e-7) & np.all(_close(got[:, 8:11], r[:, 8:11], 1e-4, 1e-6), 1)
        print(f"{tag} light {k}: {ok.mean():.4f} of {len(ok)} probes agree")
        assert ok.mean() >= 0.99


@pytest.mark.parametrize("tag", ["env", "envis"])
def test_frame_per_sample_matches_reference(ctx, tag):
    from banshee_amd import hip
    from banshee_amd.scenes import ENV_LIGHT, ENV_LIGHT_IS, open_scene
    sc = open_scene(ENV_LIGHT if tag == "env" else ENV_LIGHT_IS)
    W, grid, depth, seed = (int(FX[k]) for k in ("W", "grid", "depth", "seed"))
    s = hip.Scene(ctx, sc)
    S, Q, N, PS = s.render(W, W, grid, depth, seed, per_sample=True)
    ref = FX[f"frame_{tag}"]
    close = np.all(np.abs(PS - ref) <= 1e-4 + 1e-3 * np.abs(ref), axis=-1)
    st = s.stats()
    print(f"{tag}: {close.mean():.5f} of {close.size} camera samples agree; rays "
          f"{st['rays_closest']} {st['rays_any']} vs {FX['rays_' + tag]}")
    assert close.mean() >= 0.97
    n = N[..., None]
    m_g, m_r = S / n, ref.mean(axis=2)
    se = np.sqrt(2 * (ref.var(axis=2, ddof=1) + 1e-12) / n)
    assert (np.abs(m_g - m_r) <= 3 * se + 1e-6).mean() >= 0.99
    assert st["rays_any"] == pytest.approx(FX[f"rays_{tag}"][1], rel=0.01)


def _floor_render(ctx, sc, quirk_free, spp_grid=8):
    """Image mean and its standard error of a 32 x 32, depth-1 render (BSH_RENDER_NO_QUIRKS when
    quirk_free), run as a one-job batch so the estimator flag can be set."""
    from banshee_amd import hip
    s = hip.Scene(ctx, sc)
    job = hip.BatchJob(s, "render", width=32, height=32, img_grid=spp_grid, max_depth=1, seed=11)
    if quirk_free:
        job.params.flags |= 2
    ctx.batch_run([job])
    a = job.accum.numpy().astype(np.float64)
    s_, n_, q_ = a[..., 0:3], a[..., 3:4], a[..., 4:7]
    mean = s_ / n_
    var = np.maximum(q_ / n_ - mean * mean, 0) * n_ / np.maximum(n_ - 1, 1)
    return mean.mean(axis=(0, 1)), np.sqrt((var / n_).sum(axis=(0, 1))) / (32 * 32)


@pytest.mark.parametrize("quirk_free", [False, True])
def test_point_and_directional_closed_form(ctx, quirk_free):
    """Singular lights have no MIS and no randomness at depth 1: albedo / pi e cos / d^2 and
    albedo / pi e cos, in either estimator."""
    from banshee_amd.scenes import DIRECTIONAL_LIGHT, POINT_LIGHT, floor_under
    target = np.array([0.3, 0.0, 0.2])
    P, e = np.array([0.8, 1.7, -0.4]), np.array([3.0, 2.0, 1.0])
    m, _ = _floor_render(ctx, floor_under([POINT_LIGHT, *P, *e, 0, 0]), quirk_free)
    d = P - target
    want = 0.5 / np.pi * e * (d[1] / np.linalg.norm(d)) / np.dot(d, d)
    print("point", m, want)
    np.testing.assert_allclose(m, want, rtol=2e-4)
    D = np.array([-0.3, -1.0, -0.4])
    m, _ = _floor_render(ctx, floor_under([DIRECTIONAL_LIGHT, *D, *e, 0, 0]), quirk_free)
    want = 0.5 / np.pi * e * (-D[1] / np.linalg.norm(D))
    print("directional", m, want)
    np.testing.assert_allclose(m, want, rtol=2e-4)


@pytest.mark.parametrize("kind", ["env", "envis"])
def test_constant_environment_closed_form(ctx, kind):
    """Quirk-free, depth 1: a Lambert floor under an unoccluded environment of constant radiance c
    (gamma 1, scale s) reflects albedo s c, whichever of the light's two strategies draws it."""
    from banshee_amd.scenes import ENV_LIGHT, ENV_LIGHT_IS, floor_under
    c = np.array([0.7, 1.1, 1.6], np.float32)
    env = np.broadcast_to(c, (16, 32, 3)).copy()
    row = [ENV_LIGHT if kind == "env" else ENV_LIGHT_IS, 0, 0, 0, 0, 0, 0, 1.5, 1.0]
    m, se = _floor_render(ctx, floor_under(row, env=env), True, spp_grid=16)
    want = 0.5 * 1.5 * c
    # the cosine-sampled light and the Lambert BSDF draw the same density, so with MIS every sample
    # is albedo s c up to rounding (standard error 0): the bound adds FP32 rounding, 1e-5 relative
    print(kind, m, want, se)
    assert np.all(np.abs(m - want) <= 4.5 * se + 1e-5 * want)
    assert np.all(np.abs(m / want - 1) < 0.01)


def test_batched_open_scenes_equal_solo(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import ENV_LIGHT, ENV_LIGHT_IS, cornell_box, open_scene
    scenes = [hip.Scene(ctx, open_scene(ENV_LIGHT)), hip.Scene(ctx, open_scene(ENV_LIGHT_IS)),
              hip.Scene(ctx, cornell_box())]
    jobs = [hip.BatchJob(s, "render", width=24, height=24, img_grid=4, max_depth=4, seed=4) for s in scenes]
    solo = [s.render(24, 24, 4, 4, 4) for s in scenes]
    ctx.batch_run(jobs, tuning={"use_graph": 1})
    for j, (S, Q, N, _) in zip(jobs, solo):
        a = j.accum.numpy()
        np.testing.assert_array_equal(a[..., 0:3], S)
        np.testing.assert_array_equal(a[..., 3], N)
