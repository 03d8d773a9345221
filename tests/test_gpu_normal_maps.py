"""Normal maps on the GPU (bsh_scene_set_normal_maps): NormalMapping around Lambert, Microfacet
and OrenNayar rows agrees per camera sample with the reference rendering the same box through
its NormalMapping wrapper (tests/golden/normalmaps.npz). The reference rewrites the hit normal
in place in Sample and Evaluate, so a vertex's BSDF-sampled NEE ray sees the normal mapped twice
or once depending on its light sample's shadow ray; the GPU traces both and keeps the one the
shadow ray selects, which this test exercises on every normal-mapped vertex."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

FX = golden("normalmaps")


def test_frame_per_sample_matches_reference(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import normal_mapped_cornell
    W, grid, depth, seed = (int(FX[k]) for k in ("W", "grid", "depth", "seed"))
    sc = normal_mapped_cornell(W, W)
    assert sc.digest() == str(FX["digest"])
    s = hip.Scene(ctx, sc)
    S, Q, N, PS = s.render(W, W, grid, depth, seed, per_sample=True)
    ref = FX["frame"]
    close = np.all(np.abs(PS - ref) <= 1e-4 + 1e-3 * np.abs(ref), axis=-1)
    st = s.stats()
    print(f"normal maps: {close.mean():.5f} of {close.size} camera samples agree; rays "
          f"{st['rays_closest']} {st['rays_any']} vs {FX['ray_counts']}")
    assert close.mean() >= 0.97
    n = N[..., None]
    se = np.sqrt(2 * (ref.var(axis=2, ddof=1) + 1e-12) / n)
    assert (np.abs(S / n - ref.mean(axis=2)) <= 3 * se + 1e-6).mean() >= 0.99


def test_flat_map_in_the_quirk_free_estimator_is_no_map(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import normal_mapped_cornell, textured_cornell
    flat = normal_mapped_cornell(24, 24)
    flat.textures = list(flat.textures[:-1]) + [np.tile(np.array([0.5, 0.5, 1.0], np.float32), (16, 16, 1))]
    a = hip.Scene(ctx, flat).render_guided(24, 24, 2, 4, 5)[0]
    b = hip.Scene(ctx, textured_cornell(24, 24)).render_guided(24, 24, 2, 4, 5)[0]
    close = np.all(np.abs(a - b) <= 1e-5 + 1e-4 * np.abs(b), axis=-1)
    assert close.mean() >= 0.99, close.mean()


def test_bad_normal_maps_are_refused(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import normal_mapped_cornell
    sc = normal_mapped_cornell(16, 16)
    sc.normal_maps = sc.normal_maps.copy()
    sc.normal_maps[5] = 0                      # row 5 is Glass
    with pytest.raises(hip.BshError, match="takes no normal map"):
        hip.Scene(ctx, sc)
    sc = normal_mapped_cornell(16, 16)
    sc.normal_maps = sc.normal_maps.copy()
    sc.normal_maps[0] = len(sc.textures)
    with pytest.raises(hip.BshError, match="normal map out of range"):
        hip.Scene(ctx, sc)
    sc = normal_mapped_cornell(16, 16)
    sc.normal_maps = sc.normal_maps[:-1].copy()
    with pytest.raises(hip.BshError, match="one entry per material row"):
        hip.Scene(ctx, sc)
