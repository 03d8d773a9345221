"""Texture coordinates and texture maps on the GPU (bsh_scene_set_uvs, bsh_scene_set_textures,
bsh_scene_set_material_maps): a Cornell box with UVs on every face and kd / ks / kr maps on
Lambert, Microfacet, OrenNayar and Glass rows agrees per camera sample with the reference
rendering it through its texture system (tests/golden/textures.npz), and bad maps are refused."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

FX = golden("textures")


def test_frame_per_sample_matches_reference(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import textured_cornell
    W, grid, depth, seed = (int(FX[k]) for k in ("W", "grid", "depth", "seed"))
    sc = textured_cornell(W, W)
    assert sc.digest() == str(FX["digest"])
    s = hip.Scene(ctx, sc)
    S, Q, N, PS = s.render(W, W, grid, depth, seed, per_sample=True)
    ref = FX["frame"]
    close = np.all(np.abs(PS - ref) <= 1e-4 + 1e-3 * np.abs(ref), axis=-1)
    st = s.stats()
    print(f"textured: {close.mean():.5f} of {close.size} camera samples agree; rays "
          f"{st['rays_closest']} {st['rays_any']} vs {FX['ray_counts']}")
    assert close.mean() >= 0.97
    n = N[..., None]
    se = np.sqrt(2 * (ref.var(axis=2, ddof=1) + 1e-12) / n)
    assert (np.abs(S / n - ref.mean(axis=2)) <= 3 * se + 1e-6).mean() >= 0.99


def test_bad_maps_are_refused(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import textured_cornell
    sc = textured_cornell(16, 16)
    sc.mat_maps = sc.mat_maps.copy()
    sc.mat_maps[1, 0] = 9                      # no map 9
    with pytest.raises(hip.BshError, match="texture map out of range"):
        hip.Scene(ctx, sc)
    sc = textured_cornell(16, 16)
    sc.mat_maps = sc.mat_maps.copy()
    sc.mat_maps[0, 1] = 0                      # a kr map on a Lambert row
    with pytest.raises(hip.BshError, match="takes no such map"):
        hip.Scene(ctx, sc)
    sc = textured_cornell(16, 16)
    sc.tri_uv = sc.tri_uv.copy()
    sc.tri_uv[3, 1] = sc.uvs.shape[0]
    with pytest.raises(hip.BshError, match="uv index out of range at triangle 3"):
        hip.Scene(ctx, sc)
