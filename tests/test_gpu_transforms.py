"""Per-bundle transforms on the GPU (bsh_scene_set_transforms, baked into world space at commit)
against the reference, which intersects and shades in object space through m / minv: recorded
reference queries hit the same triangles at the same distance (disagreements adjudicated in FP64
on the world-space triangles), and a whole frame of the transformed Cornell box agrees per camera
sample (tests/golden/transforms.npz)."""
import numpy as np
import pytest

from adjudicate import classify
from conftest import golden
from test_transform_oracle import world_scene

pytestmark = pytest.mark.gpu

FX = golden("transforms")


def _scene(ctx, width=4):
    from banshee_amd import hip
    from banshee_amd.scenes import transformed_cornell
    sc = transformed_cornell(int(FX["W"]), int(FX["W"]))
    assert sc.digest() == str(FX["digest"])
    return sc, hip.Scene(ctx, sc, width=width)


@pytest.mark.parametrize("width", [2, 4])
def test_recorded_rays_match_reference(ctx, width):
    sc, s = _scene(ctx, width)
    rays, kind, prim, thit = FX["rays"], FX["kind"], FX["prim"], FX["thit"]
    cl = kind == 0
    gp, gt = s.intersect(rays[cl])
    rp, rt = prim[cl], thit[cl]
    same = gp == rp
    print(f"transformed width {width}: closest {cl.sum()} rays, prim agreement {same.mean():.6f}")
    assert same.mean() >= 0.9999 or (~same).sum() <= 2
    hit = same & (rp >= 0)
    # world-space against object-space arithmetic: t to 1e-5 relative or 1e-6 absolute (a few
    # ulp of the box's unit-sized coordinates, which set the rounding of a hit very near the
    # origin: 2 of 7880 hits at t ~ 0.025 differ by 5e-7), barycentrics to 1e-4
    np.testing.assert_allclose(gt[hit, 0], rt[hit, 0], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(gt[hit, 1:], rt[hit, 1:], rtol=1e-4, atol=1e-5)
    ws = world_scene(sc)
    bad = np.nonzero(~same)[0]
    verdicts = [classify(ws, rays[cl][i], int(gp[i]), int(rp[i])) for i in bad]
    assert "fail" not in verdicts, list(zip(bad, verdicts))
    an = kind == 1
    go, _ = s.intersect(rays[an], any_hit=True)
    agree = go == prim[an]
    print(f"transformed: any-hit {an.sum()} rays, agreement {agree.mean():.6f}")
    assert agree.mean() >= 0.9999 or (~agree).sum() <= 2


def test_frame_per_sample_matches_reference(ctx):
    sc, s = _scene(ctx)
    W, grid, depth, seed = (int(FX[k]) for k in ("W", "grid", "depth", "seed"))
    S, Q, N, PS = s.render(W, W, grid, depth, seed, per_sample=True)
    ref = FX["frame"]
    close = np.all(np.abs(PS - ref) <= 1e-4 + 1e-3 * np.abs(ref), axis=-1)
    st = s.stats()
    print(f"transformed: {close.mean():.5f} of {close.size} camera samples agree; rays "
          f"{st['rays_closest']} {st['rays_any']} vs {FX['ray_counts']}")
    assert close.mean() >= 0.97
    n = N[..., None]
    se = np.sqrt(2 * (ref.var(axis=2, ddof=1) + 1e-12) / n)
    assert (np.abs(S / n - ref.mean(axis=2)) <= 3 * se + 1e-6).mean() >= 0.99


def test_transform_count_must_match_bundles(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import transformed_cornell
    sc = transformed_cornell(16, 16)
    sc.transforms = sc.transforms[:-1]
    with pytest.raises(hip.BshError, match="one transform pair per bundle"):
        hip.Scene(ctx, sc)
