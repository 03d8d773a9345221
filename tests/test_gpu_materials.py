"""OrenNayar, PerfectReflect, PerfectRefract, Glass and MixedMaterial on the GPU (material rows
3-7): every row's Sample / Evaluate / GetPdf against the compiled reference's on 1024 cases
(Glass and MixedMaterial drawing their rand() from the same keyed stream), and a whole frame of a
box holding all of them per camera sample against the reference's GiTracer
(tests/golden/materials.npz)."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

FX = golden("materials")


def test_materials_match_reference(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import material_scene
    sc = material_scene()
    assert sc.digest() == str(FX["digest"])
    s = hip.Scene(ctx, sc)
    for k in range(sc.materials.shape[0]):
        got = s.diag_material(k, int(FX["probe_seed"]), FX["frames"], FX["view"], FX["u"], FX["ea"], FX["eb"])
        ref = FX["probes"][k]
        both_nan = np.isnan(got) & np.isnan(ref)
        close = (np.abs(got - ref) <= 1e-5 + 1e-4 * np.abs(ref)) | both_nan
        ok = close.all(axis=1)
        print(f"material {k} (type {int(sc.materials[k, 0])}): {ok.mean():.4f} of {len(ok)} cases agree")
        assert ok.mean() >= 0.99
        # the BSDF type flags and which branch Glass / MixedMaterial took agree exactly
        assert np.array_equal(got[:, 7], ref[:, 7])


def test_frame_per_sample_matches_reference(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import material_scene
    W, grid, depth, seed = (int(FX[k]) for k in ("W", "grid", "depth", "seed"))
    s = hip.Scene(ctx, material_scene())
    S, Q, N, PS = s.render(W, W, grid, depth, seed, per_sample=True)
    ref = FX["frame"]
    close = np.all(np.abs(PS - ref) <= 1e-4 + 1e-3 * np.abs(ref), axis=-1)
    st = s.stats()
    print(f"materials: {close.mean():.5f} of {close.size} camera samples agree; rays "
          f"{st['rays_closest']} {st['rays_any']} vs {FX['rays']}")
    assert close.mean() >= 0.97
    n = N[..., None]
    m_g, m_r = S / n, ref.mean(axis=2)
    se = np.sqrt(2 * (ref.var(axis=2, ddof=1) + 1e-12) / n)
    assert (np.abs(m_g - m_r) <= 3 * se + 1e-6).mean() >= 0.99
    assert st["rays_any"] == pytest.approx(FX["rays"][1], rel=0.01)


def test_bad_mixed_rows_are_rejected(ctx):
    from banshee_amd import hip
    from banshee_amd.scenes import material_scene
    sc = material_scene()
    sc.materials = sc.materials.copy()
    sc.materials[sc.meta["rows"]["mixed_floor"], 5] = sc.meta["rows"]["glass"]   # a Glass component
    with pytest.raises(hip.BshError, match="single BSDF"):
        hip.Scene(ctx, sc)
    sc.materials[sc.meta["rows"]["mixed_floor"], 5] = 99                          # out of range
    with pytest.raises(hip.BshError, match="out of range"):
        hip.Scene(ctx, sc)
