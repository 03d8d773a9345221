"""The reference's own Cornell asset, imported with bsh_import_obj and rendered on the GPU, against
the compiled reference's GiTracer on the same World (tests/golden/gi_cornell_obj.npz, written by
make_golden.py with BuildWorldCornellBox's camera, UnitTest/basic_features.h:235-270): sample for
sample on a crop, per pixel within 3 standard errors on the full frame."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("width", [2, 4])
def test_imported_cornell_matches_reference(ctx, width):
    from banshee_amd import hip
    from test_gpu_render import _per_sample_agreement
    fx = golden("gi_cornell_obj")
    sc = hip.import_obj(os.path.join(GOLDEN, "cornell_obj", "orig.objm"), fx["camera"])
    assert sc.digest() == str(fx["digest"]), "the import no longer builds the World the fixture was made on"
    s = hip.Scene(ctx, sc, width=width)
    W, H, grid, depth, seed = (int(fx[k]) for k in ("W", "H", "grid", "depth", "seed"))
    crop = tuple(int(v) for v in fx["crop"])
    _, _, _, PS = s.render(W, H, grid, depth, seed, crop=crop, per_sample=True)
    frac, _ = _per_sample_agreement(PS, fx["crop_samples"])
    print(f"imported Cornell box, width {width}: {frac:.5f} of the camera samples agree with the reference")
    assert frac >= 0.97
    S, Q, N, _ = s.render(W, H, grid, depth, seed)
    n = N[..., None]
    mg, mr = S / n, fx["sum"] / n
    vr = np.maximum(fx["sumsq"] / n - mr * mr, 0) * n / (n - 1)
    vg = np.maximum(Q / n - mg * mg, 0) * n / (n - 1)
    within = np.abs(mg - mr) <= 3 * np.sqrt((vr + vg) / n) + 1e-6
    print(f"imported Cornell box: per-pixel 3SE agreement {within.mean():.5f}")
    assert within.mean() >= 0.997
