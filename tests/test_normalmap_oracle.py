"""Normal maps on the reference side (reference oracle, live): NormalMapping around the single
BSDFs of the normal-mapped Cornell box (banshee_amd.scenes.normal_mapped_cornell).

* A flat normal map (every texel (0.5, 0.5, 1): the tangent-space normal (0, 0, 1)) leaves
  Bsdf::MapNormal's result n' = normalize(n 1 + du 0 - dv 0) = n, however often it is applied
  (bsdf/bsdf.h:97-113), so the box renders, sample for sample, as the same box without normal
  maps, to the rounding of the extra normalisations.
* The bump map changes the shading of the mapped surfaces: the harness applies the wrapper to
  the rows the scene names and only to those."""
import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.skipif(not oracle_lib.available("ref"), reason="reference oracle not built")

W, GRID, DEPTH, SEED = 24, 2, 4, 3


def _frame(sc):
    s = oracle_lib.Oracle("ref").scene(sc)
    _, _, _, ps, rays = s.render_keyed_frame(W, W, ("stratified", "stratified", "stratified"), GRID, DEPTH, SEED,
                                             per_sample=True)
    return ps, rays


def test_flat_normal_map_renders_as_no_map():
    from banshee_amd.scenes import normal_mapped_cornell, textured_cornell
    flat = normal_mapped_cornell(W, W)
    flat.textures = list(flat.textures[:-1]) + [np.tile(np.array([0.5, 0.5, 1.0], np.float32), (16, 16, 1))]
    a, ra = _frame(flat)
    b, rb = _frame(textured_cornell(W, W))
    close = np.all(np.abs(a - b) <= 1e-5 + 1e-4 * np.abs(b), axis=-1)
    assert close.mean() >= 0.995, close.mean()
    assert abs(int(ra[0]) - int(rb[0])) <= 0.005 * rb[0]


def test_bump_map_changes_the_mapped_surfaces():
    from banshee_amd.scenes import normal_mapped_cornell, textured_cornell
    a, _ = _frame(normal_mapped_cornell(W, W))
    b, _ = _frame(textured_cornell(W, W))
    differ = np.any(np.abs(a - b) > 1e-3 + 1e-2 * np.abs(b), axis=-1)
    assert differ.mean() > 0.3, differ.mean()
    assert np.all(np.isfinite(a))
