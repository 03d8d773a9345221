"""Texture coordinates and maps on the reference side (reference oracle, live), pinned by closed
forms on the textured Cornell box (banshee_amd.scenes.textured_cornell):

* rays straight down onto the floor quad (corners (0,0,0), (2,0,0), (2,0,2), (0,0,2), uv 0..2)
  report uv = (x, z) and FillHit's UV frame dpdu = +x, dpdv = +z (mesh.cpp:72-94);
* a Lambert row whose kd map is one constant colour renders, sample for sample, as the same row
  with that colour and no map (GET_VALUE, bsdf/bsdf.h:42), to the rounding of the bilinear blend
  of four equal texels ((1 - s) c + s c is c to an ulp)."""
import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.skipif(not oracle_lib.available("ref"), reason="reference oracle not built")


def test_floor_uv_and_frame():
    from banshee_amd.scenes import textured_cornell
    sc = textured_cornell()
    s = oracle_lib.Oracle("ref").scene(sc)
    rng = np.random.default_rng(4)
    xz = rng.uniform(0.05, 0.28, (200, 2))     # under no box, away from the quad's diagonal
    rays = np.zeros((200, 8), np.float32)
    rays[:, 0], rays[:, 1], rays[:, 2] = xz[:, 0], 0.5, xz[:, 1]
    rays[:, 3], rays[:, 5], rays[:, 7] = 0.01, -1.0, 1e6
    out = s.intersect_fill(rays)
    assert np.all(out[:, 19] >= 0)
    np.testing.assert_allclose(out[:, 16], xz[:, 0], atol=2e-6)
    np.testing.assert_allclose(out[:, 17], xz[:, 1], atol=2e-6)
    np.testing.assert_allclose(out[:, 10:13], np.tile([1, 0, 0], (200, 1)), atol=1e-6)
    np.testing.assert_allclose(out[:, 13:16], np.tile([0, 0, 1], (200, 1)), atol=1e-6)


def test_constant_map_equals_plain_colour():
    from banshee_amd.scenes import textured_cornell
    sc = textured_cornell()
    c = np.array([0.35, 0.6, 0.45], np.float32)
    a = textured_cornell()
    a.textures = [np.broadcast_to(c, (4, 4, 3)).copy()] + list(sc.textures[1:])
    b = textured_cornell()
    b.materials = b.materials.copy()
    b.materials[0, 1:4] = c
    b.mat_maps = b.mat_maps.copy()
    b.mat_maps[0, 0] = -1
    o = oracle_lib.Oracle("ref")
    pa = o.scene(a).render_keyed_frame(16, 16, ("stratified", "stratified", "stratified"), 2, 3, 5, per_sample=True)[3]
    pb = o.scene(b).render_keyed_frame(16, 16, ("stratified", "stratified", "stratified"), 2, 3, 5, per_sample=True)[3]
    np.testing.assert_allclose(pa, pb, rtol=1e-6, atol=1e-7)
