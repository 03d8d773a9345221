"""OBJ/MTL import (bsh_import_obj, host only) of the reference's own Cornell asset
(Resources/CornellBox/orig.objm + orig.mtl, copied under tests/golden/cornell_obj as data), checked
against what AssimpAssetImporter::Import (import/assimp_assetimporter.cpp:20-146) builds from it,
worked out by hand from the file: 16 quads in 8 (object, material) meshes, 32 triangles, the
"light" material Emissive(0.6 (60, 55, 40)) = (36, 33, 24), one AreaLight per light triangle."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

ASSET = os.path.join(GOLDEN, "cornell_obj", "orig.objm")
CAMERA = [0, 1, 3.5, 0, 1, 0, 0, 1, 0, 0.01, 10000, np.float32(np.pi / 4), 1.0]


def _lib_or_skip():
    from banshee_amd import hip
    if not os.path.exists(hip.LIB_PATH):
        pytest.skip("library not built")
    return hip


def test_cornell_asset_structure():
    hip = _lib_or_skip()
    sc = hip.import_obj(ASSET, CAMERA)
    assert sc.ntris == 32
    # meshes in file order: light | leftWall object: shortBox faces, leftWall face | backWall |
    # rightWall | ceiling | shortBox object: tallBox faces | floor
    assert sc.bundles[:, 1].tolist() == [2, 10, 2, 2, 2, 2, 10, 2]
    assert sc.bundles[:, 0].tolist() == [0, 2, 12, 14, 16, 18, 20, 30]
    names = ["Default", "backWall", "ceiling", "floor", "leftWall", "light", "rightWall", "shortBox", "tallBox"]
    mat_of_bundle = [names[sc.tri_m[b0]] for b0 in sc.bundles[:, 0]]
    assert mat_of_bundle == ["light", "shortBox", "leftWall", "backWall", "rightWall", "ceiling", "tallBox", "floor"]
    m = sc.materials
    assert m[5, 0] == 2 and m[5, 1:4].tolist() == [np.float32(0.6) * 60, np.float32(0.6) * 55, np.float32(0.6) * 40]
    assert m[5, 1:4].tolist() == [36.0, 33.0, 24.0]
    assert np.all(m[[0, 1, 2, 3, 4, 6, 7, 8], 0] == 0)   # Lambert
    np.testing.assert_array_equal(m[4, 1:4], np.array([0.504, 0.052, 0.040], np.float32))
    np.testing.assert_array_equal(m[6, 1:4], np.array([0.112, 0.360, 0.0728], np.float32))
    np.testing.assert_array_equal(m[0, 1:4], np.float32(0.6))
    # one AreaLight per triangle of the emissive mesh
    assert sc.lights.tolist() == [[0, 0], [0, 1]]
    # the light quad "f 1//1 2//1 3//1 4//1" becomes (1 2 3), (1 3 4) with normal (0, -1, 0)
    lines = [l.split() for l in open(ASSET) if l.startswith("v ")]
    P = np.array([[float(x) for x in l[1:4]] for l in lines], np.float32)
    tri = sc.verts[sc.tri_v[:2]]
    np.testing.assert_array_equal(tri[0], P[[0, 1, 2]])
    np.testing.assert_array_equal(tri[1], P[[0, 2, 3]])
    np.testing.assert_array_equal(sc.normals[sc.tri_n[:2]].reshape(-1, 3), np.tile([0, -1, 0], (6, 1)))
    # every quad of the file is convex, so it splits as (a b c), (a c d); the bundles keep the
    # file's face order
    quads = [l.split()[1:] for l in open(ASSET) if l.startswith("f ")]
    qidx = [[int(t.split("/")[0]) - 1 for t in q] for q in quads]
    for k, qi in enumerate(qidx):
        t = sc.verts[sc.tri_v[2 * k:2 * k + 2]]
        np.testing.assert_array_equal(t[0], P[[qi[0], qi[1], qi[2]]])
        np.testing.assert_array_equal(t[1], P[[qi[0], qi[2], qi[3]]])
    # vertices are shared inside a bundle (JoinIdenticalVertices): 4 per quad, not 6
    assert sc.verts.shape[0] == 64


def test_faces_without_normals_get_their_triangle_normal(tmp_path):
    hip = _lib_or_skip()
    (tmp_path / "a.mtl").write_text("newmtl red\nKd 0.5 0.1 0.1\nnewmtl glow\nKe 1 1 1\n")
    (tmp_path / "a.obj").write_text(
        "mtllib a.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 2 0 0\n"
        "usemtl red\nf 1 2 3 4\nusemtl glow\nf -1 -4 -3\nusemtl nosuch\nf 1 2 5\n")
    sc = hip.import_obj(str(tmp_path / "a.obj"), CAMERA)
    assert sc.ntris == 4 and sc.bundles.shape[0] == 3
    np.testing.assert_allclose(sc.normals[sc.tri_n[:2]].reshape(-1, 3), np.tile([0, 0, 1], (6, 1)))
    # Ke > 0 makes a material emissive whatever its name; an unknown usemtl takes the default
    assert sc.materials[sc.tri_m[2], 0] == 2 and sc.lights.tolist() == [[1, 0]]
    assert sc.tri_m[3] == 0 and sc.materials[0, 0] == 0


@pytest.mark.parametrize("body,msg", [("v 0 0 0\nf 1 2 3\n", "bad face index"),
                                      ("v 0 0 0\nv 1 0 0\nf 1 2\n", "fewer than 3"),
                                      ("mtllib missing.mtl\n", "cannot open material library")])
def test_malformed_files_are_rejected(tmp_path, body, msg):
    hip = _lib_or_skip()
    (tmp_path / "b.obj").write_text(body)
    with pytest.raises(hip.BshError, match=msg):
        hip.import_obj(str(tmp_path / "b.obj"), CAMERA)
    with pytest.raises(hip.BshError, match="cannot open"):
        hip.import_obj(str(tmp_path / "none.obj"), CAMERA)
