"""Per-bundle transforms on the reference side, pinned by FP64 geometry: every recorded
closest-hit query of the reference rendering the transformed Cornell box (tests/golden/
transforms.npz; the reference intersects in object space through minv, mesh.cpp:216-261) hits
the triangle and distance that FP64 brute force finds among the world-space triangles
p' = m p (row-major m, translation in its last column, math/mathutils.h:129-136)."""
import numpy as np

from adjudicate import mt64
from conftest import golden


def world_scene(sc):
    """A copy of `sc` with its vertices moved to world space in FP64."""
    from banshee_amd.scenes import Scene
    out = Scene(**{k: getattr(sc, k) for k in sc.__dataclass_fields__})
    v = sc.verts.astype(np.float64).copy()
    for b, row in enumerate(sc.bundles):
        m = sc.transforms[b, :16].astype(np.float64).reshape(4, 4)
        vb, vc = int(row[2]), int(row[3])
        v[vb:vb + vc] = v[vb:vb + vc] @ m[:3, :3].T + m[:3, 3]
    out.verts = v
    out.transforms = None
    return out


def brute_force(ws, ray):
    best, bt = -1, np.inf
    for k in range(ws.tri_v.shape[0]):
        p1, p2, p3 = (ws.verts[i] for i in ws.tri_v[k])
        t, b1, b2, _ = mt64(ray[0:3], ray[4:7], p1, p2, p3)
        if np.isfinite(t) and b1 >= 0 and b2 >= 0 and b1 + b2 <= 1 and ray[3] < t <= ray[7] and t <= bt:
            best, bt = k, t
    return best, bt


def test_reference_transforms_are_world_space_geometry():
    from banshee_amd.scenes import transformed_cornell
    fx = golden("transforms")
    sc = transformed_cornell(int(fx["W"]), int(fx["W"]))
    assert sc.digest() == str(fx["digest"])
    ws = world_scene(sc)
    cl = np.nonzero(fx["kind"] == 0)[0][:1000]
    agree, close_t = 0, 0
    for i in cl:
        ray = fx["rays"][i].astype(np.float64)
        p, t = brute_force(ws, ray)
        agree += p == fx["prim"][i]
        if p == fx["prim"][i] and p >= 0:
            close_t += abs(t - fx["thit"][i, 0]) <= 1e-5 * t
    hits = sum(1 for i in cl if fx["prim"][i] >= 0)
    assert agree >= len(cl) - 2, (agree, len(cl))
    assert close_t >= hits - 2
