"""Ray-query parity (GPU): hits of recorded reference rays, adjudicated in FP64."""
import numpy as np
import pytest

from adjudicate import classify
from conftest import golden, scene_of

pytestmark = pytest.mark.gpu


def _scene(ctx, sc, width=2):
    return scene_of(ctx, sc, width)


@pytest.mark.parametrize("width", [2, 4, "4q"])
@pytest.mark.parametrize("name", ["cornell", "terrain"])
def test_recorded_rays_match_reference(ctx, cornell, terrain100k, name, width):
    fx = golden(f"rays_{name}")
    sc = cornell if name == "cornell" else terrain100k
    assert str(fx["digest"]) == sc.digest(), "scene generator drifted from the fixture"
    s = _scene(ctx, sc, width)
    rays, kind, prim, thit = fx["rays"], fx["kind"], fx["prim"], fx["thit"]
    cl = kind == 0
    gp, gt = s.intersect(rays[cl])
    rp, rt = prim[cl], thit[cl]
    same = gp == rp
    print(f"{name} width {width}: closest {cl.sum()} rays, prim agreement {same.mean():.6f}")
    assert same.mean() >= 0.9999 or (~same).sum() <= 1
    hit = same & (rp >= 0)
    # identical Moller-Trumbore arithmetic: t and barycentrics must agree to 1e-5 relative
    np.testing.assert_allclose(gt[hit, 0], rt[hit, 0], rtol=1e-5)
    np.testing.assert_allclose(gt[hit, 1:], rt[hit, 1:], rtol=1e-5, atol=1e-6)
    bad = np.nonzero(~same)[0]
    verdicts = [classify(sc, rays[cl][i], int(gp[i]), int(rp[i])) for i in bad]
    assert "fail" not in verdicts, list(zip(bad, verdicts))
    an = kind == 1
    go, _ = s.intersect(rays[an], any_hit=True)
    agree = go == prim[an]
    print(f"{name}: any-hit {an.sum()} rays, agreement {agree.mean():.6f}")
    assert agree.mean() >= 0.9999 or (~agree).sum() <= 1


def _brute_force_fp32(sc, rays):
    """All-pairs Moller-Trumbore in float32 numpy with the reference's operation order
    (mesh.cpp:7-45); closest accepted hit, later index wins ties like t <= tmax."""
    V = sc.verts.astype(np.float32)
    p1 = V[sc.tri_v[:, 0]][None]
    e1 = (V[sc.tri_v[:, 1]] - V[sc.tri_v[:, 0]])[None]
    e2 = (V[sc.tri_v[:, 2]] - V[sc.tri_v[:, 0]])[None]
    o = rays[:, None, 0:3]
    d = rays[:, None, 4:7]

    def cross(a, b):
        return np.stack([a[..., 1] * b[..., 2] - b[..., 1] * a[..., 2], b[..., 0] * a[..., 2] - a[..., 0] * b[..., 2],
                         a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)

    def dot(a, b):
        return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]

    with np.errstate(all="ignore"):
        s1 = cross(d, e2)
        invd = np.float32(1) / dot(s1, e1)
        dd = o - p1
        b1 = dot(dd, s1) * invd
        s2 = cross(dd, e1)
        b2 = dot(d, s2) * invd
        t = dot(e2, s2) * invd
        ok = (b1 >= 0) & (b1 <= 1) & (b2 >= 0) & (b1 + b2 <= 1) & (t > rays[:, None, 3]) & (t <= rays[:, None, 7])
    t = np.where(ok, t, np.inf)
    best = np.full(len(rays), -1)
    tb = np.full(len(rays), np.inf, np.float32)
    for k in range(t.shape[1]):
        upd = t[:, k] <= tb
        best[upd] = k
        tb[upd] = t[upd, k]
    best[~np.isfinite(tb)] = -1
    return best, tb


@pytest.mark.parametrize("width", [2, 4, "4q"])
def test_brute_force_cornell(ctx, cornell, width):
    rng = np.random.default_rng(5)
    n = 20000
    o = rng.uniform(0.05, 1.95, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.full((n, 1), 0.01), d, np.full((n, 1), 1e6)], 1).astype(np.float32)
    s = _scene(ctx, cornell, width)
    gp, gt = s.intersect(rays)
    bp, bt = _brute_force_fp32(cornell, rays)
    same = gp == bp
    bad = np.nonzero(~same)[0]
    verdicts = [classify(cornell, rays[i], int(gp[i]), int(bp[i])) for i in bad]
    assert "fail" not in verdicts
    np.testing.assert_array_equal(gt[same & (bp >= 0), 0], bt[same & (bp >= 0)])
    occ, _ = s.intersect(rays, any_hit=True)
    assert np.array_equal(occ == 1, bp >= 0) or (occ != (bp >= 0)).sum() <= len(bad)


@pytest.mark.parametrize("width", [2, 4, "4q"])
def test_edge_cases(ctx, cornell, width):
    s = _scene(ctx, cornell, width)
    # tmax == t accepted, tmin == t rejected (mesh.cpp:33); floor at y = 0
    ray_down = np.array([[1.0, 1.0, 1.0, 0.01, 0.0, -1.0, 0.0, 1.0]], np.float32)
    p, t = s.intersect(ray_down)
    assert p[0] >= 0 and t[0, 0] == np.float32(1.0)
    ray_short = ray_down.copy()
    ray_short[0, 7] = 0.999
    p, _ = s.intersect(ray_short)
    assert p[0] == -1
    ray_tmin = ray_down.copy()
    ray_tmin[0, 3] = 1.0
    ray_tmin[0, 7] = 5.0
    p, _ = s.intersect(ray_tmin)
    assert p[0] == -1
    # empty input
    p, t = s.intersect(np.zeros((0, 8), np.float32))
    assert p.shape == (0,)
