"""Closed-form and brute-force pins of the CPU oracles (CPU only).

The restatement (oracle/cpu_ref) is bit-exact with the compiled reference on the golden
fixtures; these tests pin both against what the mathematics fixes, so that a wrong sign,
index or dropped term in either would show: Moller-Trumbore on hand triangles, the BVH
against brute force, Lambert and GGX closed forms and sample/pdf consistency, the area
light's pdfs, glibc rand / MWC sequences from their published definitions, the camera
frustum, and the closed forms of the SH tables (Y_lm, the cosine-lobe
convolution A_l = pi, 2 pi/3, pi/4, 0, -pi/24).
"""
import math

import numpy as np
import pytest

import oracle_lib
from banshee_amd.scenes import EMISSIVE, LAMBERT, MICROFACET, _Builder, terrain_for_tris

KINDS = [k for k in ("cpu", "ref") if oracle_lib.available(k)]


@pytest.fixture(scope="module", params=KINDS)
def orc(request):
    return oracle_lib.Oracle(request.param)


def one_triangle_scene():
    B = _Builder()
    B.bundle([[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 0, 1]] * 3, [[0, 1, 2]], B.material(LAMBERT, (0.5, 0.5, 0.5)))
    return B.build([0, 0, 5, 0, 0, 0, 0, 1, 0, 0.01, 1e4, 0.8, 1.0], "tri")


def test_triangle_closed_forms(orc):
    s = orc.scene(one_triangle_scene())
    rays = np.array([
        [0.25, 0.25, 1, 0.0, 0, 0, -1, 10],     # interior: t = 1, b = (0.25, 0.25)
        [0.5, 0.5, 1, 0.0, 0, 0, -1, 10],       # on the hypotenuse b1 + b2 = 1: accepted
        [0.25, 0.25, 1, 0.0, 1, 0, 0, 10],      # parallel to the plane: miss
        [0.25, 0.25, 1, 0.0, 0, 0, -1, 1.0],    # t == tmax: accepted
        [0.25, 0.25, 1, 1.0, 0, 0, -1, 10],     # t == tmin: rejected
        [0.25, 0.25, -1, 0.0, 0, 0, 1, 10],     # from behind: no culling
        [0.8, 0.8, 1, 0.0, 0, 0, -1, 10],       # outside
    ], np.float32)
    prim, th = s.intersect(rays)
    assert list(prim) == [0, 0, -1, 0, -1, 0, -1]
    assert np.allclose(th[0], [1.0, 0.25, 0.25])
    assert abs(th[1, 1] + th[1, 2] - 1) < 1e-6
    assert th[5, 0] == 1.0
    occ, _ = s.intersect(rays, any_hit=True)
    assert list(occ) == [1, 1, 0, 1, 0, 1, 0]


def brute_force(scene, rays):
    """FP64 closest hit over every triangle, with the accept rule tmin < t <= tmax."""
    v = scene.verts.astype(np.float64)
    p1, p2, p3 = v[scene.tri_v[:, 0]], v[scene.tri_v[:, 1]], v[scene.tri_v[:, 2]]
    e1, e2 = p2 - p1, p3 - p1
    best = np.full(len(rays), -1)
    bt = np.full(len(rays), np.inf)
    for i, r in enumerate(rays.astype(np.float64)):
        o, d = r[0:3], r[4:7]
        s1 = np.cross(d, e2)
        det = np.einsum("ij,ij->i", s1, e1)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1 / det
            dd = o - p1
            b1 = np.einsum("ij,ij->i", dd, s1) * inv
            s2 = np.cross(dd, e1)
            b2 = (s2 @ d) * inv
            t = np.einsum("ij,ij->i", e2, s2) * inv
        ok = (b1 >= 0) & (b1 <= 1) & (b2 >= 0) & (b1 + b2 <= 1) & (t > r[3]) & (t <= r[7])
        if ok.any():
            j = np.argmin(np.where(ok, t, np.inf))
            best[i], bt[i] = j, t[j]
    return best, bt


def test_bvh_equals_brute_force(orc):
    sc = terrain_for_tris(3000, seed=4, width=64, height=64)
    s = orc.scene(sc)
    rng = np.random.default_rng(0)
    lo, hi = sc.verts.min(0), sc.verts.max(0)
    n = 400
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.full((n, 1), 0.01), d, np.full((n, 1), 1e4)], 1).astype(np.float32)
    prim, th = s.intersect(rays)
    bp, bt = brute_force(sc, rays)
    from adjudicate import classify
    hit = bp >= 0
    assert hit.sum() > n // 4
    assert np.array_equal(prim >= 0, hit)
    assert np.allclose(th[hit, 0], bt[hit], rtol=1e-5)
    for i in np.nonzero(prim != bp)[0]:
        assert classify(sc, rays[i], prim[i], bp[i]) != "fail"


def frames_up(n):
    f = np.zeros((n, 12), np.float32)
    f[:, 2] = f[:, 5] = 1          # n = ng = +z
    f[:, 6] = 1                    # dpdu = +x
    f[:, 10] = 1                   # dpdv = +y
    return f


def upper(rng, n):
    d = rng.normal(size=(n, 3))
    d[:, 2] = np.abs(d[:, 2]) + 1e-3
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def test_lambert_closed_forms(orc):
    rng = np.random.default_rng(1)
    n = 20000
    kd = np.array([0.3, 0.5, 0.7])
    mat = np.array([LAMBERT, *kd, 0, 0, 0, 0], np.float32)
    view, ea, eb = upper(rng, n), upper(rng, n), upper(rng, n)
    out = orc.bsdf(mat, frames_up(n), view, rng.random((n, 2)), ea, eb)
    wo = out[:, 4:7]
    assert np.allclose(out[:, 0:3], kd / np.pi, rtol=1e-6)
    assert np.all(wo[:, 2] >= 0)
    assert np.allclose(out[:, 3], wo[:, 2] / np.pi, rtol=1e-4, atol=1e-6)
    assert np.allclose(out[:, 8:11], kd / np.pi, rtol=1e-6)
    # cosine-weighted: cos^2(theta) is uniform on [0, 1]
    c2 = np.clip(wo[:, 2].astype(np.float64) ** 2, 0, 1)
    counts = np.bincount(np.minimum((c2 * 20).astype(int), 19), minlength=20)
    chi2 = np.sum((counts - n / 20) ** 2 / (n / 20))
    assert chi2 < 55   # 19 dof


def test_ggx_closed_forms(orc):
    rng = np.random.default_rng(2)
    a, ks = 0.3, np.array([0.8, 0.6, 0.4])
    mat = np.array([MICROFACET, *ks, 1.5, a, 0, 0], np.float32)
    # normal incidence: D(n) = 1/(pi a^2), G = 1, F = ((eta-1)/(eta+1))^2
    nrm = np.tile(np.float32([0, 0, 1]), (1, 1))
    out = orc.bsdf(mat, frames_up(1), nrm, np.full((1, 2), 0.5, np.float32), nrm, nrm)
    expect = ks * 0.04 / (4 * np.pi * a * a)
    assert np.allclose(out[0, 8:11], expect, rtol=1e-5)
    # reciprocity
    n = 2000
    ea, eb = upper(rng, n), upper(rng, n)
    o1 = orc.bsdf(mat, frames_up(n), ea, rng.random((n, 2)), ea, eb)
    o2 = orc.bsdf(mat, frames_up(n), ea, rng.random((n, 2)), eb, ea)
    assert np.allclose(o1[:, 8:11], o2[:, 8:11], rtol=1e-4, atol=1e-7)
    # sample/pdf consistency at a fixed view: the pdf's integral over the upper hemisphere
    # (quadrature) equals the fraction of samples landing there, and chi-square in cos bins
    view = np.float32([[0.5, 0.1, math.sqrt(1 - 0.26)]])
    m = 60000
    s = orc.bsdf(mat, frames_up(m), np.repeat(view, m, 0), rng.random((m, 2)), np.repeat(view, m, 0),
                 np.repeat(view, m, 0))
    wo = s[:, 4:7].astype(np.float64)
    up = wo[:, 2] > 0
    nt, npp = 400, 800
    th = (np.arange(nt) + 0.5) * (np.pi / 2) / nt
    ph = (np.arange(npp) + 0.5) * 2 * np.pi / npp
    T, P = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3).astype(np.float32)
    dA = (np.sin(T) * (np.pi / 2 / nt) * (2 * np.pi / npp)).ravel()
    k = dirs.shape[0]
    q = orc.bsdf(mat, frames_up(k), np.repeat(view, k, 0), np.zeros((k, 2), np.float32), np.repeat(view, k, 0),
                 dirs)[:, 11].astype(np.float64)
    mass = np.sum(q * dA)
    frac = up.mean()
    assert abs(mass - frac) < 4 * math.sqrt(frac * (1 - frac) / m) + 2e-3, (mass, frac)
    bins = np.minimum((wo[up, 2] * 10).astype(int), 9)
    counts = np.bincount(bins, minlength=10)
    qb = np.bincount(np.minimum((dirs[:, 2] * 10).astype(int), 9), weights=q * dA, minlength=10) * m
    sel = qb > 20
    chi2 = np.sum((counts[sel] - qb[sel]) ** 2 / qb[sel])
    assert chi2 < 40


def test_area_light_closed_forms(orc):
    B = _Builder()
    B.bundle([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], [[0, 1, 0]] * 4, [[0, 1, 2], [0, 2, 3]],
             B.material(LAMBERT, (0.5, 0.5, 0.5)))
    tri = np.array([[0, 2, 0], [1, 2, 0], [0, 2, 1]], np.float32)
    lb = B.bundle(tri, [[0, -1, 0]] * 3, [[0, 1, 2]], B.material(EMISSIVE, (3, 3, 3)))
    B.area_lights(lb)
    sc = B.build([0, 1, 5, 0, 0, 0, 0, 1, 0, 0.01, 1e4, 0.8, 1.0], "light")
    s = orc.scene(sc)
    rng = np.random.default_rng(3)
    n = 20000
    p = np.tile(np.float32([0.2, 0.0, 0.3]), (n, 1))
    u = rng.random((n, 2)).astype(np.float32)
    w = np.tile(np.float32([0, 1, 0]), (n, 1))
    out = s.light(0, p, u, w)
    A = 0.5
    assert np.allclose(out[:, 3], 1 / A, rtol=1e-6)                  # area pdf
    y = p + out[:, 4:7]                                              # sampled points
    assert np.allclose(y[:, 1], 2.0, atol=1e-5)
    assert np.all((y[:, 0] >= -1e-5) & (y[:, 2] >= -1e-5) & (y[:, 0] + y[:, 2] <= 1 + 1e-5))
    assert np.allclose(y.mean(0), [1 / 3, 2, 1 / 3], atol=0.01)      # uniform: mean = centroid
    d = out[:, 4:7].astype(np.float64)
    r2 = np.sum(d * d, 1)
    cosl = d[:, 1] / np.sqrt(r2)
    assert np.allclose(out[:, 0], 3 * cosl / r2, rtol=1e-4)          # Le(cos_l) / d^2
    # solid-angle pdf toward the centroid: d^2 / (cos_l A)
    c = np.float32([1 / 3, 2, 1 / 3])
    wc = (c - p[0]) / np.linalg.norm(c - p[0])
    g = s.light(0, p[:1], u[:1], wc[None].astype(np.float32))[0, 7]
    dc = np.linalg.norm(c - p[0])
    assert abs(g - dc * dc / ((2 / dc) * A)) < 1e-4 * g


def test_glibc_rand_and_mwc(orc):
    # srand(1): the first rand() values of glibc's TYPE_3 generator are public
    r = orc.mcrng(1, 6).astype(np.uint64)
    z, w = 1804289383, 846930886
    expect = []
    for _ in range(6):
        z = 36969 * (z & 0xFFFF) + (z >> 16)
        w = 18000 * (w & 0xFFFF) + (w >> 16)
        expect.append(((z << 16) + w) & 0xFFFFFFFF)
        z &= 0xFFFFFFFF
        w &= 0xFFFFFFFF
    assert list(r) == expect


def test_stratified_cells(orc):
    g = 4
    s = orc.stratified(g, 7, 64).astype(np.float64)
    idx = np.arange(64) % (g * g)
    assert np.all(np.floor(s[:, 0] * g) == idx % g)
    assert np.all(np.floor(s[:, 1] * g) == idx // g)


def test_camera_frustum(orc):
    cam = np.float32([1, 2, 3, 0, 0, 0, 0, 1, 0, 0.01, 1e4, 0.7, 1.5])
    r = orc.camera_rays(cam, np.float32([[0.5, 0.5], [0.5, 0.0], [0.5, 1.0], [0.0, 0.5], [1.0, 0.5]]))
    f = -cam[0:3] / np.linalg.norm(cam[0:3])
    assert np.allclose(r[0, 4:7], f, atol=1e-6)
    assert np.allclose(r[:, 0:3], cam[0:3]) and np.allclose(r[:, 3], 0.01) and np.allclose(r[:, 7], 1e4)
    ang = lambda a, b: math.acos(np.clip(np.dot(a, b), -1, 1))
    # dim.x = 2 tan(fovy/2) znear, dim.y = dim.x / aspect; right_ = cross(forward_, up) is not
    # normalised (perspective_camera.cpp:14-15), so both extents scale by |forward x up|
    k = np.linalg.norm(np.cross(f, [0, 1, 0]))
    assert abs(ang(r[1, 4:7], r[2, 4:7]) - 2 * math.atan(k * math.tan(0.35) / 1.5)) < 1e-5
    assert abs(ang(r[3, 4:7], r[4, 4:7]) - 2 * math.atan(k * math.tan(0.35))) < 1e-5
    r2 = orc.camera_rays(np.float32([0, 0, 5, 0, 0, 0, 0, 1, 0, 0.01, 1e4, 0.7, 1.0]), np.float32([[0, 0.5], [1, 0.5]]))
    assert abs(ang(r2[0, 4:7], r2[1, 4:7]) - 0.7) < 1e-5      # level camera: the nominal fovy


def test_sh_closed_forms(orc):
    rng = np.random.default_rng(4)
    d = rng.normal(size=(500, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    Y = orc.sh_eval(d, 2).astype(np.float64)
    x, y, z = d.T
    c1, c2, c3, c4 = 0.48860251, 1.09254843, 0.31539157, 0.54627422
    ref = np.stack([np.full_like(x, 0.28209479), -c1 * y, c1 * z, -c1 * x, c2 * x * y, -c2 * y * z,
                    c3 * (3 * z * z - 1), -c2 * x * z, c4 * (x * x - y * y)], 1)
    assert np.allclose(Y, ref, atol=2e-6)
    # orthonormality up to l = 4 on a fine grid
    nt, npp = 200, 400
    th = (np.arange(nt) + 0.5) * np.pi / nt
    ph = (np.arange(npp) + 0.5) * 2 * np.pi / npp
    T, P = np.meshgrid(th, ph, indexing="ij")
    g = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dA = (np.sin(T) * (np.pi / nt) * (2 * np.pi / npp)).ravel()
    Yg = orc.sh_eval(g, 4).astype(np.float64)
    G = (Yg * dA[:, None]).T @ Yg
    assert np.allclose(G, np.eye(25), atol=2e-4)
    # cosine-lobe convolution: cout_l / cin_l = A_l
    cin = np.ones((25, 3), np.float32)
    cout = orc.sh_convolve(4, cin)[:, 0]
    A = [np.pi, 2 * np.pi / 3, np.pi / 4, 0.0, -np.pi / 24]
    for l in range(5):
        for m in range(-l, l + 1):
            assert abs(cout[l * l + l + m] - A[l]) < 1e-5, (l, m, cout[l * l + l + m], A[l])
    # a constant map of radiance 1 projects to c00 = sqrt(4 pi) and gives irradiance 0.7 (the
    # ShTracer albedo 0.7 times E / pi = 1)
    env = np.ones((64, 128, 3), np.float32)
    c = orc.sh_project(env, 2)
    assert abs(c[0, 0] - math.sqrt(4 * math.pi)) < 2e-3
    assert np.all(np.abs(c[1:]) < 2e-3)
    e = orc.sh_irradiance(2, c, d[:10].astype(np.float32))
    assert np.allclose(e, 0.7, atol=2e-3)
