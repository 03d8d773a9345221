"""FP64 adjudication of ray/triangle disagreements (test infrastructure).

A disagreement between two tracers about which triangle a ray hits first is
acceptable only when double precision shows it is an edge, vertex, grazing or
equal-distance case: the reference itself resolves those by floating-point
rounding and test order (Mesh::IntersectFace accepts t <= tmax, mesh.cpp:33).
"""
import numpy as np

EDGE_EPS = 2e-5     # barycentric distance to an edge, relative
TIE_EPS = 2e-5      # relative t difference counted as a tie
GRAZE_EPS = 2e-4    # |cos| between ray and triangle normal counted as grazing


def mt64(o, d, p1, p2, p3):
    """Double-precision Moller-Trumbore (same formulas as mesh.cpp:7-45, no rejection)."""
    o, d, p1, p2, p3 = (np.asarray(x, np.float64) for x in (o, d, p1, p2, p3))
    e1 = p2 - p1
    e2 = p3 - p1
    s1 = np.cross(d, e2)
    det = np.dot(s1, e1)
    if det == 0.0:
        return np.inf, np.nan, np.nan, 0.0
    inv = 1.0 / det
    dd = o - p1
    b1 = np.dot(dd, s1) * inv
    s2 = np.cross(dd, e1)
    b2 = np.dot(d, s2) * inv
    t = np.dot(e2, s2) * inv
    n = np.cross(e1, e2)
    cosang = abs(np.dot(n, d)) / (np.linalg.norm(n) * np.linalg.norm(d) + 1e-300)
    return t, b1, b2, cosang


def tri(scene, prim):
    v = scene.verts
    i = scene.tri_v[prim]
    return v[i[0]], v[i[1]], v[i[2]]


def near_edge(b1, b2):
    b0 = 1.0 - b1 - b2
    return min(abs(b0), abs(b1), abs(b2)) <= EDGE_EPS


def classify(scene, ray, prim_a, prim_b):
    """Returns 'tie', 'edge', 'grazing' or 'fail' for a closest-hit disagreement."""
    o, d = ray[0:3], ray[4:7]
    tmin, tmax = ray[3], ray[7]
    ta = tb = None
    for p, key in ((prim_a, "a"), (prim_b, "b")):
        if p < 0:
            continue
        t, b1, b2, c = mt64(o, d, *tri(scene, p))
        if near_edge(b1, b2):
            return "edge"
        if c <= GRAZE_EPS:
            return "grazing"
        if abs(t - tmin) <= TIE_EPS * max(1.0, abs(tmin)) or abs(t - tmax) <= TIE_EPS * max(1.0, abs(tmax)):
            return "edge"
        if key == "a":
            ta = t
        else:
            tb = t
    if ta is not None and tb is not None and abs(ta - tb) <= TIE_EPS * max(abs(ta), abs(tb), 1e-30):
        return "tie"
    return "fail"
