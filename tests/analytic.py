"""Closed forms and brute-force quadratures used as independent checks (test helpers)."""
import numpy as np


def triangle_points(tri, m):
    """Centroids and areas of the m^2 congruent sub-triangles of `tri` (3 x 3)."""
    p1, p2, p3 = (np.asarray(v, np.float64) for v in tri)
    e1, e2 = (p2 - p1) / m, (p3 - p1) / m
    i, j = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    i, j = i.ravel(), j.ravel()
    up = i + j < m                        # upward sub-triangles: (i, j), (i+1, j), (i, j+1)
    c_up = p1 + np.outer(i[up] + 1 / 3, e1) + np.outer(j[up] + 1 / 3, e2)
    dn = i + j < m - 1                    # downward: (i+1, j), (i+1, j+1), (i, j+1)
    c_dn = p1 + np.outer(i[dn] + 2 / 3, e1) + np.outer(j[dn] + 2 / 3, e2)
    area = 0.5 * np.linalg.norm(np.cross(p2 - p1, p3 - p1)) / (m * m)
    return np.concatenate([c_up, c_dn]), area


def segment_blocked(p, q, occluders):
    """For each point of q (M x 3): does the open segment p -> q cross any triangle of
    `occluders` (K x 3 x 3)? Moller-Trumbore in FP64, both faces blocking."""
    p = np.asarray(p, np.float64)
    d = np.asarray(q, np.float64) - p
    out = np.zeros(len(d), bool)
    for tri in np.asarray(occluders, np.float64):
        p1, p2, p3 = tri
        e1, e2 = p2 - p1, p3 - p1
        s1 = np.cross(d, e2)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1.0 / (s1 @ e1)
            s = p - p1
            b1 = (s1 @ s) * inv
            s2 = np.cross(s, e1)
            b2 = (d @ s2) * inv
            t = (s2 @ e2) * inv
        out |= (b1 >= 0) & (b2 >= 0) & (b1 + b2 <= 1) & (t > 1e-9) & (t < 1 - 1e-9)
    return out


def emitter_irradiance(p, n, tri, light_normal, cos_power, m=600, occluders=None):
    """E(p) = integral over the emitter of cos_l^cos_power cos_p / r^2 dA (unit emission),
    for a one-sided emitter facing `light_normal`; midpoint rule on m^2 sub-triangles. With
    `occluders`, sub-triangles whose centroid is hidden from p by one of them count zero."""
    c, dA = triangle_points(tri, m)
    d = c - np.asarray(p, np.float64)
    r2 = np.sum(d * d, axis=1)
    w = d / np.sqrt(r2)[:, None]
    cos_p = np.maximum(w @ np.asarray(n, np.float64), 0)
    cos_l = np.maximum(-(w @ np.asarray(light_normal, np.float64)), 0)
    f = cos_l ** cos_power * cos_p / r2
    if occluders is not None and len(occluders):
        f = np.where(segment_blocked(p, c, occluders), 0.0, f)
    return float(np.sum(f) * dA)


def lambert_polygon_irradiance(p, n, poly):
    """Lambert's closed form for a uniform unit-radiance polygon seen from p with normal n:
    E = 1/2 sum_i theta_i (n . Gamma_i), Gamma_i the unit normal of the plane through p,
    v_i and v_{i+1}, theta_i the angle between them."""
    v = np.asarray(poly, np.float64) - np.asarray(p, np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    s = 0.0
    for i in range(len(v)):
        a, b = v[i], v[(i + 1) % len(v)]
        g = np.cross(a, b)
        s += np.arccos(np.clip(a @ b, -1, 1)) * (g / np.linalg.norm(g)) @ np.asarray(n, np.float64)
    return abs(0.5 * s)


def camera_footprint(cam13, k=5):
    """Points where a k x k grid of rays across a (narrow) camera's image meets y = 0."""
    eye, at, up = (np.asarray(cam13[i:i + 3], np.float64) for i in (0, 3, 6))
    znear, fovy, aspect = cam13[9], cam13[11], cam13[12]
    f = (at - eye) / np.linalg.norm(at - eye)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    wx = 2 * np.tan(fovy / 2) * znear
    wy = wx / aspect
    out = []
    for sx in (np.arange(k) + 0.5) / k - 0.5:
        for sy in (np.arange(k) + 0.5) / k - 0.5:
            d = znear * f + sx * wx * r + sy * wy * u
            t = -eye[1] / d[1]
            out.append(eye + t * d)
    return np.asarray(out)
