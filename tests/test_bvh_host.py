"""The host BVH builder and the flattened node layouts (CPU only, through the C ABI).

Invariants: every primitive is in exactly one leaf; every child box contains its subtree's
triangles (boxes are padded outward); inner references are in range; the depth respects the
traversal stack bound. A numpy traversal of the flattened nodes (visiting every child whose
box the ray enters) finds the same closest hits as FP64 brute force, so a wrong box, ref,
count or collapse step in any layout fails here before it reaches a GPU. The quantised 4-wide
layout ("4q") must hold the float 4-wide tree node for node, with every child box containing
the float box it replaces by at least half a grid step (outward rounding), on grids no finer
than 2^-17 of the scene's largest coordinate.
"""
import numpy as np
import pytest

from banshee_amd import hip
from banshee_amd.scenes import terrain_for_tris

pytestmark = pytest.mark.skipif(not __import__("os").path.exists(hip.LIB_PATH), reason="library not built")


def tri_verts(sc):
    return sc.verts[sc.tri_v].reshape(-1, 9).astype(np.float32)


def decode_q(nodes):
    """GpuNodeQ rows (16 words): lo [4,3], hi [4,3], ref [4], cnt [4], origin [3], step [3], in FP64."""
    u = nodes.view(np.uint32)
    origin = nodes[:, 0:3].astype(np.float64)
    step = np.stack([2.0 ** ((u[:, 3] >> (8 * a)) & 255).astype(np.uint8).view(np.int8).astype(np.int64)
                     for a in range(3)], -1)
    def byte(w, c):
        return ((w >> (8 * c)) & 255).astype(np.float64)
    lo = np.stack([np.stack([origin[:, a] + byte(u[:, 4 + 2 * a], c) * step[:, a] for a in range(3)], -1)
                   for c in range(4)], 1)
    hi = np.stack([np.stack([origin[:, a] + byte(u[:, 5 + 2 * a], c) * step[:, a] for a in range(3)], -1)
                   for c in range(4)], 1)
    cnt = np.stack([((u[:, 10] >> (8 * c)) & 255).astype(np.uint8).view(np.int8).astype(np.int32) for c in range(4)], -1)
    return lo, hi, nodes.view(np.int32)[:, 12:16], cnt, origin, step


def build(tv, layout):
    """(nodes, order, w, depth) for layout 2, 4 or "4q"."""
    if layout == "4q":
        return hip.bvh_build_host(tv, width=4, quantized=True)
    return hip.bvh_build_host(tv, width=layout, quantized=False)


def decode(nodes, width):
    """Per node: lo [k,3], hi [k,3], ref [k], cnt [k] for k = width children."""
    if width == "4q":
        return decode_q(nodes)[:4]
    if width == 2:
        f = nodes
        i = nodes.view(np.int32)
        lo = np.stack([np.stack([f[:, 0], f[:, 2], f[:, 8]], -1), np.stack([f[:, 4], f[:, 6], f[:, 10]], -1)], 1)
        hi = np.stack([np.stack([f[:, 1], f[:, 3], f[:, 9]], -1), np.stack([f[:, 5], f[:, 7], f[:, 11]], -1)], 1)
        return lo, hi, i[:, 12:14], i[:, 14:16]
    f = nodes.reshape(-1, 8, 4)
    i = nodes.view(np.int32).reshape(-1, 8, 4)
    lo = np.stack([f[:, 0], f[:, 2], f[:, 4]], -1)
    hi = np.stack([f[:, 1], f[:, 3], f[:, 5]], -1)
    return lo, hi, i[:, 6], i[:, 7]


@pytest.fixture(scope="module")
def scene():
    return terrain_for_tris(4000, seed=9, width=64, height=64)


@pytest.mark.parametrize("width", [2, 4, "4q"])
def test_layout_invariants(scene, width):
    tv = tri_verts(scene)
    nodes, order, w, depth = build(tv, width)
    assert w == (4 if width == "4q" else width)
    assert sorted(order.tolist()) == list(range(len(tv)))
    lo, hi, ref, cnt = decode(nodes, width)
    n = nodes.shape[0]
    seen = np.zeros(len(tv), np.int32)
    tris = tv.reshape(-1, 3, 3)
    for k in range(n):
        for c in range(w):
            if cnt[k, c] > 0:
                prims = order[ref[k, c]:ref[k, c] + cnt[k, c]]
                seen[prims] += 1
                p = tris[prims].reshape(-1, 3)
                assert np.all(p >= lo[k, c]) and np.all(p <= hi[k, c])
            elif cnt[k, c] == 0:
                assert 0 < ref[k, c] < n
                child = ref[k, c]
                occupied = cnt[child] >= 0
                # quantised boxes are not nested (each node rounds on its own grid); that they
                # contain the float boxes, which are, is test_quantized_boxes_contain_float_boxes
                if width != "4q":
                    assert np.all(lo[child][occupied] >= lo[k, c]) and np.all(hi[child][occupied] <= hi[k, c])
    assert np.all(seen == 1)
    bound = 32 if w == 2 else 48
    assert (depth + 2 if w == 2 else 3 * depth + 1) <= bound


def traverse_np(nodes, layout, order, tris, o, d, tmin, tmax):
    lo, hi, ref, cnt = decode(nodes, layout)
    w = lo.shape[1]
    inv = 1.0 / np.where(np.abs(d) < 1e-20, np.copysign(1e-20, d), d)
    best, bt = -1, tmax
    stack = [0]
    while stack:
        k = stack.pop()
        for c in range(w):
            if cnt[k, c] < 0:
                continue
            t0 = (lo[k, c] - o) * inv
            t1 = (hi[k, c] - o) * inv
            tn = max(np.max(np.minimum(t0, t1)), tmin)
            tf = min(np.min(np.maximum(t0, t1)), bt)
            if tn > tf:
                continue
            if cnt[k, c] == 0:
                stack.append(ref[k, c])
                continue
            for q in range(ref[k, c], ref[k, c] + cnt[k, c]):
                p1, p2, p3 = tris[order[q]].astype(np.float64)
                e1, e2 = p2 - p1, p3 - p1
                s1 = np.cross(d, e2)
                det = s1 @ e1
                if det == 0:
                    continue
                dd = o - p1
                b1 = (dd @ s1) / det
                s2 = np.cross(dd, e1)
                b2 = (d @ s2) / det
                t = (e2 @ s2) / det
                if b1 >= 0 and b1 <= 1 and b2 >= 0 and b1 + b2 <= 1 and tmin < t <= bt:
                    best, bt = order[q], t
    return best, bt


@pytest.mark.parametrize("width", [2, 4, "4q"])
def test_layout_traversal_matches_brute_force(scene, width):
    tv = tri_verts(scene)
    nodes, order, w, _ = build(tv, width)
    tris = tv.reshape(-1, 3, 3)
    rng = np.random.default_rng(7 if width == "4q" else width)
    lo, hi = scene.verts.min(0), scene.verts.max(0)
    from test_oracle_pins import brute_force
    n = 150
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.full((n, 1), 0.01), d, np.full((n, 1), 1e4)], 1).astype(np.float32)
    bp, bt = brute_force(scene, rays)
    r64 = rays.astype(np.float64)
    hits = 0
    for i in range(n):
        p, t = traverse_np(nodes, width, order, tris, r64[i, 0:3], r64[i, 4:7], r64[i, 3], r64[i, 7])
        assert (p >= 0) == (bp[i] >= 0), i
        if p >= 0:
            hits += 1
            assert abs(t - bt[i]) <= 1e-9 * max(1.0, bt[i]), (i, t, bt[i])
    assert hits > n // 4


def test_four_wide_collapses_levels(scene):
    tv = tri_verts(scene)
    n2, _, _, d2 = hip.bvh_build_host(tv, width=2)
    n4, _, w, d4 = hip.bvh_build_host(tv, width=4)
    assert w == 4
    assert len(n4) < 0.6 * len(n2)
    assert d4 < d2


@pytest.mark.parametrize("flat,offset", [(False, 0.0), (False, 3.0e4), (True, 0.0)])
def test_quantized_boxes_contain_float_boxes(scene, flat, offset):
    """The quantised tree is the float 4-wide tree with each child box grown outward: same refs
    and counts, every plane at least half a grid step outside the float plane (or, for a lower
    plane, on the node's origin, which is the lowest child's own float plane), and grid steps no
    finer than 2^-17 of the scene's largest coordinate (the offset scene puts the terrain far
    from the origin, where that floor decides the grid of the small nodes). The flat case lays
    the terrain on a plane (its thinnest axis set to 1.5): along that axis every float box is
    then only its builder padding thick (a relative 2^-20 each way), far below the floor, which
    then sets the grid there."""
    tv = tri_verts(scene) + np.float32(offset)
    if flat:
        v = tv.reshape(-1, 3)
        v[:, np.argmin(np.ptp(v, axis=0))] = 1.5
    n4, o4, w4, _ = hip.bvh_build_host(tv, width=4, quantized=False)
    nq, oq, wq, _ = hip.bvh_build_host(tv, width=4, quantized=True)
    assert w4 == wq == 4 and nq.shape == (n4.shape[0], 16)
    assert np.array_equal(o4, oq)
    lo, hi, ref, cnt = decode(n4, 4)
    qlo, qhi, qref, qcnt, origin, step = decode_q(nq)
    assert np.array_equal(ref, qref) and np.array_equal(cnt, qcnt)
    live = cnt >= 0
    lo64, hi64 = lo.astype(np.float64), hi.astype(np.float64)
    half = 0.5 * step[:, None, :]
    on_origin = qlo == origin[:, None, :]
    assert np.all(((lo64 - qlo >= half) | (on_origin & (qlo <= lo64)))[live])
    assert np.all((qhi - hi64 >= half)[live])
    # the origin is the lowest live child's float plane; 254 steps cover the node
    olo = np.where(live[..., None], lo64, np.inf).min(1)
    ohi = np.where(live[..., None], hi64, -np.inf).max(1)
    assert np.array_equal(origin, olo)
    assert np.all(ohi - olo <= 254 * step)
    mag = np.abs(np.concatenate([lo64[0][live[0]], hi64[0][live[0]]])).max()
    assert np.all(step >= mag * 2.0 ** -17)
    # the grid is the finest the two rules allow: half the step would break one of them
    assert np.all((ohi - olo > 127 * step) | (step / 2 < mag * 2.0 ** -17))


def test_default_layout_is_quantized(scene):
    """quantized=None (the default) builds the quantised 4-wide nodes."""
    tv = tri_verts(scene)
    na, _, wa, _ = hip.bvh_build_host(tv, width=4)
    nq, _, wq, _ = hip.bvh_build_host(tv, width=4, quantized=True)
    assert wa == wq == 4 and na.shape == nq.shape and np.array_equal(na.view(np.int32), nq.view(np.int32))


def test_quantized_layout_falls_back_outside_its_exponent_range(scene):
    """Coordinates near 1e30 need grid steps past 2^60, where 2^e / d could leave the float
    range: the builder keeps float nodes (128 B) for them. Coordinates near 1e-18 (grid steps
    near 2^-77) still quantise, with boxes that contain every float box. (Below about 1e-24
    the SAH areas underflow and the tree stays binary, so the grid's 2^-100 floor is not
    reached through a 4-wide build.)"""
    tv = tri_verts(scene)
    big = (tv.astype(np.float64) * 1e28).astype(np.float32)
    nodes, _, w, _ = hip.bvh_build_host(big, width=4)
    assert w == 4 and nodes.shape[1] == 32
    tiny = (tv.astype(np.float64) * 1e-20).astype(np.float32)
    nq, _, wq, _ = hip.bvh_build_host(tiny, width=4)
    nf, _, _, _ = hip.bvh_build_host(tiny, width=4, quantized=False)
    assert wq == 4 and nq.shape[1] == 16
    lo, hi, _, cnt = decode(nf, 4)
    qlo, qhi, _, qcnt, _, _ = decode_q(nq)
    live = cnt >= 0
    assert np.array_equal(cnt, qcnt)
    assert np.all((qlo <= lo.astype(np.float64))[live]) and np.all((qhi >= hi.astype(np.float64))[live])


def test_invalid_quantized_value_is_rejected(scene):
    """bsh_bvh_params.quantized takes -1, 0 or 1; anything else is BSH_ERR_INVALID with a message."""
    import ctypes
    tv = np.ascontiguousarray(tri_verts(scene))
    n, w, d, nb = ctypes.c_size_t(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    for bad in (2, -2):
        bp = hip.BvhParams(4, 16, 0, 4, bad)
        st = hip.lib().bsh_bvh_build_host(tv.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(tv.shape[0]),
                                          ctypes.byref(bp), None, ctypes.c_size_t(0), None, ctypes.byref(n),
                                          ctypes.byref(w), ctypes.byref(d), ctypes.byref(nb))
        assert st != 0
        assert b"quantized must be -1, 0 or 1" in hip.lib().bsh_last_error()
