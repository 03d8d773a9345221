"""The C ABI's error paths: every check below returns its status code and message before any
kernel runs, and leaves nothing half-built (the reference reports the same misuse by throwing
std::runtime_error, e.g. an area light on a non-emissive shape, light/arealight.cpp:16-19; a
NaN vertex in the reference silently yields misses)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INVALID, UNSUPPORTED, STATE = 1, 3, 5


def _raw_scene(ctx, sc):
    """bsh_scene_create + materials + lights + camera, without commit."""
    from banshee_amd import hip
    from banshee_amd.hip import _p, lib
    h = ctypes.c_void_p()
    a = [np.ascontiguousarray(x, t) for x, t in ((sc.verts, np.float32), (sc.normals, np.float32),
                                                 (sc.tri_v, np.int32), (sc.tri_n, np.int32), (sc.tri_m, np.int32),
                                                 (sc.bundles, np.int32))]
    code = lib().bsh_scene_create(ctx.h, _p(a[0]), ctypes.c_size_t(a[0].shape[0]), _p(a[1]),
                                  ctypes.c_size_t(a[1].shape[0]), _p(a[2]), _p(a[3]), _p(a[4]),
                                  ctypes.c_size_t(a[4].shape[0]), _p(a[5]), ctypes.c_size_t(a[5].shape[0]),
                                  ctypes.byref(h))
    return code, h, a


def _status(code):
    from banshee_amd.hip import lib
    return code, lib().bsh_last_error().decode()


def test_scene_creation_checks(ctx, cornell):
    from banshee_amd import hip
    bad = cornell.with_camera(cornell.camera)
    bad.verts = cornell.verts.copy()
    bad.verts[3, 1] = np.nan
    with pytest.raises(hip.BshError, match="non-finite vertex"):
        hip.Scene(ctx, bad)
    bad = cornell.with_camera(cornell.camera)
    bad.tri_v = cornell.tri_v.copy()
    bad.tri_v[5, 2] = cornell.verts.shape[0]
    with pytest.raises(hip.BshError, match="vertex index out of range at triangle 5"):
        hip.Scene(ctx, bad)
    bad = cornell.with_camera(cornell.camera)
    bad.tri_n = cornell.tri_n.copy()
    bad.tri_n[7, 0] = -1
    with pytest.raises(hip.BshError, match="normal index out of range at triangle 7"):
        hip.Scene(ctx, bad)
    bad = cornell.with_camera(cornell.camera)
    bad.bundles = cornell.bundles.copy()
    bad.bundles[1, 1] += 1          # bundle 1 claims one triangle too many
    with pytest.raises(hip.BshError, match="bundles"):
        hip.Scene(ctx, bad)


def test_material_and_light_checks(ctx, cornell):
    from banshee_amd import hip
    bad = cornell.with_camera(cornell.camera)
    bad.tri_m = cornell.tri_m.copy()
    bad.tri_m[0] = cornell.materials.shape[0]
    with pytest.raises(hip.BshError, match="material index out of range at triangle 0"):
        hip.Scene(ctx, bad)
    bad = cornell.with_camera(cornell.camera)
    bad.materials = cornell.materials.copy()
    bad.materials[0, 0] = 9
    with pytest.raises(hip.BshError, match="unknown material type"):
        hip.Scene(ctx, bad)
    # an area light on a non-emissive triangle: AreaLight's own message (arealight.cpp:16-19)
    bad = cornell.with_camera(cornell.camera)
    bad.lights = np.array([[0, 0]], np.int32)       # bundle 0 is a Lambert wall
    with pytest.raises(hip.BshError, match="The material passed is not an emissive one"):
        hip.Scene(ctx, bad)
    bad = cornell.with_camera(cornell.camera)
    bad.lights = np.array([[0, 99]], np.int32)
    with pytest.raises(hip.BshError, match="out of range"):
        hip.Scene(ctx, bad)
    # an environment light without a map, and an unknown light type
    from banshee_amd.scenes import ENV_LIGHT
    bad = cornell.with_camera(cornell.camera)
    bad.other_lights = np.array([[ENV_LIGHT, 0, 0, 0, 0, 0, 0, 1, 2.2]], np.float32)
    with pytest.raises(hip.BshError, match="need bsh_scene_set_env_map"):
        hip.Scene(ctx, bad)
    bad.other_lights = np.array([[9, 0, 0, 0, 1, 1, 1, 1, 2.2]], np.float32)
    with pytest.raises(hip.BshError, match="unknown type"):
        hip.Scene(ctx, bad)
    bad.other_lights = None
    bad.env = np.full((4, 8, 3), np.inf, np.float32)
    with pytest.raises(hip.BshError, match="non-finite texel"):
        hip.Scene(ctx, bad)


def test_call_order_and_frame_checks(ctx, cornell):
    from banshee_amd import hip
    from banshee_amd.hip import _p, lib
    code, h, keep = _raw_scene(ctx, cornell)
    assert code == 0
    try:
        acc = ctx.alloc((4, 4, 8), np.float32)
        p = hip.render_params(4, 4, 1, 2, 1)
        code, msg = _status(lib().bsh_render(h, ctypes.byref(p), acc.ptr, None, None))
        assert code == STATE and "not committed" in msg
        rays = ctx.to_device(np.zeros((4, 8), np.float32))
        hits = ctx.alloc((4, 4), np.float32)
        code, msg = _status(lib().bsh_intersect_closest(h, rays.ptr, ctypes.c_size_t(4), hits.ptr, None))
        assert code == STATE and "not committed" in msg
    finally:
        lib().bsh_scene_destroy(h)
    s = hip.Scene(ctx, cornell)
    for kw, msg in ((dict(width=0), "bad frame parameters"), (dict(crop=(60, 60, 8, 8)), "crop outside the image"),
                    (dict(spp_range=(3, 5)), "bad sample range"), (dict(chunk_spp=4), "chunk_spp is reserved")):
        args = dict(width=64, height=64, img_grid=2, max_depth=2, seed=1)
        args.update(kw)
        with pytest.raises(hip.BshError, match=msg):
            s.render(args.pop("width"), args.pop("height"), args.pop("img_grid"), args.pop("max_depth"),
                     args.pop("seed"), **args)
    probes = np.array([[1.0, 1.0, 1.0]], np.float32)
    with pytest.raises(hip.BshError, match="rays_per_probe must be a square"):
        s.bake_sh(probes, 30, 2, 3, 1)
    with pytest.raises(hip.BshError, match="lmax must be 0..4"):
        s.bake_sh(probes, 16, 5, 3, 1)
    with pytest.raises(hip.BshError, match="chunk_probes is reserved"):
        s.bake_sh(probes, 16, 2, 3, 1, chunk_probes=8)
    with pytest.raises(hip.BshError, match="bad max_depth"):
        s.bake_sh(probes, 16, 2, 0, 1)
