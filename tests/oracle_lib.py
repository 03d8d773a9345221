"""ctypes bindings for the two CPU oracles (test infrastructure only).

* ``oracle/_ref/liboracle_ref.so`` -- the reference's own sources compiled in place
  plus ``oracle/harness`` (prefix ``oref_``). Built only where the reference is present.
* ``oracle/_build/libcpuref.so``  -- this repository's CPU restatement
  (``oracle/cpu_ref``, prefix ``cref_``), pinned against the fixtures the first one wrote.

Both expose the same entry points, so ``Oracle("ref")`` and ``Oracle("cpu")`` are
interchangeable in tests.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "liboracle_ref.so")
CPU_SO = os.path.join(ROOT, "oracle", "_build", "libcpuref.so")

_libs = {}


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def available(kind: str) -> bool:
    return os.path.exists(REF_SO if kind == "ref" else CPU_SO)


def _load(kind):
    if kind not in _libs:
        path = REF_SO if kind == "ref" else CPU_SO
        lib = ctypes.CDLL(path)
        pre = "oref_" if kind == "ref" else "cref_"
        vp = ctypes.c_void_p
        i = ctypes.c_int
        getattr(lib, pre + "scene_create").restype = vp
        getattr(lib, pre + "scene_create").argtypes = [vp, i, vp, i, vp, vp, vp, i, vp, i, vp, i, vp, i, vp]
        getattr(lib, pre + "scene_destroy").argtypes = [vp]
        getattr(lib, pre + "intersect").argtypes = [vp, vp, i, i, vp, vp]
        getattr(lib, pre + "render_keyed").argtypes = [vp, i, i, i, i, i, i, i, i, i, ctypes.c_uint, vp, vp, vp,
                                                          vp, vp]
        getattr(lib, pre + "time_render").restype = ctypes.c_double
        getattr(lib, pre + "time_render").argtypes = [vp, i, i, i, i, i, ctypes.c_uint, i, vp]
        getattr(lib, pre + "bsdf").argtypes = [vp, i, vp, vp, vp, vp, vp, vp]
        getattr(lib, pre + "light").argtypes = [vp, i, i, vp, vp, vp, vp]
        getattr(lib, pre + "camera_rays").argtypes = [vp, vp, i, vp]
        getattr(lib, pre + "mcrng").argtypes = [ctypes.c_uint, i, vp]
        getattr(lib, pre + "stratified").argtypes = [i, ctypes.c_uint, i, vp]
        getattr(lib, pre + "sh_eval").argtypes = [vp, i, i, vp]
        getattr(lib, pre + "sh_convolve").argtypes = [i, vp, vp]
        getattr(lib, pre + "sh_project").argtypes = [vp, i, i, i, vp]
        getattr(lib, pre + "sh_irradiance").argtypes = [i, vp, vp, i, vp]
        if kind == "cpu":
            lib.cref_bake_sh.argtypes = [vp, vp, i, i, i, i, ctypes.c_uint, ctypes.c_float, vp, vp, vp, i]
            lib.cref_render_bsdf_only.argtypes = [vp, i, i, i, i, i, i, i, i, ctypes.c_uint, vp, vp, vp, vp, i]
        if kind == "ref":
            lib.oref_render_keyed_frame.argtypes = [vp, i, i, i, i, i, i, i, ctypes.c_uint, vp, vp, vp, vp, vp]
            lib.oref_scene_create_textured.restype = vp
            lib.oref_scene_create_textured.argtypes = [vp, i, vp, i, vp, vp, vp, i, vp, i, vp, i, vp, i, vp,
                                                       vp, i, vp, vp, vp, i, vp]
            lib.oref_scene_create_normal_mapped.restype = vp
            lib.oref_scene_create_normal_mapped.argtypes = lib.oref_scene_create_textured.argtypes + [vp]
            lib.oref_scene_set_env.argtypes = [vp, vp, i, i]
            lib.oref_material.argtypes = [vp, i, ctypes.c_uint, i, vp, vp, vp, vp, vp, vp]
            lib.oref_scene_set_transforms.argtypes = [vp, vp, i]
            lib.oref_scene_add_light.argtypes = [vp, i, vp, vp, ctypes.c_float, ctypes.c_float]
            lib.oref_light_full.argtypes = [vp, i, i, vp, vp, vp, vp, vp]
            lib.oref_texture_sample.argtypes = [vp, i, vp, i, i, vp]
            lib.oref_dist2d.argtypes = [i, i, vp, i, vp, vp, vp, vp]
            lib.oref_image_samples.argtypes = [i, i, ctypes.c_uint, i, vp]
            lib.oref_nee_samples.argtypes = [i, ctypes.c_uint, i, vp]
            lib.oref_record_begin.argtypes = [vp, i, i]
            lib.oref_record_fetch.argtypes = [vp, vp, vp, vp, vp]
            lib.oref_record_fetch.restype = i
            lib.oref_intersect_fill.argtypes = [vp, vp, i, vp]
        _libs[kind] = (lib, pre)
    return _libs[kind]


def f32(a):
    return np.ascontiguousarray(a, np.float32)


def i32(a):
    return np.ascontiguousarray(a, np.int32)


class Oracle:
    def __init__(self, kind: str):
        self.kind = kind
        self.lib, self.pre = _load(kind)

    def _fn(self, name):
        return getattr(self.lib, self.pre + name)

    # ---- scene ----
    def scene(self, sc):
        return OracleScene(self, sc)

    # ---- stateless probes ----
    def bsdf(self, mat, frames, view, u, ea, eb):
        n = frames.shape[0]
        out = np.zeros((n, 12), np.float32)
        keep = [f32(mat), f32(frames), f32(view), f32(u), f32(ea), f32(eb)]
        self._fn("bsdf")(_p(keep[0]), n, *[_p(k) for k in keep[1:]], _p(out))
        return out

    def camera_rays(self, cam, samples):
        samples = f32(samples)
        out = np.zeros((samples.shape[0], 8), np.float32)
        c = f32(cam)
        self._fn("camera_rays")(_p(c), _p(samples), samples.shape[0], _p(out))
        return out

    def mcrng(self, seed, n):
        out = np.zeros(n, np.uint32)
        self._fn("mcrng")(seed, n, _p(out))
        return out

    def stratified(self, grid, seed, n):
        out = np.zeros((n, 2), np.float32)
        self._fn("stratified")(grid, seed, n, _p(out))
        return out

    # ---- the reference's samplers on the keyed stream (reference oracle only) ----
    def image_samples(self, sampler, arg, seed, npix):
        """[npix, spp, 2]: the image samples of <sampler>(arg) over a one-row image, reset
        after every pixel as ImageRenderer::Render does."""
        from banshee_amd.hip import SAMPLERS, num_samples
        spp = num_samples(sampler, arg)
        out = np.zeros((npix, spp, 2), np.float32)
        self.lib.oref_image_samples(SAMPLERS[sampler], arg, seed, npix, _p(out))
        return out

    def nee_samples(self, sampler, seed, n):
        """[n, 2]: successive samples of a light sampler <sampler>(1) in camera sample 0."""
        from banshee_amd.hip import SAMPLERS
        out = np.zeros((n, 2), np.float32)
        self.lib.oref_nee_samples(SAMPLERS[sampler], seed, n, _p(out))
        return out

    def dist2d(self, values, u, q):
        """Reference only: Distribution2D(n, m, values [m, n]): Sample2D at u (k x 2) -> (k x 3:
        uv, pdf) and Pdf at q (k x 2)."""
        values = f32(values)
        m, n = values.shape
        u, q = f32(u), f32(q)
        s = np.zeros((u.shape[0], 3), np.float32)
        p = np.zeros(q.shape[0], np.float32)
        self.lib.oref_dist2d(n, m, _p(values), u.shape[0], _p(u), _p(q), _p(s), _p(p))
        return s, p

    def sh_eval(self, dirs, lmax):
        dirs = f32(dirs)
        out = np.zeros((dirs.shape[0], (lmax + 1) ** 2), np.float32)
        self._fn("sh_eval")(_p(dirs), dirs.shape[0], lmax, _p(out))
        return out

    def sh_convolve(self, lmax, cin):
        cin = f32(cin)
        out = np.zeros_like(cin)
        self._fn("sh_convolve")(lmax, _p(cin), _p(out))
        return out

    def sh_project(self, env, lmax):
        env = f32(env)
        H, W = env.shape[:2]
        out = np.zeros(((lmax + 1) ** 2, 3), np.float32)
        self._fn("sh_project")(_p(env), W, H, lmax, _p(out))
        return out

    def sh_irradiance(self, lmax, coeffs, normals):
        coeffs, normals = f32(coeffs), f32(normals)
        out = np.zeros((normals.shape[0], 3), np.float32)
        self._fn("sh_irradiance")(lmax, _p(coeffs), _p(normals), normals.shape[0], _p(out))
        return out


class OracleScene:
    def __init__(self, orc: Oracle, sc):
        self.o = orc
        self.sc = sc
        self._keep = [f32(sc.verts), f32(sc.normals), i32(sc.tri_v), i32(sc.tri_n), i32(sc.tri_m),
                      i32(sc.bundles), f32(sc.materials), i32(sc.lights), f32(sc.camera)]
        k = self._keep
        textured = any(getattr(sc, a, None) is not None for a in ("uvs", "textures", "mat_maps", "normal_maps"))
        if textured:
            if orc.kind != "ref":
                raise RuntimeError("texture coordinates and maps need the reference oracle")
            uv = f32(sc.uvs).reshape(-1, 2) if sc.uvs is not None else None
            tu = i32(sc.tri_uv).reshape(-1, 3) if sc.tri_uv is not None else None
            texs = sc.textures or []
            flat = f32(np.concatenate([np.asarray(t, np.float32).ravel() for t in texs])) if texs else None
            desc, off = [], 0
            for t in texs:
                desc += [t.shape[1], t.shape[0], off]
                off += t.size
            desc = i32(desc) if texs else None
            mm = i32(sc.mat_maps).reshape(-1, 2) if sc.mat_maps is not None else None
            nm = i32(sc.normal_maps).ravel() if getattr(sc, "normal_maps", None) is not None else None
            self._keep += [a for a in (uv, tu, flat, desc, mm, nm) if a is not None]
            args = (_p(k[0]), k[0].shape[0], _p(k[1]), k[1].shape[0], _p(k[2]), _p(k[3]), _p(k[4]), k[2].shape[0],
                    _p(k[5]), k[5].shape[0], _p(k[6]), k[6].shape[0], _p(k[7]), k[7].shape[0], _p(k[8]),
                    _p(uv), 0 if uv is None else uv.shape[0], _p(tu), _p(flat), _p(desc), len(texs), _p(mm))
            if nm is not None:
                self.h = orc.lib.oref_scene_create_normal_mapped(*args, _p(nm))
            else:
                self.h = orc.lib.oref_scene_create_textured(*args)
        else:
            self.h = orc._fn("scene_create")(
                _p(k[0]), k[0].shape[0], _p(k[1]), k[1].shape[0], _p(k[2]), _p(k[3]), _p(k[4]), k[2].shape[0],
                _p(k[5]), k[5].shape[0], _p(k[6]), k[6].shape[0], _p(k[7]), k[7].shape[0], _p(k[8]))
        if not self.h:
            raise RuntimeError("oracle scene creation failed")
        xf = getattr(sc, "transforms", None)
        if xf is not None:
            if orc.kind != "ref":
                raise RuntimeError("bundle transforms need the reference oracle")
            x = f32(xf).reshape(-1, 32)
            self._keep.append(x)
            orc.lib.oref_scene_set_transforms(self.h, _p(x), x.shape[0])
        others = getattr(sc, "other_lights", None)
        if others is not None and len(others):
            if orc.kind != "ref":
                raise RuntimeError("point, directional and environment lights need the reference oracle")
            env = getattr(sc, "env", None)
            if env is not None:
                e = f32(env)
                self._keep.append(e)
                orc.lib.oref_scene_set_env(self.h, _p(e), e.shape[1], e.shape[0])
            for r in np.asarray(others, np.float32).reshape(-1, 9):
                v, em = f32(r[1:4]), f32(r[4:7])
                orc.lib.oref_scene_add_light(self.h, int(r[0]), _p(v), _p(em), float(r[7]), float(r[8]))

    def __del__(self):
        try:
            if self.h:
                self.o._fn("scene_destroy")(self.h)
        except Exception:
            pass

    def intersect(self, rays, any_hit=False):
        rays = f32(rays)
        n = rays.shape[0]
        prim = np.zeros(n, np.int32)
        thit = np.zeros((n, 3), np.float32)
        self.o._fn("intersect")(self.h, _p(rays), n, 1 if any_hit else 0, _p(prim), _p(thit))
        return prim, thit

    def intersect_fill(self, rays):
        rays = f32(rays)
        out = np.zeros((rays.shape[0], 20), np.float32)
        self.o.lib.oref_intersect_fill(self.h, _p(rays), rays.shape[0], _p(out))
        return out

    def render_keyed(self, W, H, x0, y0, tw, th, img_grid, nee_grid, depth, seed, per_sample=False):
        spp = img_grid * img_grid
        s = np.zeros((th, tw, 3), np.float32)
        q = np.zeros((th, tw, 3), np.float32)
        c = np.zeros((th, tw), np.float32)
        ps = np.zeros((th, tw, spp, 3), np.float32) if per_sample else None
        rays = np.zeros(2, np.int64)
        self.o._fn("render_keyed")(self.h, W, H, x0, y0, tw, th, img_grid, nee_grid, depth, seed, _p(s), _p(q),
                                   _p(c), _p(ps), _p(rays))
        return s, q, c, ps, rays

    def render_keyed_frame(self, W, H, samplers, img_arg, depth, seed, per_sample=False):
        """The whole frame through ImageRenderer::Render with samplers = (image, light, BSDF)
        kinds, the image sampler built with img_arg, the others with 1 (reference only)."""
        from banshee_amd.hip import SAMPLERS, num_samples
        spp = num_samples(samplers[0], img_arg)
        s = np.zeros((H, W, 3), np.float32)
        q = np.zeros((H, W, 3), np.float32)
        c = np.zeros((H, W), np.float32)
        ps = np.zeros((H, W, spp, 3), np.float32) if per_sample else None
        rays = np.zeros(2, np.int64)
        self.o.lib.oref_render_keyed_frame(self.h, W, H, SAMPLERS[samplers[0]], img_arg, SAMPLERS[samplers[1]],
                                           SAMPLERS[samplers[2]], depth, seed, _p(s), _p(q), _p(c), _p(ps),
                                           _p(rays))
        return s, q, c, ps, rays

    def time_render(self, W, H, img_grid, nee_grid, depth, seed, mt=True):
        rays = np.zeros(2, np.int64)
        t = self.o._fn("time_render")(self.h, W, H, img_grid, nee_grid, depth, seed, 1 if mt else 0, _p(rays))
        return t, rays

    def bake_sh(self, probes, N, lmax, depth, seed, tmin=0.0, threads=0, bg=(0.0, 0.0, 0.0)):
        """CPU restatement only: (coeffs float64 [P, nt, 3], radiance float32 [P, N, 3])."""
        probes = f32(probes).reshape(-1, 3)
        P = probes.shape[0]
        nt = (lmax + 1) ** 2
        c = np.zeros((P, nt, 3), np.float64)
        r = np.zeros((P, N, 3), np.float32)
        b = f32(bg)
        self.o.lib.cref_bake_sh(self.h, _p(probes), P, N, lmax, depth, seed, tmin, _p(b), _p(c), _p(r), threads)
        return c, r

    def render_bsdf_only(self, W, H, img_grid, depth, seed, crop=None, bg=(0.0, 0.0, 0.0), threads=0):
        """CPU restatement only: the BSDF-sampling-only estimator of the quirk-free integral.
        Returns (sum float64 [h, w, 3], sumsq float64 [h, w, 3], n float32 [h, w])."""
        x0, y0, tw, th = crop if crop is not None else (0, 0, W, H)
        s = np.zeros((th, tw, 3), np.float64)
        q = np.zeros((th, tw, 3), np.float64)
        n = np.zeros((th, tw), np.float32)
        b = f32(bg)
        self.o.lib.cref_render_bsdf_only(self.h, W, H, x0, y0, tw, th, img_grid, depth, seed, _p(b), _p(s), _p(q),
                                         _p(n), threads)
        return s, q, n

    def light_full(self, idx, p, nrm, u, w):
        """Reference only: Light::GetSample / GetPdf / GetLe, n x 16 (bsh_diag_light_full's layout)."""
        arrs = [f32(a) for a in (p, nrm, u, w)]
        out = np.zeros((arrs[0].shape[0], 16), np.float32)
        self.o.lib.oref_light_full(self.h, idx, arrs[0].shape[0], *[_p(a) for a in arrs], _p(out))
        return out

    def material(self, idx, seed, frames, view, u, ea, eb):
        """Reference only: Material::Sample / Evaluate / GetPdf of world.materials_[idx], rand()
        of case i keyed by ks_base(seed, i): n x 12 (bsh_diag_material's layout)."""
        arrs = [f32(a) for a in (frames, view, u, ea, eb)]
        out = np.zeros((arrs[0].shape[0], 12), np.float32)
        self.o.lib.oref_material(self.h, idx, seed, arrs[0].shape[0], *[_p(a) for a in arrs], _p(out))
        return out

    def texture_sample(self, uv, filt=1, wrap=0):
        """Reference only: the texture system's lookup of the scene's environment map."""
        uv = f32(uv)
        out = np.zeros((uv.shape[0], 3), np.float32)
        self.o.lib.oref_texture_sample(self.h, uv.shape[0], _p(uv), filt, wrap, _p(out))
        return out

    def light(self, idx, p, u, w):
        p, u, w = f32(p), f32(u), f32(w)
        out = np.zeros((p.shape[0], 8), np.float32)
        self.o._fn("light")(self.h, idx, p.shape[0], _p(p), _p(u), _p(w), _p(out))
        return out

    def record(self, max_rays, stride):
        self.o.lib.oref_record_begin(self.h, max_rays, stride)

    def fetch(self, max_rays):
        rays = np.zeros((max_rays, 8), np.float32)
        kind = np.zeros(max_rays, np.int32)
        prim = np.zeros(max_rays, np.int32)
        thit = np.zeros((max_rays, 3), np.float32)
        n = self.o.lib.oref_record_fetch(self.h, _p(rays), _p(kind), _p(prim), _p(thit))
        return rays[:n], kind[:n], prim[:n], thit[:n]
