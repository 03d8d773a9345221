"""ctypes binding of libbanshee_hip.so (the C ABI in include/banshee_hip.h).

The library is the only compute path: if it is missing or no gfx950 device is
present, every call raises; nothing falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbanshee_hip.so")

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_size_t = ctypes.c_size_t
c_uint32 = ctypes.c_uint32


class BshError(RuntimeError):
    pass


class Ray(ctypes.Structure):
    _fields_ = [("ox", ctypes.c_float), ("oy", ctypes.c_float), ("oz", ctypes.c_float), ("tmin", ctypes.c_float),
                ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float), ("tmax", ctypes.c_float)]


class Bundle(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("tri_begin", "tri_count", "vert_begin", "vert_count", "norm_begin",
                                              "norm_count")]


class Material(ctypes.Structure):
    _fields_ = [("type", ctypes.c_float), ("color", ctypes.c_float * 3), ("eta", ctypes.c_float),
                ("alpha", ctypes.c_float), ("pad", ctypes.c_float * 2)]


class AreaLight(ctypes.Structure):
    _fields_ = [("bundle", ctypes.c_int32), ("shape", ctypes.c_int32)]


class Camera(ctypes.Structure):
    _fields_ = [("eye", ctypes.c_float * 3), ("at", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("znear", ctypes.c_float), ("zfar", ctypes.c_float), ("fovy", ctypes.c_float),
                ("aspect", ctypes.c_float)]


class BvhParams(ctypes.Structure):
    _fields_ = [("max_leaf", ctypes.c_int32), ("sah_bins", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("width", ctypes.c_int32), ("quantized", ctypes.c_int32)]


class RenderParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("crop_x", ctypes.c_int32),
                ("crop_y", ctypes.c_int32), ("crop_w", ctypes.c_int32), ("crop_h", ctypes.c_int32),
                ("img_grid", ctypes.c_int32), ("spp_begin", ctypes.c_int32), ("spp_count", ctypes.c_int32),
                ("max_depth", ctypes.c_int32), ("seed", ctypes.c_uint32), ("chunk_spp", ctypes.c_int32),
                ("bg", ctypes.c_float * 3), ("flags", ctypes.c_int32), ("img_sampler", ctypes.c_int32),
                ("light_sampler", ctypes.c_int32), ("brdf_sampler", ctypes.c_int32)]


# BSH_SAMPLER_* kinds
SAMPLERS = {"stratified": 0, "regular": 1, "random": 2, "cmj": 3, "sobol": 4}


def num_samples(sampler: str, arg: int) -> int:
    """Samples per pixel of the image sampler `sampler` built with argument `arg`
    (Sampler::num_samples)."""
    if sampler in ("stratified", "regular", "cmj"):
        return arg * arg
    if sampler == "random":
        return arg
    if sampler == "sobol":
        return 1 << max(0, int(arg - 1).bit_length())
    raise ValueError(sampler)


def _set_samplers(p, samplers):
    img, light, brdf = samplers
    p.img_sampler, p.light_sampler, p.brdf_sampler = SAMPLERS[img], SAMPLERS[light], SAMPLERS[brdf]


class Texture(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("offset", ctypes.c_uint64)]


class Light(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("v", ctypes.c_float * 3), ("e", ctypes.c_float * 3),
                ("scale", ctypes.c_float), ("gamma", ctypes.c_float)]


class ProbeParams(ctypes.Structure):
    _fields_ = [("rays_per_probe", ctypes.c_int32), ("lmax", ctypes.c_int32), ("max_depth", ctypes.c_int32),
                ("seed", ctypes.c_uint32), ("chunk_probes", ctypes.c_int32), ("tmin", ctypes.c_float),
                ("bg", ctypes.c_float * 3), ("flags", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    _fields_ = [("paths_per_chunk", ctypes.c_int32), ("trace_blocks_per_cu", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("refill_lanes", ctypes.c_int32), ("claim_rays", ctypes.c_int32),
                ("lanes", ctypes.c_int32)]


class Job(ctypes.Structure):
    _fields_ = [("scene", c_void_p), ("kind", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("render", RenderParams), ("accum", c_void_p), ("samples", c_void_p), ("probe", ProbeParams),
                ("probes", c_void_p), ("nprobes", c_size_t), ("coeffs", c_void_p), ("radiance", c_void_p)]


class GuideParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint32), ("lr", ctypes.c_float), ("beta1", ctypes.c_float),
                ("beta2", ctypes.c_float), ("eps", ctypes.c_float), ("bsdf_fraction", ctypes.c_float),
                ("guided_bounces", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class RenderStats(ctypes.Structure):
    _fields_ = [("rays_closest", ctypes.c_uint64), ("rays_any", ctypes.c_uint64),
                ("camera_samples", ctypes.c_uint64), ("node_visits", ctypes.c_uint64),
                ("tri_tests", ctypes.c_uint64), ("trace_ms", ctypes.c_double), ("shade_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("trace_launches", ctypes.c_uint32),
                ("shade_launches", ctypes.c_uint32), ("guide_launches", ctypes.c_uint32),
                ("guide_rows", ctypes.c_uint64), ("guide_ms", ctypes.c_double),
                ("graph_replayed", ctypes.c_uint32)]


class LaunchStats(ctypes.Structure):
    _fields_ = [("rays_closest", ctypes.c_uint64), ("rays_any", ctypes.c_uint64),
                ("nodes_closest", ctypes.c_uint64), ("tris_closest", ctypes.c_uint64),
                ("nodes_any", ctypes.c_uint64), ("tris_any", ctypes.c_uint64), ("trace_ms", ctypes.c_double)]


# Every symbol include/banshee_hip.h declares.
EXPORTS = [
    "bsh_context_create", "bsh_context_destroy", "bsh_last_error", "bsh_device_alloc", "bsh_device_free",
    "bsh_copy_to_device", "bsh_copy_to_host", "bsh_device_sync", "bsh_scene_create", "bsh_scene_set_materials",
    "bsh_scene_set_lights", "bsh_scene_set_camera", "bsh_scene_commit", "bsh_scene_destroy", "bsh_scene_info",
    "bsh_intersect_closest", "bsh_intersect_any", "bsh_intersect_count", "bsh_render", "bsh_render_stats_get",
    "bsh_diag_bsdf", "bsh_diag_light", "bsh_diag_camera", "bsh_diag_keyed",
    "bsh_bake_sh", "bsh_sh_project_env", "bsh_sh_irradiance", "bsh_diag_sh_eval",
    "bsh_batch_run", "bsh_context_set_tuning", "bsh_context_get_tuning", "bsh_context_stats_get", "bsh_autotune",
    "bsh_guide_num_params", "bsh_guide_create", "bsh_guide_destroy", "bsh_guide_get_params", "bsh_guide_set_params",
    "bsh_guide_forward", "bsh_guide_gradient", "bsh_guide_train_step", "bsh_render_guided", "bsh_diag_mfma",
    "bsh_diag_guide", "bsh_bvh_build_host", "bsh_guide_set_bsdf_fraction", "bsh_guide_train",
    "bsh_bake_sh_guided", "bsh_guide_forward_fp32", "bsh_guide_gradient_fp32",
    "bsh_context_launch_stats", "bsh_diag_stream_read", "bsh_import_obj", "bsh_sh_project_env_noalloc",
    "bsh_diag_samples", "bsh_scene_set_other_lights", "bsh_scene_set_env_map", "bsh_diag_light_full",
    "bsh_diag_material", "bsh_guide_get_state", "bsh_guide_set_state", "bsh_scene_set_transforms",
    "bsh_scene_set_uvs", "bsh_scene_set_textures", "bsh_scene_set_material_maps", "bsh_scene_set_normal_maps",
]


class AssetCounts(ctypes.Structure):
    _fields_ = [(n, c_size_t) for n in ("nverts", "nnormals", "ntris", "nbundles", "nmaterials", "nlights",
                                         "ignored_maps")]


def import_obj(path: str, camera, name: str = ""):
    """bsh_import_obj (host only): the World arrays the reference's AssimpAssetImporter builds from
    an OBJ/MTL file, as a banshee_amd.scenes.Scene with the given camera (13 floats: eye, at, up,
    znear, zfar, fovy, aspect)."""
    from .scenes import Scene as SceneData
    c = AssetCounts()
    p = os.fsencode(path)
    _check(lib().bsh_import_obj(p, ctypes.byref(c), *([None] * 8)), "bsh_import_obj")
    v = np.zeros((c.nverts, 3), np.float32)
    n = np.zeros((c.nnormals, 3), np.float32)
    tv = np.zeros((c.ntris, 3), np.int32)
    tn = np.zeros((c.ntris, 3), np.int32)
    tm = np.zeros(c.ntris, np.int32)
    b = np.zeros((c.nbundles, 6), np.int32)
    m = np.zeros((c.nmaterials, 8), np.float32)
    li = np.zeros((max(c.nlights, 1), 2), np.int32)
    _check(lib().bsh_import_obj(p, ctypes.byref(c), _p(v), _p(n), _p(tv), _p(tn), _p(tm), _p(b), _p(m), _p(li)),
           "bsh_import_obj")
    return SceneData(verts=v, normals=n, tri_v=tv, tri_n=tn, tri_m=tm, bundles=b, materials=m,
                     lights=np.ascontiguousarray(li[:c.nlights]), camera=np.asarray(camera, np.float32),
                     name=name or os.path.basename(path), meta={"ignored_maps": int(c.ignored_maps)})

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BshError(f"{LIB_PATH} is not built; run `python -m banshee_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.bsh_last_error.restype = ctypes.c_char_p
        for name in EXPORTS:
            f = getattr(L, name)
            if name == "bsh_guide_num_params":
                f.restype = c_size_t
            elif name != "bsh_last_error":
                f.restype = c_int
        _lib = L
    return _lib


def _check(code: int, what: str):
    if code != 0:
        msg = lib().bsh_last_error().decode()
        raise BshError(f"{what} failed ({code}): {msg}")


def _p(a):
    return a.ctypes.data_as(c_void_p)


class DeviceArray:
    """A device allocation owned by a Context, with numpy-shaped host copies."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = c_void_p()
        _check(lib().bsh_device_alloc(ctx.h, c_size_t(self.nbytes), ctypes.byref(self.ptr)), "bsh_device_alloc")

    @classmethod
    def from_host(cls, ctx, a: np.ndarray):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        _check(lib().bsh_copy_to_device(self.ctx.h, self.ptr, _p(a), c_size_t(self.nbytes)), "bsh_copy_to_device")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _check(lib().bsh_copy_to_host(self.ctx.h, _p(out), self.ptr, c_size_t(self.nbytes)), "bsh_copy_to_host")
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            lib().bsh_device_free(self.ctx.h, self.ptr)
            self.ptr = c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    def __init__(self, device: int = 0):
        self.h = c_void_p()
        _check(lib().bsh_context_create(c_int(device), ctypes.byref(self.h)), "bsh_context_create")

    def sync(self):
        _check(lib().bsh_device_sync(self.h), "bsh_device_sync")

    def alloc(self, shape, dtype) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def to_device(self, a) -> DeviceArray:
        return DeviceArray.from_host(self, a)

    # ---- diagnostics ----
    def diag_bsdf(self, mat_row, frames, view, u, ea, eb):
        m = Material()
        m.type = float(mat_row[0])
        for k in range(3):
            m.color[k] = float(mat_row[1 + k])
        m.eta = float(mat_row[4])
        m.alpha = float(mat_row[5])
        arrs = [np.ascontiguousarray(x, np.float32) for x in (frames, view, u, ea, eb)]
        n = arrs[0].shape[0]
        out = np.zeros((n, 12), np.float32)
        _check(lib().bsh_diag_bsdf(self.h, ctypes.byref(m), c_size_t(n), *[_p(a) for a in arrs], _p(out)),
               "bsh_diag_bsdf")
        return out

    def diag_camera(self, cam13, samples):
        c = camera_struct(cam13)
        s = np.ascontiguousarray(samples, np.float32)
        out = np.zeros((s.shape[0], 8), np.float32)
        _check(lib().bsh_diag_camera(self.h, ctypes.byref(c), c_size_t(s.shape[0]), _p(s), _p(out)),
               "bsh_diag_camera")
        return out

    def diag_keyed(self, seed, camsample, stream, counter):
        a = np.ascontiguousarray(camsample, np.uint32)
        b = np.ascontiguousarray(stream, np.uint32)
        c = np.ascontiguousarray(counter, np.uint32)
        out = np.zeros(a.shape[0], np.uint32)
        _check(lib().bsh_diag_keyed(self.h, c_uint32(seed), c_size_t(a.shape[0]), _p(a), _p(b), _p(c), _p(out)),
               "bsh_diag_keyed")
        return out

    def diag_samples(self, sampler, arg, seed, n, image=True):
        """Device sampler patterns (bsh_diag_samples): image samples of n pixels [n, spp, 2], or
        the first n light-sampler samples [n, 2]."""
        spp = num_samples(sampler, arg) if image else 1
        out = np.zeros((n * spp, 2), np.float32)
        _check(lib().bsh_diag_samples(self.h, c_int(SAMPLERS[sampler]), c_int(arg), c_int(1 if image else 0),
                                      c_uint32(seed), c_size_t(n * spp), _p(out)), "bsh_diag_samples")
        return out.reshape(n, spp, 2) if image else out

    def diag_sh_eval(self, dirs, lmax):
        d = np.ascontiguousarray(dirs, np.float32)
        out = np.zeros((d.shape[0], (lmax + 1) ** 2), np.float32)
        _check(lib().bsh_diag_sh_eval(self.h, c_int(lmax), c_size_t(d.shape[0]), _p(d), _p(out)), "bsh_diag_sh_eval")
        return out

    def sh_project_env(self, env, lmax):
        env = np.ascontiguousarray(env, np.float32)
        H, W = env.shape[:2]
        de = self.to_device(env)
        dc = self.alloc(((lmax + 1) ** 2, 3), np.float32)
        _check(lib().bsh_sh_project_env(self.h, de.ptr, c_int(W), c_int(H), c_int(lmax), dc.ptr, c_void_p(None)),
               "bsh_sh_project_env")
        self.sync()
        return dc.numpy()

    def sh_project_env_noalloc(self, env, lmax, env_dev=None):
        """bsh_sh_project_env_noalloc on `env` (H x W x 3; or on `env_dev`, a device copy of it)."""
        env = np.ascontiguousarray(env, np.float32)
        H, W = env.shape[:2]
        de = env_dev if env_dev is not None else self.to_device(env)
        dc = self.alloc(((lmax + 1) ** 2, 3), np.float32)
        _check(lib().bsh_sh_project_env_noalloc(self.h, de.ptr, c_int(W), c_int(H), c_int(lmax), dc.ptr,
                                                c_void_p(None)), "bsh_sh_project_env_noalloc")
        self.sync()
        return dc.numpy()

    def sh_irradiance(self, lmax, coeffs, normals):
        dc = self.to_device(np.ascontiguousarray(coeffs, np.float32))
        dn = self.to_device(np.ascontiguousarray(normals, np.float32))
        n = dn.shape[0]
        do = self.alloc((n, 3), np.float32)
        _check(lib().bsh_sh_irradiance(self.h, c_int(lmax), dc.ptr, dn.ptr, c_size_t(n), do.ptr, c_void_p(None)),
               "bsh_sh_irradiance")
        self.sync()
        return do.numpy()

    def diag_mfma(self, a, b):
        """One 16x16x16 FP16 MFMA tile on the device: returns a @ b in FP32 (a, b float16 16x16)."""
        a = np.ascontiguousarray(a, np.float16).view(np.uint16)
        b = np.ascontiguousarray(b, np.float16).view(np.uint16)
        d = np.zeros((16, 16), np.float32)
        _check(lib().bsh_diag_mfma(self.h, _p(a), _p(b), _p(d)), "bsh_diag_mfma")
        return d

    def diag_guide(self, outputs, u, dirs):
        """Device lobe decoding / mixture pdf / vMF sampling on given outputs (n x 48).
        Returns n x 8: sample xyz, pdf(sample), pdf(dirs), sum of lobe weights."""
        o = np.ascontiguousarray(outputs, np.float32).reshape(-1, 48)
        u = np.ascontiguousarray(u, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros((o.shape[0], 8), np.float32)
        _check(lib().bsh_diag_guide(self.h, c_size_t(o.shape[0]), _p(o), _p(u), _p(d), _p(out)), "bsh_diag_guide")
        return out

    # ---- batching ----
    def tuning(self) -> dict:
        t = Tuning()
        _check(lib().bsh_context_get_tuning(self.h, ctypes.byref(t)), "bsh_context_get_tuning")
        return {k: getattr(t, k) for k, _ in t._fields_ if k != "reserved"}

    def set_tuning(self, paths_per_chunk=0, trace_blocks_per_cu=0, use_graph=0, refill_lanes=0, claim_rays=0, lanes=0):
        t = Tuning(paths_per_chunk, trace_blocks_per_cu, use_graph, refill_lanes, claim_rays, lanes)
        _check(lib().bsh_context_set_tuning(self.h, ctypes.byref(t)), "bsh_context_set_tuning")

    def stats(self) -> dict:
        st = RenderStats()
        _check(lib().bsh_context_stats_get(self.h, ctypes.byref(st)), "bsh_context_stats_get")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def stream_read_gbs(self, buf: DeviceArray, reps: int, per_xcd: bool = False, nontemporal: bool = False,
                        small_grid: bool = False) -> float:
        """GB/s of bsh_diag_stream_read over all of `buf` (per_xcd: every XCD reads all of it;
        nontemporal: non-temporal loads; small_grid: 8 workgroups per CU instead of 32)."""
        ms = ctypes.c_double()
        flags = int(per_xcd) | (2 if nontemporal else 0) | (4 if small_grid else 0)
        _check(lib().bsh_diag_stream_read(self.h, buf.ptr, c_size_t(buf.nbytes), c_int(reps), c_int(flags),
                                          ctypes.byref(ms)), "bsh_diag_stream_read")
        moved = buf.nbytes * reps * (8 if per_xcd else 1)
        return moved / (ms.value * 1e-3) / 1e9

    def launch_stats(self) -> list:
        """Per bounce launch of the last run: dicts of LaunchStats' fields."""
        n = c_size_t()
        _check(lib().bsh_context_launch_stats(self.h, None, c_size_t(0), ctypes.byref(n)), "bsh_context_launch_stats")
        arr = (LaunchStats * max(n.value, 1))()
        _check(lib().bsh_context_launch_stats(self.h, arr, c_size_t(n.value), ctypes.byref(n)),
               "bsh_context_launch_stats")
        return [{k: getattr(arr[i], k) for k, _ in LaunchStats._fields_} for i in range(n.value)]

    def batch_run(self, jobs, tuning: dict | None = None, stream=None):
        """jobs: list of BatchJob. Runs them in one wavefront (one hipGraph when use_graph)."""
        arr = (Job * len(jobs))(*[j.struct() for j in jobs])
        t = None
        if tuning is not None:
            t = Tuning(tuning.get("paths_per_chunk", 0), tuning.get("trace_blocks_per_cu", 0),
                       tuning.get("use_graph", 0), tuning.get("refill_lanes", 0), tuning.get("claim_rays", 0),
                       tuning.get("lanes", 0))
        _check(lib().bsh_batch_run(self.h, arr, c_size_t(len(jobs)), ctypes.byref(t) if t else None,
                                   c_void_p(stream)), "bsh_batch_run")

    def autotune(self, jobs, paths=None, blocks=None):
        arr = (Job * len(jobs))(*[j.struct() for j in jobs])
        pa = np.ascontiguousarray(paths if paths is not None else [], np.int32)
        bl = np.ascontiguousarray(blocks if blocks is not None else [], np.int32)
        best = Tuning()
        ms = ctypes.c_double()
        _check(lib().bsh_autotune(self.h, arr, c_size_t(len(jobs)), _p(pa) if pa.size else None, c_size_t(pa.size),
                                  _p(bl) if bl.size else None, c_size_t(bl.size), ctypes.byref(best),
                                  ctypes.byref(ms)), "bsh_autotune")
        return {"paths_per_chunk": best.paths_per_chunk, "trace_blocks_per_cu": best.trace_blocks_per_cu,
                "refill_lanes": best.refill_lanes, "claim_rays": best.claim_rays, "lanes": best.lanes}, ms.value

    def close(self):
        if self.h:
            lib().bsh_context_destroy(self.h)
            self.h = c_void_p()


def render_params(width, height, img_grid, max_depth, seed, crop=None, spp_range=None, chunk_spp=0,
                  bg=(0.0, 0.0, 0.0), count=False, accumulate=False,
                  samplers=("stratified", "stratified", "stratified")) -> RenderParams:
    p = RenderParams()
    p.width, p.height = width, height
    if crop is not None:
        p.crop_x, p.crop_y, p.crop_w, p.crop_h = crop
    p.img_grid = img_grid
    _set_samplers(p, samplers)
    spp = num_samples(samplers[0], img_grid)
    sb, sc = (spp_range if spp_range else (0, spp))
    p.spp_begin, p.spp_count = sb, sc
    p.max_depth = max_depth
    p.seed = seed
    p.chunk_spp = chunk_spp
    for k in range(3):
        p.bg[k] = bg[k]
    p.flags = (4 if count else 0) | (1 if accumulate else 0)
    return p


class BatchJob:
    """A render or probe-bake job for Context.batch_run; owns its device outputs."""

    def __init__(self, scene: "Scene", kind: str, **kw):
        self.scene = scene
        self.kind = kind
        ctx = scene.ctx
        if kind == "render":
            self.params = render_params(**{k: v for k, v in kw.items() if k != "per_sample"})
            crop = kw.get("crop")
            cw = crop[2] if crop else kw["width"]
            ch = crop[3] if crop else kw["height"]
            sc = self.params.spp_count
            self.accum = ctx.alloc((ch, cw, 8), np.float32)
            self.samples = ctx.alloc((ch, cw, sc, 3), np.float32) if kw.get("per_sample") else None
        else:
            self.probes = ctx.to_device(np.ascontiguousarray(kw["probes"], np.float32).reshape(-1, 3))
            n = self.probes.shape[0]
            pp = ProbeParams()
            pp.rays_per_probe, pp.lmax, pp.max_depth, pp.seed = kw["rays_per_probe"], kw["lmax"], kw["max_depth"], \
                kw["seed"]
            pp.tmin = kw.get("tmin", 0.0)
            self.pparams = pp
            self.coeffs = ctx.alloc((n, (kw["lmax"] + 1) ** 2, 3), np.float32)
            self.radiance = ctx.alloc((n, kw["rays_per_probe"], 3), np.float32) if kw.get("radiance") else None

    def struct(self) -> Job:
        j = Job()
        j.scene = self.scene.h
        if self.kind == "render":
            j.kind = 0
            j.render = self.params
            j.accum = self.accum.ptr
            j.samples = self.samples.ptr if self.samples else None
        else:
            j.kind = 1
            j.probe = self.pparams
            j.probes = self.probes.ptr
            j.nprobes = self.probes.shape[0]
            j.coeffs = self.coeffs.ptr
            j.radiance = self.radiance.ptr if self.radiance else None
        return j

    def result(self):
        if self.kind == "render":
            a = self.accum.numpy()
            return a[..., 0:3], a[..., 4:7], a[..., 3], (self.samples.numpy() if self.samples else None)
        return self.coeffs.numpy(), (self.radiance.numpy() if self.radiance else None)


RECORD_FLOATS = 12
GUIDE_OUT = 48


class Guide:
    """The path-guiding network (vMF mixture MLP) owned by a Context."""

    def __init__(self, ctx: Context, seed=1, lr=5e-3, beta1=0.9, beta2=0.99, eps=1e-8, bsdf_fraction=0.5,
                 guided_bounces=0):
        self.ctx = ctx
        self.h = c_void_p()
        p = GuideParams()
        p.seed, p.lr, p.beta1, p.beta2, p.eps, p.bsdf_fraction = seed, lr, beta1, beta2, eps, bsdf_fraction
        p.guided_bounces = guided_bounces
        _check(lib().bsh_guide_create(ctx.h, ctypes.byref(p), ctypes.byref(self.h)), "bsh_guide_create")

    def set_bsdf_fraction(self, a: float):
        _check(lib().bsh_guide_set_bsdf_fraction(self.h, ctypes.c_float(a)), "bsh_guide_set_bsdf_fraction")

    @staticmethod
    def num_params() -> int:
        return int(lib().bsh_guide_num_params())

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params(), np.float32)
        _check(lib().bsh_guide_get_params(self.h, _p(out)), "bsh_guide_get_params")
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.num_params()
        _check(lib().bsh_guide_set_params(self.h, _p(w)), "bsh_guide_set_params")

    def get_state(self) -> dict:
        """Training state (bsh_guide_get_state): parameters, Adam moments and step count."""
        n = self.num_params()
        p, m1, m2 = (np.zeros(n, np.float32) for _ in range(3))
        step = ctypes.c_int64()
        _check(lib().bsh_guide_get_state(self.h, _p(p), _p(m1), _p(m2), ctypes.byref(step)), "bsh_guide_get_state")
        return {"params": p, "m1": m1, "m2": m2, "step": int(step.value)}

    def set_state(self, st: dict):
        arrs = [np.ascontiguousarray(st[k], np.float32) for k in ("params", "m1", "m2")]
        assert all(a.size == self.num_params() for a in arrs)
        _check(lib().bsh_guide_set_state(self.h, *[_p(a) for a in arrs], ctypes.c_int64(int(st["step"]))),
               "bsh_guide_set_state")

    def forward(self, inputs, fp32=False):
        """inputs: n x 8 raw rows (position in [0,1]^3, roughness, normal, unused) -> n x 48 outputs.
        fp32=True: the FP32 verification mode (bsh_guide_forward_fp32)."""
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 8)
        n = x.shape[0]
        dx = self.ctx.to_device(x)
        do = self.ctx.alloc((n, GUIDE_OUT), np.float32)
        name = "bsh_guide_forward_fp32" if fp32 else "bsh_guide_forward"
        _check(getattr(lib(), name)(self.h, dx.ptr, c_size_t(n), do.ptr, c_void_p(None)), name)
        self.ctx.sync()
        return do.numpy()

    def gradient(self, records, fp32=False):
        """records: n x 12 (host array or DeviceArray) -> (gradient float32[num_params], loss).
        fp32=True: the FP32 verification mode (bsh_guide_gradient_fp32)."""
        dr = records if isinstance(records, DeviceArray) else self.ctx.to_device(
            np.ascontiguousarray(records, np.float32).reshape(-1, RECORD_FLOATS))
        n = dr.shape[0]
        g = np.zeros(self.num_params(), np.float32)
        loss = ctypes.c_float()
        name = "bsh_guide_gradient_fp32" if fp32 else "bsh_guide_gradient"
        _check(getattr(lib(), name)(self.h, dr.ptr, c_size_t(n), _p(g), ctypes.byref(loss)), name)
        return g, loss.value

    def train_step(self, records, n=None, offset=0, want_loss=True):
        """One Adam step on rows [offset, offset + n) of `records` (host array or DeviceArray).
        Returns the weighted mean loss (want_loss=False skips its device-to-host read)."""
        dr = records if isinstance(records, DeviceArray) else self.ctx.to_device(
            np.ascontiguousarray(records, np.float32).reshape(-1, RECORD_FLOATS))
        n = dr.shape[0] - offset if n is None else n
        assert 0 <= offset and offset + n <= dr.shape[0]
        ptr = c_void_p(dr.ptr.value + offset * RECORD_FLOATS * 4)
        loss = ctypes.c_float()
        _check(lib().bsh_guide_train_step(self.h, ptr, c_size_t(n), ctypes.byref(loss) if want_loss else None,
                                          c_void_p(None)), "bsh_guide_train_step")
        return loss.value if want_loss else None

    def train(self, records: DeviceArray, n: int, steps: int, batch: int, want_loss=False):
        """`steps` Adam steps on batches of up to `batch` of the first n device records: step s
        takes every stride-th record from offset s mod stride, stride = max(1, n // batch)."""
        loss = ctypes.c_float()
        _check(lib().bsh_guide_train(self.h, records.ptr, c_size_t(n), c_int(steps), c_size_t(batch),
                                     ctypes.byref(loss) if want_loss else None, c_void_p(None)), "bsh_guide_train")
        return loss.value if want_loss else None

    def close(self):
        if self.h:
            lib().bsh_guide_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _quant_flag(quantized):
    """bsh_bvh_params.quantized: None -> 0 (the default: quantised 4-wide nodes), True -> 1,
    False -> -1 (float 4-wide nodes)."""
    return 0 if quantized is None else 1 if quantized else -1


def bvh_build_host(tri_verts, max_leaf=4, sah_bins=16, threads=0, width=0, quantized=None):
    """Host-only BVH build (no device): returns (nodes float32 [n, 16 or 32], tri_order, width, depth).
    Node rows are the raw 64-B (binary, or quantised 4-wide) or 128-B (float 4-wide) records; the
    int fields are returned by viewing the row as int32."""
    tv = np.ascontiguousarray(tri_verts, np.float32).reshape(-1, 9)
    bp = BvhParams(max_leaf, sah_bins, threads, width, _quant_flag(quantized))
    n, w, d, nb = c_size_t(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    _check(lib().bsh_bvh_build_host(_p(tv), c_size_t(tv.shape[0]), ctypes.byref(bp), None, c_size_t(0), None,
                                    ctypes.byref(n), ctypes.byref(w), ctypes.byref(d), ctypes.byref(nb)),
           "bsh_bvh_build_host")
    cols = nb.value // 4
    nodes = np.zeros((n.value, cols), np.float32)
    order = np.zeros(tv.shape[0], np.int32)
    _check(lib().bsh_bvh_build_host(_p(tv), c_size_t(tv.shape[0]), ctypes.byref(bp), _p(nodes), c_size_t(n.value),
                                    _p(order), ctypes.byref(n), ctypes.byref(w), ctypes.byref(d), None),
           "bsh_bvh_build_host")
    return nodes, order, w.value, d.value


def camera_struct(cam13) -> Camera:
    c = Camera()
    v = [float(x) for x in cam13]
    for k in range(3):
        c.eye[k], c.at[k], c.up[k] = v[k], v[3 + k], v[6 + k]
    c.znear, c.zfar, c.fovy, c.aspect = v[9], v[10], v[11], v[12]
    return c


class Scene:
    """A committed device scene (World after Commit) built from a banshee_amd.scenes.Scene."""

    def __init__(self, ctx: Context, sc, max_leaf=4, sah_bins=16, threads=0, width=0, quantized=None):
        """quantized: None (the default) or True build quantised 64-B 4-wide nodes, False float
        128-B 4-wide nodes."""
        self.ctx = ctx
        self.sc = sc
        self.h = c_void_p()
        v = np.ascontiguousarray(sc.verts, np.float32)
        n = np.ascontiguousarray(sc.normals, np.float32)
        tv = np.ascontiguousarray(sc.tri_v, np.int32)
        tn = np.ascontiguousarray(sc.tri_n, np.int32)
        tm = np.ascontiguousarray(sc.tri_m, np.int32)
        b = np.ascontiguousarray(sc.bundles, np.int32)
        _check(lib().bsh_scene_create(ctx.h, _p(v), c_size_t(v.shape[0]), _p(n), c_size_t(n.shape[0]), _p(tv), _p(tn),
                                      _p(tm), c_size_t(tm.shape[0]), _p(b), c_size_t(b.shape[0]),
                                      ctypes.byref(self.h)), "bsh_scene_create")
        mats = np.zeros((sc.materials.shape[0], 8), np.float32)
        mats[:, :sc.materials.shape[1]] = sc.materials[:, :8]
        _check(lib().bsh_scene_set_materials(self.h, _p(mats), c_size_t(mats.shape[0])), "bsh_scene_set_materials")
        lights = np.ascontiguousarray(sc.lights, np.int32).reshape(-1, 2)
        _check(lib().bsh_scene_set_lights(self.h, _p(lights), c_size_t(lights.shape[0])), "bsh_scene_set_lights")
        uv = getattr(sc, "uvs", None)
        if uv is not None:
            u = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
            tu = np.ascontiguousarray(sc.tri_uv, np.int32).reshape(-1, 3)
            _check(lib().bsh_scene_set_uvs(self.h, _p(u), c_size_t(u.shape[0]), _p(tu)), "bsh_scene_set_uvs")
        texs = getattr(sc, "textures", None)
        if texs is not None and len(texs):
            flat = np.concatenate([np.ascontiguousarray(t, np.float32).ravel() for t in texs])
            descs = (Texture * len(texs))()
            off = 0
            for k, t in enumerate(texs):
                descs[k].height, descs[k].width = int(t.shape[0]), int(t.shape[1])
                descs[k].offset = off
                off += t.size
            _check(lib().bsh_scene_set_textures(self.h, _p(flat), c_size_t(flat.size), descs, c_size_t(len(texs))),
                   "bsh_scene_set_textures")
        mm = getattr(sc, "mat_maps", None)
        if mm is not None:
            m2 = np.ascontiguousarray(mm, np.int32).reshape(-1, 2)
            _check(lib().bsh_scene_set_material_maps(self.h, _p(m2), c_size_t(m2.shape[0])),
                   "bsh_scene_set_material_maps")
        nm = getattr(sc, "normal_maps", None)
        if nm is not None:
            n1 = np.ascontiguousarray(nm, np.int32).ravel()
            _check(lib().bsh_scene_set_normal_maps(self.h, _p(n1), c_size_t(n1.size)), "bsh_scene_set_normal_maps")
        xf = getattr(sc, "transforms", None)
        if xf is not None:
            x = np.ascontiguousarray(xf, np.float32).reshape(-1, 32)
            _check(lib().bsh_scene_set_transforms(self.h, _p(x), c_size_t(x.shape[0])), "bsh_scene_set_transforms")
        env = getattr(sc, "env", None)
        if env is not None:
            e = np.ascontiguousarray(env, np.float32)
            _check(lib().bsh_scene_set_env_map(self.h, _p(e), c_int(e.shape[1]), c_int(e.shape[0])),
                   "bsh_scene_set_env_map")
        others = getattr(sc, "other_lights", None)
        if others is not None and len(others):
            rows = np.asarray(others, np.float32).reshape(-1, 9)
            arr = (Light * rows.shape[0])()
            for k, r in enumerate(rows):
                arr[k].type = int(r[0])
                for c in range(3):
                    arr[k].v[c] = float(r[1 + c])
                    arr[k].e[c] = float(r[4 + c])
                arr[k].scale, arr[k].gamma = float(r[7]), float(r[8])
            _check(lib().bsh_scene_set_other_lights(self.h, arr, c_size_t(rows.shape[0])), "bsh_scene_set_other_lights")
        self.set_camera(sc.camera)
        bp = BvhParams(max_leaf, sah_bins, threads, width, _quant_flag(quantized))
        _check(lib().bsh_scene_commit(self.h, ctypes.byref(bp)), "bsh_scene_commit")

    def set_camera(self, cam13):
        c = camera_struct(cam13)
        _check(lib().bsh_scene_set_camera(self.h, ctypes.byref(c)), "bsh_scene_set_camera")

    def info(self):
        nodes, tris, nbytes = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        ms, depth, width, nb = ctypes.c_double(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib().bsh_scene_info(self.h, ctypes.byref(nodes), ctypes.byref(tris), ctypes.byref(nbytes),
                                    ctypes.byref(ms), ctypes.byref(depth), ctypes.byref(width), ctypes.byref(nb)),
               "bsh_scene_info")
        return {"nodes": nodes.value, "tris": tris.value, "bytes": nbytes.value, "build_ms": ms.value,
                "depth": depth.value, "width": width.value, "quantized": width.value == 4 and nb.value == 64,
                "node_bytes": nb.value}

    def diag_material(self, material, seed, frames, view, u, ea, eb):
        """Device Material::Sample / Evaluate / GetPdf of row `material` (bsh_diag_material): n x 12."""
        arrs = [np.ascontiguousarray(a, np.float32) for a in (frames, view, u, ea, eb)]
        n = arrs[0].shape[0]
        out = np.zeros((n, 12), np.float32)
        _check(lib().bsh_diag_material(self.h, c_int(material), c_uint32(seed), c_size_t(n), *[_p(a) for a in arrs],
                                       _p(out)), "bsh_diag_material")
        return out

    def diag_light_full(self, light, p, nrm, u, w):
        """Device Light::GetSample / GetPdf / GetLe of light `light` (area lights first, then the
        other lights): n x 16 = le(3), pdf, d(3), GetPdf(w), GetLe(w)(3), 0 x 5."""
        arrs = [np.ascontiguousarray(a, np.float32) for a in (p, nrm, u, w)]
        n = arrs[0].shape[0]
        out = np.zeros((n, 16), np.float32)
        _check(lib().bsh_diag_light_full(self.h, c_int(light), c_size_t(n), *[_p(a) for a in arrs], _p(out)),
               "bsh_diag_light_full")
        return out

    # ---- queries on device buffers ----
    def intersect_closest_dev(self, rays: DeviceArray, n: int, hits: DeviceArray, stream=None):
        _check(lib().bsh_intersect_closest(self.h, rays.ptr, c_size_t(n), hits.ptr, c_void_p(stream)),
               "bsh_intersect_closest")

    def intersect_any_dev(self, rays: DeviceArray, n: int, occ: DeviceArray, stream=None):
        _check(lib().bsh_intersect_any(self.h, rays.ptr, c_size_t(n), occ.ptr, c_void_p(stream)), "bsh_intersect_any")

    # ---- host conveniences ----
    def intersect(self, rays: np.ndarray, any_hit=False):
        """rays: n x 8 float32. Returns (prim int32[n], thit float32[n,3]) like the oracles."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        if n == 0:
            return np.zeros(0, np.int32), np.zeros((0, 3), np.float32)
        dr = self.ctx.to_device(rays)
        if any_hit:
            occ = self.ctx.alloc(n, np.int32)
            self.intersect_any_dev(dr, n, occ)
            self.ctx.sync()
            return occ.numpy(), np.zeros((n, 3), np.float32)
        hits = self.ctx.alloc((n, 4), np.float32)
        self.intersect_closest_dev(dr, n, hits)
        self.ctx.sync()
        h = hits.numpy()
        prim = h[:, 0].view(np.int32).copy()
        return prim, np.ascontiguousarray(h[:, 1:4])

    def intersect_count(self, rays: np.ndarray, any_hit=False):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        dr = self.ctx.to_device(rays)
        hits = self.ctx.alloc((n, 4), np.float32)
        vis = self.ctx.alloc((n, 2), np.uint32)
        _check(lib().bsh_intersect_count(self.h, dr.ptr, c_size_t(n), c_int(1 if any_hit else 0), hits.ptr, vis.ptr,
                                         c_void_p(None)), "bsh_intersect_count")
        self.ctx.sync()
        h = hits.numpy()
        return h[:, 0].view(np.int32).copy(), h[:, 1:4].copy(), vis.numpy()

    def render(self, width, height, img_grid, max_depth, seed, crop=None, spp_range=None, chunk_spp=0,
               bg=(0.0, 0.0, 0.0), per_sample=False, accum: DeviceArray | None = None, stream=None, download=True,
               count=False, samplers=("stratified", "stratified", "stratified")):
        """Returns (sum [h,w,3], sumsq [h,w,3], count [h,w], per_sample [h,w,s,3] or None).
        samplers: kinds of the image sampler (built with img_grid), the light and the BSDF sampler
        (built with 1), from SAMPLERS."""
        p = RenderParams()
        p.width, p.height = width, height
        if crop is not None:
            p.crop_x, p.crop_y, p.crop_w, p.crop_h = crop
        cw = crop[2] if crop else width
        ch = crop[3] if crop else height
        p.img_grid = img_grid
        _set_samplers(p, samplers)
        spp = num_samples(samplers[0], img_grid)
        sb, sc = (spp_range if spp_range else (0, spp))
        p.spp_begin, p.spp_count = sb, sc
        p.max_depth = max_depth
        p.seed = seed
        p.chunk_spp = chunk_spp
        for k in range(3):
            p.bg[k] = bg[k]
        p.flags = 4 if count else 0
        own = accum is None
        if own:
            accum = self.ctx.alloc((ch, cw, 8), np.float32)
        samp = self.ctx.alloc((ch, cw, sc, 3), np.float32) if per_sample else None
        _check(lib().bsh_render(self.h, ctypes.byref(p), accum.ptr, samp.ptr if samp else c_void_p(None),
                                c_void_p(stream)), "bsh_render")
        if not download:
            return None
        self.ctx.sync()
        a = accum.numpy()
        return a[..., 0:3], a[..., 4:7], a[..., 3], (samp.numpy() if samp else None)

    def bake_sh(self, probes, rays_per_probe, lmax, max_depth, seed, tmin=0.0, chunk_probes=0, radiance=False,
                count=False, coeffs: DeviceArray | None = None, download=True, quirk_free=False,
                bg=(0.0, 0.0, 0.0)):
        """Returns (coeffs [nprobes, (lmax+1)^2, 3], per-ray radiance [nprobes, N, 3] or None).
        quirk_free selects the unbiased estimator (BSH_RENDER_NO_QUIRKS) instead of GiTracer's;
        bg is the radiance of a probe ray that hits nothing (World::bgcolor_)."""
        probes = np.ascontiguousarray(probes, np.float32).reshape(-1, 3)
        n = probes.shape[0]
        dp = self.ctx.to_device(probes)
        nt = (lmax + 1) ** 2
        if coeffs is None:
            coeffs = self.ctx.alloc((n, nt, 3), np.float32)
        rad = self.ctx.alloc((n, rays_per_probe, 3), np.float32) if radiance else None
        p = ProbeParams()
        p.rays_per_probe, p.lmax, p.max_depth, p.seed = rays_per_probe, lmax, max_depth, seed
        p.chunk_probes, p.tmin = chunk_probes, tmin
        for k in range(3):
            p.bg[k] = bg[k]
        p.flags = (4 if count else 0) | (2 if quirk_free else 0)
        _check(lib().bsh_bake_sh(self.h, dp.ptr, c_size_t(n), ctypes.byref(p), coeffs.ptr,
                                 rad.ptr if rad else c_void_p(None), c_void_p(None)), "bsh_bake_sh")
        if not download:
            return None
        self.ctx.sync()
        return coeffs.numpy(), (rad.numpy() if rad else None)

    def bake_sh_guided(self, probes, rays_per_probe, lmax, max_depth, seed, guide: "Guide | None" = None,
                       tmin=0.0, chunk_probes=0, radiance=False, coeffs: DeviceArray | None = None,
                       records: DeviceArray | None = None, download=True, ray_dirs=False):
        """Quirk-free probe bake guided at every path vertex, the probe included (its rays are
        drawn from the uniform/guide mixture and weighted). Returns (coeffs, per-ray radiance or
        None, nrecords) when download, else nrecords; with ray_dirs, (coeffs, radiance, dirs
        [n, N, 4] = direction and projection weight, nrecords). probes may be a host array or a
        DeviceArray of shape [n, 3]."""
        if isinstance(probes, DeviceArray):
            dp = probes
            n = probes.shape[0]
        else:
            probes = np.ascontiguousarray(probes, np.float32).reshape(-1, 3)
            n = probes.shape[0]
            dp = self.ctx.to_device(probes)
        nt = (lmax + 1) ** 2
        if coeffs is None:
            coeffs = self.ctx.alloc((n, nt, 3), np.float32)
        rad = self.ctx.alloc((n, rays_per_probe, 3), np.float32) if radiance else None
        dirs = self.ctx.alloc((n, rays_per_probe, 4), np.float32) if ray_dirs else None
        p = ProbeParams()
        p.rays_per_probe, p.lmax, p.max_depth, p.seed = rays_per_probe, lmax, max_depth, seed
        p.chunk_probes, p.tmin = chunk_probes, tmin
        p.flags = 2
        nrec = c_size_t(0)
        cap = records.shape[0] if records is not None else 0
        _check(lib().bsh_bake_sh_guided(self.h, dp.ptr, c_size_t(n), ctypes.byref(p), guide.h if guide else None,
                                        coeffs.ptr, rad.ptr if rad else c_void_p(None),
                                        dirs.ptr if dirs else c_void_p(None),
                                        records.ptr if records is not None else None, c_size_t(cap),
                                        ctypes.byref(nrec), c_void_p(None)), "bsh_bake_sh_guided")
        if not download:
            return nrec.value
        self.ctx.sync()
        if ray_dirs:
            return coeffs.numpy(), (rad.numpy() if rad else None), dirs.numpy(), nrec.value
        return coeffs.numpy(), (rad.numpy() if rad else None), nrec.value

    def render_guided(self, width, height, img_grid, max_depth, seed, guide: "Guide | None" = None, crop=None,
                      spp_range=None, chunk_spp=0, accum: DeviceArray | None = None, records: DeviceArray | None = None,
                      download=True, accumulate=False):
        """Quirk-free render whose continuation samples the BSDF/guide mixture. Returns
        (sum, sumsq, count, nrecords) when download, else nrecords."""
        p = render_params(width, height, img_grid, max_depth, seed, crop=crop, spp_range=spp_range,
                          chunk_spp=chunk_spp, accumulate=accumulate)
        p.flags |= 2
        cw = crop[2] if crop else width
        ch = crop[3] if crop else height
        if accum is None:
            accum = self.ctx.alloc((ch, cw, 8), np.float32)
        nrec = c_size_t(0)
        cap = records.shape[0] if records is not None else 0
        _check(lib().bsh_render_guided(self.h, ctypes.byref(p), guide.h if guide else None, accum.ptr,
                                       records.ptr if records is not None else None, c_size_t(cap),
                                       ctypes.byref(nrec), c_void_p(None)), "bsh_render_guided")
        if not download:
            return nrec.value
        self.ctx.sync()
        a = accum.numpy()
        return a[..., 0:3], a[..., 4:7], a[..., 3], nrec.value

    def stats(self) -> dict:
        s = RenderStats()
        _check(lib().bsh_render_stats_get(self.h, ctypes.byref(s)), "bsh_render_stats_get")
        return {k: getattr(s, k) for k, _ in s._fields_}

    def diag_light(self, light, p, u, w):
        p, u, w = (np.ascontiguousarray(x, np.float32) for x in (p, u, w))
        out = np.zeros((p.shape[0], 8), np.float32)
        _check(lib().bsh_diag_light(self.h, c_int(light), c_size_t(p.shape[0]), _p(p), _p(u), _p(w), _p(out)),
               "bsh_diag_light")
        return out

    def close(self):
        if self.h:
            lib().bsh_scene_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
