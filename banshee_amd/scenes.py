"""Seeded procedural scenes in the flat layout the C-ABI takes.

Every generator uses numpy's PCG64 bit generator (``numpy.random.Generator(
numpy.random.PCG64(seed))``) so a (generator, seed, size) triple always yields the
same triangles.  Transforms are baked into vertices, so every bundle has the
identity transform; a reference ``Mesh`` built from the same arrays therefore sees
object space == world space (``Banshee/math/mathutils.h:129-157`` with identity
matrices).

Flat layout (all little-endian, C-contiguous):
    verts    float32 [V,3]    positions
    normals  float32 [N,3]    shading normals
    tri_v    int32   [T,3]    global vertex indices
    tri_n    int32   [T,3]    global normal indices
    tri_m    int32   [T]      material index (``Mesh::Face::m``, mesh.h:57-63)
    bundles  int32   [B,6]    tri_begin, tri_count, vert_begin, vert_count,
                              norm_begin, norm_count  (one reference Mesh each)
    materials float32 [M,8]   type, c0, c1, c2, eta, alpha, 0, 0
                              type 0 Lambert(kd=c)        bsdf/lambert.h:45-135
                              type 1 Microfacet(eta, ks=c, GGX(alpha))
                                                          bsdf/microfacet.h:53-77
                              type 2 Emissive(e=c)        material/emissive.h:45-99
    lights   int32   [L,2]    (bundle, shape index in bundle): one AreaLight per
                              emissive triangle, arealight.cpp:11-23
    camera   float32 [13]     eye3, at3, up3, znear, zfar, fovy, aspect
                              (PerscpectiveCamera ctor, perspective_camera.cpp:5-23)
The global primitive index of a triangle is its row in ``tri_v``; bundles are
concatenated in order, which is the order ``Bvh::Build`` numbers shapes
(bvh.cpp:26-51).
"""
from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, field

import numpy as np

LAMBERT, MICROFACET, EMISSIVE = 0, 1, 2
# material rows 3-7: OrenNayar(kd, kr = alpha), PerfectReflect / PerfectRefract(eta, ks, Fresnel when
# x != 0), Glass(eta, ks), MixedMaterial(eta) over the component rows alpha, x, y (-1: none)
ORENNAYAR, PERFECT_REFLECT, PERFECT_REFRACT, GLASS, MIXED = 3, 4, 5, 6, 7
# bsh_light types; a Scene's other_lights rows are (type, v xyz, e xyz, scale, gamma)
POINT_LIGHT, DIRECTIONAL_LIGHT, ENV_LIGHT, ENV_LIGHT_IS = 1, 2, 3, 4


@dataclass
class Scene:
    verts: np.ndarray
    normals: np.ndarray
    tri_v: np.ndarray
    tri_n: np.ndarray
    tri_m: np.ndarray
    bundles: np.ndarray
    materials: np.ndarray
    lights: np.ndarray
    camera: np.ndarray
    name: str = ""
    meta: dict = field(default_factory=dict)
    # lights after the area lights (PointLight, DirectionalLight, EnvironmentLight[Is]) as rows
    # (type, v xyz, e xyz, scale, gamma), and the environment lights' lat-long map [h, w, 3]
    other_lights: np.ndarray | None = None
    env: np.ndarray | None = None
    # per bundle (ShapeBundle::SetTransform): row-major m then minv, [nbundles, 32]; None: identities
    transforms: np.ndarray | None = None
    # texture coordinates: (u, v) pairs and three indices per triangle; texture maps [h, w, 3];
    # per material row (colour map, kr map), -1 for none
    uvs: np.ndarray | None = None
    tri_uv: np.ndarray | None = None
    textures: list | None = None
    mat_maps: np.ndarray | None = None
    # per material row the texture map used as its normal map (NormalMapping), -1 for none
    normal_maps: np.ndarray | None = None

    @property
    def ntris(self) -> int:
        return int(self.tri_v.shape[0])

    def digest(self) -> str:
        h = hashlib.sha256()
        for a in (self.verts, self.normals, self.tri_v, self.tri_n, self.tri_m, self.bundles,
                  self.materials, self.lights, self.camera):
            h.update(np.ascontiguousarray(a).tobytes())
        for a in (self.other_lights, self.env, self.transforms, self.uvs, self.tri_uv, self.mat_maps,
                  self.normal_maps):
            if a is not None:
                h.update(np.ascontiguousarray(a, np.float32).tobytes())
        for t in self.textures or ():
            h.update(np.ascontiguousarray(t, np.float32).tobytes())
        return h.hexdigest()[:16]

    def with_camera(self, camera) -> "Scene":
        s = Scene(**{k: getattr(self, k) for k in self.__dataclass_fields__})
        s.camera = np.asarray(camera, np.float32)
        return s


class _Builder:
    def __init__(self):
        self.v, self.n, self.tv, self.tn, self.tm, self.b = [], [], [], [], [], []
        self.nv = 0
        self.nt = 0
        self.mats = []
        self.lights = []

    def material(self, type_, c, eta=0.0, alpha=0.0, x=0.0, y=0.0) -> int:
        self.mats.append([type_, c[0], c[1], c[2], eta, alpha, x, y])
        return len(self.mats) - 1

    def bundle(self, verts, normals, tris, mat) -> int:
        verts = np.asarray(verts, np.float32).reshape(-1, 3)
        normals = np.asarray(normals, np.float32).reshape(-1, 3)
        assert verts.shape == normals.shape
        tris = np.asarray(tris, np.int64).reshape(-1, 3)
        mats = np.broadcast_to(np.asarray(mat, np.int32), (tris.shape[0],))
        vb = self.nv
        self.v.append(verts)
        self.n.append(normals)
        self.tv.append((tris + vb).astype(np.int32))
        self.tn.append((tris + vb).astype(np.int32))
        self.tm.append(mats.astype(np.int32))
        self.b.append([self.nt, tris.shape[0], vb, verts.shape[0], vb, verts.shape[0]])
        self.nv += verts.shape[0]
        self.nt += tris.shape[0]
        return len(self.b) - 1

    def quad(self, corners, normal, mat) -> int:
        """corners: 4 points in order around the quad; two triangles (0,1,2), (0,2,3)."""
        c = np.asarray(corners, np.float32)
        n = np.tile(np.asarray(normal, np.float32), (4, 1))
        return self.bundle(c, n, [[0, 1, 2], [0, 2, 3]], mat)

    def area_lights(self, bundle: int):
        for s in range(self.b[bundle][1]):
            self.lights.append([bundle, s])

    def build(self, camera, name, meta=None, other_lights=None, env=None) -> Scene:
        return Scene(
            other_lights=None if other_lights is None else np.asarray(other_lights, np.float32).reshape(-1, 9),
            env=None if env is None else np.ascontiguousarray(env, np.float32),
            verts=np.ascontiguousarray(np.concatenate(self.v), np.float32),
            normals=np.ascontiguousarray(np.concatenate(self.n), np.float32),
            tri_v=np.ascontiguousarray(np.concatenate(self.tv), np.int32),
            tri_n=np.ascontiguousarray(np.concatenate(self.tn), np.int32),
            tri_m=np.ascontiguousarray(np.concatenate(self.tm), np.int32),
            bundles=np.ascontiguousarray(np.asarray(self.b, np.int32).reshape(-1, 6)),
            materials=np.ascontiguousarray(np.asarray(self.mats, np.float32).reshape(-1, 8)),
            lights=np.ascontiguousarray(np.asarray(self.lights, np.int32).reshape(-1, 2)),
            camera=np.asarray(camera, np.float32),
            name=name,
            meta=meta or {},
        )


def box_mesh(lo, hi):
    """Axis-aligned box: 24 vertices (flat faces), 12 triangles, outward normals."""
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [
        ([[x0, y0, z0], [x0, y0, z1], [x0, y1, z1], [x0, y1, z0]], [-1, 0, 0]),
        ([[x1, y0, z0], [x1, y1, z0], [x1, y1, z1], [x1, y0, z1]], [1, 0, 0]),
        ([[x0, y0, z0], [x1, y0, z0], [x1, y0, z1], [x0, y0, z1]], [0, -1, 0]),
        ([[x0, y1, z0], [x0, y1, z1], [x1, y1, z1], [x1, y1, z0]], [0, 1, 0]),
        ([[x0, y0, z0], [x0, y1, z0], [x1, y1, z0], [x1, y0, z0]], [0, 0, -1]),
        ([[x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], [0, 0, 1]),
    ]
    v, n, t = [], [], []
    for k, (c, nn) in enumerate(faces):
        v.extend(c)
        n.extend([nn] * 4)
        b = 4 * k
        t.extend([[b, b + 1, b + 2], [b, b + 2, b + 3]])
    return np.asarray(v, np.float32), np.asarray(n, np.float32), np.asarray(t, np.int32)


def cornell_box(width=256, height=256, seed=0, light_radiance=15.0) -> Scene:
    """Procedural Cornell-style box: 5 walls + 2 axis-aligned boxes (34 triangles),
    Lambert albedo U[0.2,0.8]^3 per object, one 0.5x0.5 quad area light of
    radiance 15 just under the ceiling (2 emissive triangles => 2 AreaLights)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    B = _Builder()

    def albedo():
        return rng.uniform(0.2, 0.8, 3).astype(np.float32)

    walls = [
        ([[0, 0, 0], [2, 0, 0], [2, 0, 2], [0, 0, 2]], [0, 1, 0]),      # floor
        ([[0, 2, 0], [0, 2, 2], [2, 2, 2], [2, 2, 0]], [0, -1, 0]),     # ceiling
        ([[0, 0, 0], [0, 2, 0], [2, 2, 0], [2, 0, 0]], [0, 0, 1]),      # back
        ([[0, 0, 0], [0, 0, 2], [0, 2, 2], [0, 2, 0]], [1, 0, 0]),      # left
        ([[2, 0, 0], [2, 2, 0], [2, 2, 2], [2, 0, 2]], [-1, 0, 0]),     # right
    ]
    for c, n in walls:
        B.quad(c, n, B.material(LAMBERT, albedo()))
    for lo, hi in (((0.3, 0.0, 0.9), (0.9, 0.6, 1.5)), ((1.1, 0.0, 0.3), (1.7, 1.2, 0.9))):
        v, n, t = box_mesh(lo, hi)
        B.bundle(v, n, t, B.material(LAMBERT, albedo()))
    em = B.material(EMISSIVE, (light_radiance,) * 3)
    lb = B.quad([[0.75, 1.98, 0.75], [1.25, 1.98, 0.75], [1.25, 1.98, 1.25], [0.75, 1.98, 1.25]],
                [0, -1, 0], em)
    B.area_lights(lb)
    cam = [1, 1, 3.5, 1, 1, 0, 0, 1, 0, 0.01, 10000.0, math.pi / 4, width / height]
    return B.build(cam, "cornell", {"seed": seed})


def icosphere(subdiv: int):
    """Unit icosphere by repeated midpoint subdivision of an icosahedron."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    verts = [np.asarray(p, np.float64) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    return np.asarray(verts, np.float64), np.asarray(f, np.int32)


def _value_noise(rng, x, z, extent, octaves=4, base_cells=8):
    """Sum of octaves of smoothstep-interpolated lattice noise in [-1, 1]."""
    h = np.zeros_like(x, np.float64)
    amp, total = 1.0, 0.0
    cells = base_cells
    for _ in range(octaves):
        lat = rng.uniform(-1.0, 1.0, (cells + 1, cells + 1))
        u = (x + extent) / (2 * extent) * cells
        w = (z + extent) / (2 * extent) * cells
        i0 = np.clip(np.floor(u).astype(np.int64), 0, cells - 1)
        j0 = np.clip(np.floor(w).astype(np.int64), 0, cells - 1)
        fu = u - i0
        fw = w - j0
        su = fu * fu * (3 - 2 * fu)
        sw = fw * fw * (3 - 2 * fw)
        a = lat[i0, j0] * (1 - su) + lat[i0 + 1, j0] * su
        b = lat[i0, j0 + 1] * (1 - su) + lat[i0 + 1, j0 + 1] * su
        h += amp * (a * (1 - sw) + b * sw)
        total += amp
        amp *= 0.5
        cells *= 2
    return h / total


def _random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - w * z)
    R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y)
    R[:, 2, 1] = 2 * (y * z + w * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def terrain_scene(grid=512, spheres=2000, subdiv=3, seed=1, width=1920, height=1080,
                  extent=50.0, amplitude=3.0, light_radiance=80.0) -> Scene:
    """Seeded heightfield (grid x grid quads over [-extent, extent]^2, value-noise
    displaced, amplitude 3) plus `spheres` random-scaled/rotated icospheres
    (`subdiv` subdivisions) resting above it. Per sphere: 50% Lambert (albedo
    U[0.2,0.8]^3), 50% Microfacet(eta=1.5, ks=U[0.2,0.8]^3, GGX alpha=U[0.05,1]).
    Two 4x4 quad area lights (4 light triangles). Each light triangle has area 8, so its
    area pdf 1/8 clears the reference's MINPDF cut (0.05, ditracer.cpp:103); a triangle
    larger than 20 would never be light-sampled."""
    rng = np.random.Generator(np.random.PCG64(seed))
    B = _Builder()
    # heightfield
    g = grid
    xs = np.linspace(-extent, extent, g + 1)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    H = amplitude * _value_noise(rng, X, Z, extent)
    step = 2 * extent / g
    dHx = np.gradient(H, step, axis=0)
    dHz = np.gradient(H, step, axis=1)
    N = np.stack([-dHx, np.ones_like(H), -dHz], -1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    V = np.stack([X, H, Z], -1).reshape(-1, 3)
    idx = np.arange((g + 1) * (g + 1)).reshape(g + 1, g + 1)
    a = idx[:-1, :-1].ravel()
    b = idx[1:, :-1].ravel()
    c = idx[1:, 1:].ravel()
    d = idx[:-1, 1:].ravel()
    tris = np.empty((2 * g * g, 3), np.int64)
    tris[0::2] = np.stack([a, d, c], -1)
    tris[1::2] = np.stack([a, c, b], -1)
    B.bundle(V, N.reshape(-1, 3), tris, B.material(LAMBERT, rng.uniform(0.2, 0.8, 3)))
    # icospheres
    sv, sf = icosphere(subdiv)
    if spheres > 0:
        scale = rng.uniform(0.3, 2.0, spheres)
        R = _random_rotations(rng, spheres)
        px = rng.uniform(-0.9 * extent, 0.9 * extent, spheres)
        pz = rng.uniform(-0.9 * extent, 0.9 * extent, spheres)
        lift = rng.uniform(0.0, 5.0, spheres)
        is_ggx = rng.uniform(0.0, 1.0, spheres) < 0.5
        col = rng.uniform(0.2, 0.8, (spheres, 3))
        alpha = rng.uniform(0.05, 1.0, spheres)
        # ground height under each sphere from the same bilinear mesh interpolation
        gi = np.clip(((px + extent) / step).astype(np.int64), 0, g - 1)
        gj = np.clip(((pz + extent) / step).astype(np.int64), 0, g - 1)
        ground = np.maximum.reduce([H[gi, gj], H[gi + 1, gj], H[gi, gj + 1], H[gi + 1, gj + 1]])
        py = ground + scale + lift + 0.05
        rot_n = np.einsum("nij,vj->nvi", R, sv)
        pos = rot_n * scale[:, None, None] + np.stack([px, py, pz], -1)[:, None, :]
        for k in range(spheres):
            if is_ggx[k]:
                m = B.material(MICROFACET, col[k], 1.5, alpha[k])
            else:
                m = B.material(LAMBERT, col[k])
            B.bundle(pos[k], rot_n[k], sf, m)
    # two quad lights facing down
    em = B.material(EMISSIVE, (light_radiance,) * 3)
    for cx, cy, cz, s in ((-15.0, 22.0, -10.0, 4.0), (20.0, 20.0, 15.0, 4.0)):
        h = s / 2
        lb = B.quad([[cx - h, cy, cz - h], [cx + h, cy, cz - h], [cx + h, cy, cz + h], [cx - h, cy, cz + h]],
                    [0, -1, 0], em)
        B.area_lights(lb)
    cam = [0, 30, 75, 0, 0, 0, 0, 1, 0, 0.01, 10000.0, math.pi / 4, width / height]
    return B.build(cam, f"terrain{grid}_{spheres}", {"seed": seed, "grid": grid, "spheres": spheres,
                                                    "subdiv": subdiv, "extent": extent})


def terrain_for_tris(target_tris: int, seed: int, subdiv=3, width=1920, height=1080) -> Scene:
    """The terrain generator scaled to about `target_tris` triangles: the heightfield
    keeps the cfg1 share (1/6 of the triangles), the rest are icospheres."""
    per_sphere = 20 * 4 ** subdiv
    hf = max(8, target_tris // 6)
    grid = max(4, int(round(math.sqrt(hf / 2))))
    rest = max(0, target_tris - 2 * grid * grid)
    spheres = max(1, rest // per_sphere)
    return terrain_scene(grid=grid, spheres=spheres, subdiv=subdiv, seed=seed, width=width, height=height)


def batch_jobs(n=64, seed=100, min_tris=1_000, max_tris=2_000_000, min_res=128, max_res=1024, min_grid=2,
               max_grid=8, min_depth=4, max_depth=12):
    """The batched workload: n independent jobs of the terrain generator with triangle counts
    log-uniform in [min_tris, max_tris], square resolutions in [min_res, max_res] (multiples
    of 16), spp = grid^2 with grid in [min_grid, max_grid] (4..64 spp) and depth in
    [min_depth, max_depth]; job k uses scene seed seed + k. Returns a list of dicts."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(n):
        tris = int(round(math.exp(rng.uniform(math.log(min_tris), math.log(max_tris)))))
        res = int(rng.integers(min_res // 16, max_res // 16 + 1)) * 16
        grid = int(rng.integers(min_grid, max_grid + 1))
        depth = int(rng.integers(min_depth, max_depth + 1))
        out.append({"tris": tris, "res": res, "grid": grid, "depth": depth, "seed": seed + k})
    return out


def probe_grid(scene: Scene, n=16, margin=0.05, lift=0.5):
    """n^3 probe positions on a regular grid inside the scene's bounding box. For a terrain
    scene, a probe below the heightfield is moved up to `lift` above the highest corner of the
    heightfield cell under it, so that every probe sits above the ground."""
    lo = scene.verts.min(0)
    hi = scene.verts.max(0)
    ext = hi - lo
    t = (np.arange(n) + 0.5) / n
    xs = lo[0] + ext[0] * (margin + (1 - 2 * margin) * t)
    ys = lo[1] + ext[1] * (margin + (1 - 2 * margin) * t)
    zs = lo[2] + ext[2] * (margin + (1 - 2 * margin) * t)
    P = np.stack(np.meshgrid(xs, ys, zs, indexing="ij"), -1).reshape(-1, 3)
    if "grid" in scene.meta and "extent" in scene.meta:
        g, e = scene.meta["grid"], scene.meta["extent"]
        b = scene.bundles[0]                      # the heightfield, (g + 1)^2 vertices in x-major order
        H = scene.verts[b[2]:b[2] + b[3], 1].reshape(g + 1, g + 1)
        step = 2 * e / g
        i = np.clip(((P[:, 0] + e) / step).astype(np.int64), 0, g - 1)
        j = np.clip(((P[:, 2] + e) / step).astype(np.int64), 0, g - 1)
        top = np.maximum.reduce([H[i, j], H[i + 1, j], H[i, j + 1], H[i + 1, j + 1]])
        P[:, 1] = np.maximum(P[:, 1], top + lift)
    return P.astype(np.float32)


def far_speck(distance=1.0e4, size=1.0e-3) -> Scene:
    """An almost empty world: one Lambert triangle `size` across at `distance` from the origin.
    From within a few units of the origin it subtends about (size / distance)^2 sr, so rays cast
    from there all miss, and what a ray returns is the background alone."""
    B = _Builder()
    d, s = float(distance), float(size)
    B.bundle([[d, d, d], [d + s, d, d], [d, d + s, d]], np.tile([0, 0, 1], (3, 1)), [[0, 1, 2]],
             B.material(LAMBERT, (0.5, 0.5, 0.5)))
    cam = [0.0, 0.0, 0.0, 0.0, 0.0, -1.0, 0, 1, 0, 0.01, 10000.0, 1.0, 1.0]
    return B.build(cam, "far_speck")


def lit_floor(albedo=0.5, light_radiance=10.0, target=(0.3, 0.0, 0.2), fovy=1e-3, width=64, height=64) -> Scene:
    """A Lambert floor (y = 0, 20 x 20) under one emissive triangle facing down at y = 1.5,
    seen through a very narrow camera aimed at `target`, so every camera sample lands within
    a few millimetres of that point. With depth 1 the image estimates the reflected direct
    light there, which has a closed form (quadrature over the emitter)."""
    B = _Builder()
    B.quad([[-10, 0, -10], [-10, 0, 10], [10, 0, 10], [10, 0, -10]], [0, 1, 0],
           B.material(LAMBERT, (albedo, albedo, albedo)))
    em = B.material(EMISSIVE, (light_radiance,) * 3)
    lt = np.asarray([[-0.5, 1.5, -0.5], [0.5, 1.5, -0.5], [0.1, 1.5, 0.6]], np.float32)
    lb = B.bundle(lt, np.tile([0, -1, 0], (3, 1)), [[0, 1, 2]], em)
    B.area_lights(lb)
    cam = [0.0, 1.0, 3.0, target[0], target[1], target[2], 0, 1, 0, 0.01, 10000.0, fovy, width / height]
    return B.build(cam, "lit_floor", {"light": lt, "albedo": albedo, "radiance": light_radiance})


def lit_floor_many(n_lights=40, seed=5, albedo=0.5, target=(0.3, 0.0, 0.2), fovy=1e-3, width=64, height=64) -> Scene:
    """The lit_floor set-up under n_lights emissive triangles facing down at heights 1.2-2.0, each its
    own bundle and material with radiance U[1, 30] per channel, so the lights' powers differ. With
    depth 1 the image estimates the sum of the triangles' closed-form direct light at the target."""
    rng = np.random.Generator(np.random.PCG64(seed))
    B = _Builder()
    B.quad([[-10, 0, -10], [-10, 0, 10], [10, 0, 10], [10, 0, -10]], [0, 1, 0],
           B.material(LAMBERT, (albedo, albedo, albedo)))
    tris, rad = [], []
    for _ in range(n_lights):
        c = np.array([rng.uniform(-2, 2), rng.uniform(1.2, 2.0), rng.uniform(-2, 2)])
        s = rng.uniform(0.1, 0.4)
        a = rng.uniform(0, 2 * np.pi)
        # corners ordered so that the triangle's normal (p2 - p1) x (p3 - p1) faces down
        lt = np.array([c + s * np.array([np.cos(a + k * 2.0944), 0, -np.sin(a + k * 2.0944)]) for k in range(3)],
                      np.float32)[::-1]
        e = rng.uniform(1, 30, 3)
        lb = B.bundle(lt, np.tile([0, -1, 0], (3, 1)), [[0, 1, 2]], B.material(EMISSIVE, e))
        B.area_lights(lb)
        tris.append(lt)
        rad.append(np.asarray(e, np.float32))
    cam = [0.0, 1.0, 3.0, target[0], target[1], target[2], 0, 1, 0, 0.01, 10000.0, fovy, width / height]
    return B.build(cam, "lit_floor_many", {"lights": np.stack(tris), "radiance": np.stack(rad), "albedo": albedo})


def _grid_quad(B, origin, du, dv, nu, nv, normal, mat, keep=None):
    """A planar quad origin + s du + t dv, s, t in [0,1], split into nu x nv cells (2 triangles
    each); keep[nu, nv] drops the cells where it is False (holes)."""
    o = np.asarray(origin, np.float64)
    du = np.asarray(du, np.float64)
    dv = np.asarray(dv, np.float64)
    s = np.linspace(0, 1, nu + 1)
    t = np.linspace(0, 1, nv + 1)
    S, T = np.meshgrid(s, t, indexing="ij")
    V = (o + S[..., None] * du + T[..., None] * dv).reshape(-1, 3)
    idx = np.arange((nu + 1) * (nv + 1)).reshape(nu + 1, nv + 1)
    a, b = idx[:-1, :-1], idx[1:, :-1]
    c, d = idx[1:, 1:], idx[:-1, 1:]
    m = np.ones((nu, nv), bool) if keep is None else keep
    tris = np.concatenate([np.stack([a[m], b[m], c[m]], -1), np.stack([a[m], c[m], d[m]], -1)])
    N = np.tile(np.asarray(normal, np.float32), (V.shape[0], 1))
    return B.bundle(V, N, tris, mat)


def occluded_room(target_tris=500_000, seed=3, width=1024, height=1024, hole_fraction=0.05, holes=4,
                  light_radiance=400.0, subdiv=3) -> Scene:
    """The guiding workload: an 8 x 3 x 6 room split at z = 3 by a wall of 64 x 24 cells with
    `holes` rectangular openings, together `hole_fraction` of its area, at random places. The only light (a 1.2 x 1.2 quad) is in the far
    half and faces the far wall; the camera stands in the near half with its back to the
    holed wall, so everything it sees is lit indirectly, through the holes. Walls, floor and ceiling are finely tessellated Lambert surfaces
    (albedo U[0.5,0.8]^3); the rest of the triangle budget is icospheres resting on the floor
    in both halves, Microfacet(eta=1.5, ks=U[0.5,0.9]^3, GGX alpha=U[0.1,0.6])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    B = _Builder()
    X, Y, Z, ZW = 8.0, 3.0, 6.0, 3.0

    def lam():
        return B.material(LAMBERT, rng.uniform(0.5, 0.8, 3))
    # room shell, tessellated (~16k triangles)
    _grid_quad(B, (0, 0, 0), (X, 0, 0), (0, 0, Z), 64, 48, (0, 1, 0), lam())          # floor
    _grid_quad(B, (0, Y, 0), (0, 0, Z), (X, 0, 0), 48, 64, (0, -1, 0), lam())         # ceiling
    _grid_quad(B, (0, 0, 0), (0, Y, 0), (X, 0, 0), 24, 64, (0, 0, 1), lam())          # near wall
    _grid_quad(B, (0, 0, Z), (X, 0, 0), (0, Y, 0), 64, 24, (0, 0, -1), lam())         # far wall
    _grid_quad(B, (0, 0, 0), (0, 0, Z), (0, Y, 0), 48, 24, (1, 0, 0), lam())          # left
    _grid_quad(B, (X, 0, 0), (0, Y, 0), (0, 0, Z), 24, 48, (-1, 0, 0), lam())         # right
    keep = np.ones((64, 24), bool)
    cells = int(round(hole_fraction * 64 * 24))
    per = [cells // holes + (1 if k < cells % holes else 0) for k in range(holes)]
    for c in per:
        # a near-square opening of c cells (w x h, w >= h), placed where it does not overlap
        h = max(1, int(math.floor(math.sqrt(c))))
        w = int(math.ceil(c / h))
        for _ in range(1000):
            x0 = int(rng.integers(1, 64 - w))
            y0 = int(rng.integers(1, 24 - h))
            if keep[x0 - 1:x0 + w + 1, y0 - 1:y0 + h + 1].all():
                break
        n = 0
        for i in range(w):
            for j in range(h):
                if n < c:
                    keep[x0 + i, y0 + j] = False
                    n += 1
    _grid_quad(B, (0, 0, ZW), (X, 0, 0), (0, Y, 0), 64, 24, (0, 0, -1), lam(), keep)  # dividing wall
    # icospheres on the floor, clear of the walls
    sv, sf = icosphere(subdiv)
    per = sf.shape[0]
    shell = B.nt
    n = max(0, (target_tris - shell - 2) // per)
    if n:
        r = rng.uniform(0.1, 0.35, n)
        half = rng.uniform(0, 1, n) < 0.5
        px = rng.uniform(0.4, X - 0.4, n)
        pz = np.where(half, rng.uniform(0.4, ZW - 0.4, n), rng.uniform(ZW + 0.4, Z - 0.4, n))
        R = _random_rotations(rng, n)
        col = rng.uniform(0.5, 0.9, (n, 3))
        alpha = rng.uniform(0.1, 0.6, n)
        rot_n = np.einsum("nij,vj->nvi", R, sv)
        pos = rot_n * r[:, None, None] + np.stack([px, r + 0.001, pz], -1)[:, None, :]
        for k in range(n):
            B.bundle(pos[k], rot_n[k], sf, B.material(MICROFACET, col[k], 1.5, alpha[k]))
    # the light faces the far wall: no point of the near half is in front of it, so the near
    # half receives only light that has bounced in the far half and come through the holes
    em = B.material(EMISSIVE, (light_radiance,) * 3)
    cx, cy, cz, h = 4.0, 1.5, 4.2, 0.6
    lb = B.quad([[cx - h, cy - h, cz], [cx + h, cy - h, cz], [cx + h, cy + h, cz], [cx - h, cy + h, cz]],
                [0, 0, 1], em)
    B.area_lights(lb)
    # in the near half, backed up to the holed wall and looking away from it
    cam = [4.0, 1.8, 2.9, 3.6, 0.9, 0.0, 0, 1, 0, 0.01, 10000.0, math.pi / 2.5, width / height]
    return B.build(cam, "occluded_room", {"seed": seed, "openings": holes, "hole_cells": int((~keep).sum()),
                                          "spheres": n})


def env_map(width=64, height=32, seed=0, sun=(0.3, 0.25), sun_radiance=40.0) -> np.ndarray:
    """A seeded lat-long sky [height, width, 3]: a bluish gradient from the zenith (row 0) to the
    horizon, a darker ground half, a bright sun disc at (u, v) = sun, and mild value noise. Every
    texel is positive."""
    rng = np.random.Generator(np.random.PCG64(seed))
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    U, V = np.meshgrid(u, v)
    sky = np.stack([0.35 + 0.4 * V, 0.5 + 0.35 * V, 0.9 - 0.2 * V], -1)
    ground = np.stack([0.25 + 0 * V, 0.2 + 0 * V, 0.15 + 0 * V], -1)
    img = np.where((V < 0.5)[..., None], sky, ground)
    du = np.minimum(np.abs(U - sun[0]), 1 - np.abs(U - sun[0]))
    d2 = (du * 2) ** 2 + (V - sun[1]) ** 2
    img = img + sun_radiance * np.exp(-d2 / 0.002)[..., None] * np.array([1.0, 0.9, 0.7])
    img = img * (1 + 0.1 * rng.uniform(-1, 1, (height, width, 1)))
    return np.ascontiguousarray(img, np.float32)


def open_scene(env_kind=ENV_LIGHT, width=24, height=24, seed=4) -> Scene:
    """Every light type of the reference in one open scene: a 12 x 12 Lambert floor, a Lambert box
    and a GGX box, a small emissive quad (2 area lights), a PointLight, a DirectionalLight and an
    environment light of kind env_kind (EnvironmentLight or EnvironmentLightIs, scale 0.8, gamma
    2.2) over env_map(); the camera sees floor, boxes and sky."""
    rng = np.random.Generator(np.random.PCG64(seed))
    B = _Builder()
    B.quad([[-6, 0, -6], [-6, 0, 6], [6, 0, 6], [6, 0, -6]], [0, 1, 0], B.material(LAMBERT, (0.6, 0.55, 0.5)))
    for lo, hi, mat in (([-1.5, 0, -1.0], [-0.5, 1.2, 0.0], B.material(LAMBERT, rng.uniform(0.2, 0.8, 3))),
                        ([0.4, 0, -0.4], [1.4, 0.8, 0.6], B.material(MICROFACET, rng.uniform(0.2, 0.8, 3), 1.5, 0.3))):
        v, n, t = box_mesh(lo, hi)
        B.bundle(v, n, t, mat)
    em = B.material(EMISSIVE, (6.0, 5.5, 5.0))
    lq = np.asarray([[-0.4, 2.5, 1.2], [0.4, 2.5, 1.2], [0.4, 2.5, 1.8], [-0.4, 2.5, 1.8]], np.float32)
    lb = B.bundle(lq, np.tile([0, -1, 0], (4, 1)), [[0, 2, 1], [0, 3, 2]], em)
    B.area_lights(lb)
    others = [[POINT_LIGHT, 1.5, 1.8, 1.5, 3.0, 2.5, 2.0, 0, 0],
              [DIRECTIONAL_LIGHT, -0.3, -1.0, -0.4, 0.9, 0.85, 0.8, 0, 0],
              [env_kind, 0, 0, 0, 0, 0, 0, 0.8, 2.2]]
    cam = [0.0, 2.2, 6.0, 0.0, 0.6, 0.0, 0, 1, 0, 0.01, 10000.0, 0.9, width / height]
    return B.build(cam, "open_scene", {"env_kind": env_kind}, other_lights=others, env=env_map())


def floor_under(light_row, env=None, albedo=0.5, target=(0.3, 0.0, 0.2), fovy=1e-3, width=32, height=32) -> Scene:
    """A 40 x 40 Lambert floor lit by one non-area light (a bsh_light row), seen through a very
    narrow camera aimed at `target`: with depth 1 the image estimates the light reflected there,
    which has a closed form (point: albedo/pi e cos/d^2; directional: albedo/pi e cos; a constant
    environment of radiance c: albedo c)."""
    B = _Builder()
    B.quad([[-20, 0, -20], [-20, 0, 20], [20, 0, 20], [20, 0, -20]], [0, 1, 0],
           B.material(LAMBERT, (albedo, albedo, albedo)))
    cam = [0.0, 1.0, 3.0, target[0], target[1], target[2], 0, 1, 0, 0.01, 10000.0, fovy, width / height]
    return B.build(cam, "floor_under", {"albedo": albedo, "target": np.asarray(target)},
                   other_lights=[light_row], env=env)


def material_scene(width=24, height=24) -> Scene:
    """A closed box lit by a ceiling quad, holding every material the reference offers beyond
    Lambert and Microfacet: Oren-Nayar walls (two roughnesses), a glass box, a mirror panel
    (PerfectReflect without Fresnel), a refracting panel (PerfectRefract with Fresnel), a
    MixedMaterial floor (Lambert + Microfacet + PerfectReflect) and a MixedMaterial back wall
    (Oren-Nayar + PerfectRefract)."""
    B = _Builder()
    lam = B.material(LAMBERT, (0.6, 0.6, 0.55))
    on_a = B.material(ORENNAYAR, (0.7, 0.3, 0.25), alpha=0.4)
    on_b = B.material(ORENNAYAR, (0.25, 0.6, 0.3), alpha=1.2)
    ggx = B.material(MICROFACET, (0.7, 0.65, 0.5), 1.5, 0.25)
    mirror = B.material(PERFECT_REFLECT, (0.9, 0.9, 0.9), 1.5, x=0.0)
    refl_f = B.material(PERFECT_REFLECT, (0.8, 0.8, 0.8), 1.5, x=1.0)
    refr = B.material(PERFECT_REFRACT, (0.9, 0.95, 0.9), 1.4, x=1.0)
    glass = B.material(GLASS, (0.95, 0.95, 0.95), 1.5)
    mix_floor = B.material(MIXED, (0, 0, 0), 1.6, alpha=lam, x=ggx, y=refl_f)
    mix_back = B.material(MIXED, (0, 0, 0), 1.3, alpha=on_b, x=refr, y=-1)
    em = B.material(EMISSIVE, (12.0, 11.0, 9.0))
    B.quad([[0, 0, 0], [2, 0, 0], [2, 0, 2], [0, 0, 2]], [0, 1, 0], mix_floor)                 # floor
    B.quad([[0, 2, 0], [0, 2, 2], [2, 2, 2], [2, 2, 0]], [0, -1, 0], lam)                   # ceiling
    B.quad([[0, 0, 0], [0, 2, 0], [2, 2, 0], [2, 0, 0]], [0, 0, 1], mix_back)               # back
    B.quad([[0, 0, 0], [0, 0, 2], [0, 2, 2], [0, 2, 0]], [1, 0, 0], on_a)                   # left
    B.quad([[2, 0, 0], [2, 2, 0], [2, 2, 2], [2, 0, 2]], [-1, 0, 0], on_b)                  # right
    for lo, hi, mat in (([0.3, 0.0, 0.4], [0.9, 0.7, 1.0], glass),):
        v, n, t = box_mesh(lo, hi)
        B.bundle(v, n, t, mat)
    B.quad([[1.2, 0.05, 0.3], [1.8, 0.05, 0.3], [1.8, 1.0, 0.3], [1.2, 1.0, 0.3]], [0, 0, 1], mirror)
    B.quad([[1.1, 0.05, 1.2], [1.7, 0.05, 1.2], [1.7, 0.8, 1.2], [1.1, 0.8, 1.2]], [0, 0, 1], refr)
    lq = np.asarray([[0.75, 1.98, 0.75], [1.25, 1.98, 0.75], [1.25, 1.98, 1.25], [0.75, 1.98, 1.25]], np.float32)
    lb = B.bundle(lq, np.tile([0, -1, 0], (4, 1)), [[0, 1, 2], [0, 2, 3]], em)
    B.area_lights(lb)
    cam = [1.0, 1.0, 3.5, 1.0, 0.9, 0.0, 0, 1, 0, 0.01, 10000.0, 0.785398, width / height]
    return B.build(cam, "material_scene", {"rows": {"lambert": lam, "orennayar_a": on_a, "orennayar_b": on_b,
                                                    "microfacet": ggx, "mirror": mirror, "reflect_fresnel": refl_f,
                                                    "refract": refr, "glass": glass, "mixed_floor": mix_floor,
                                                    "mixed_back": mix_back}})


def transformed_cornell(width=64, height=64) -> Scene:
    """The Cornell box with its two inner boxes and its light moved by per-bundle transforms
    (ShapeBundle::SetTransform): a rotated, non-uniformly scaled and translated tall box, a
    mirrored short box (negative determinant) and a light quad scaled by 1.5 in x (its area pdf
    then differs between FillSample and GetShapeSurfaceArea)."""
    sc = cornell_box(width, height)
    nb = sc.bundles.shape[0]
    xf = np.tile(np.concatenate([np.eye(4).ravel(), np.eye(4).ravel()]), (nb, 1))

    def pair(m):
        return np.concatenate([m.ravel(), np.linalg.inv(m).ravel()])

    def rot_y(a):
        c, s_ = np.cos(a), np.sin(a)
        return np.array([[c, 0, s_, 0], [0, 1, 0, 0], [-s_, 0, c, 0], [0, 0, 0, 1]])

    def about(m, cx, cz):
        t = np.eye(4); t[0, 3], t[2, 3] = cx, cz
        ti = np.eye(4); ti[0, 3], ti[2, 3] = -cx, -cz
        return t @ m @ ti

    short, tall, light_bundle = 5, 6, int(sc.lights[0, 0])   # cornell_box's bundle order
    xf[tall] = pair(about(rot_y(0.35) @ np.diag([0.9, 1.1, 1.2, 1.0]), 1.4, 0.6) + np.array(
        [[0, 0, 0, -0.05], [0, 0, 0, 0], [0, 0, 0, 0.05], [0, 0, 0, 0]]))
    xf[short] = pair(about(np.diag([-1.0, 1.0, 1.0, 1.0]) @ rot_y(-0.2), 0.6, 1.2))
    xf[light_bundle] = pair(about(np.diag([1.5, 1.0, 1.0, 1.0]), 1.0, 1.0))
    out = Scene(**{k: getattr(sc, k) for k in sc.__dataclass_fields__})
    out.transforms = xf.astype(np.float32)
    out.name = "transformed_cornell"
    return out


def textured_cornell(width=32, height=32, seed=7) -> Scene:
    """The Cornell box with texture coordinates on every face (each quad (0,0)-(2,2), so maps
    wrap) and texture maps: a checkerboard kd on the floor (Lambert), a ks gradient on the back
    wall (Microfacet), kd and kr maps on the left wall (OrenNayar), and a ks map on a glass short
    box. UVs also give the hit frames FillHit's UV derivatives."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sc = cornell_box(width, height)
    nv = sc.verts.shape[0]
    quad_uv = np.array([[0, 0], [2, 0], [2, 2], [0, 2]], np.float32)
    uvs = np.tile(quad_uv, (nv // 4, 1))
    assert uvs.shape[0] == nv                    # every bundle is a run of 4-vertex quads
    checker = np.where(((np.arange(16)[:, None] // 4 + np.arange(16)[None, :] // 4) % 2 == 0)[..., None],
                       np.array([0.8, 0.75, 0.7]), np.array([0.15, 0.2, 0.3])).astype(np.float32)
    gy, gx = np.meshgrid(np.linspace(0, 1, 4), np.linspace(0, 1, 8), indexing="ij")
    gradient = np.stack([0.2 + 0.6 * gx, 0.7 - 0.4 * gy, 0.3 + 0.2 * gx * gy], -1).astype(np.float32)
    noise = rng.uniform(0.2, 0.9, (12, 12, 3)).astype(np.float32)
    rough = rng.uniform(0.1, 1.4, (6, 6, 3)).astype(np.float32)
    out = Scene(**{k: getattr(sc, k) for k in sc.__dataclass_fields__})
    mats = sc.materials.copy()
    # cornell_box's rows: walls 0-4 (floor, ceiling, back, left, right), boxes 5-6, light 7
    mats[2] = [MICROFACET, 0.6, 0.6, 0.6, 1.5, 0.3, 0, 0]
    mats[3] = [ORENNAYAR, 0.5, 0.5, 0.5, 0, 0.7, 0, 0]
    mats[5] = [GLASS, 0.9, 0.9, 0.9, 1.5, 0, 0, 0]
    maps = -np.ones((mats.shape[0], 2), np.int32)
    maps[0, 0] = 0          # floor kd: checkerboard
    maps[2, 0] = 1          # back wall ks: gradient
    maps[3] = (2, 3)        # left wall kd and kr
    maps[5, 0] = 2          # glass ks
    out.materials = mats
    out.uvs = uvs
    out.tri_uv = sc.tri_v.copy()
    out.textures = [checker, gradient, noise, rough]
    out.mat_maps = maps
    out.name = "textured_cornell"
    return out


def normal_mapped_cornell(width=32, height=32, seed=7) -> Scene:
    """textured_cornell with a normal map (a 16 x 16 bump pattern, tangent-space normals tilted
    up to 50 degrees) on the floor (Lambert), the back wall (Microfacet), the left wall
    (OrenNayar) and the second box (Lambert), read as NormalMapping reads them."""
    out = textured_cornell(width, height, seed)
    y, x = np.meshgrid(np.arange(16), np.arange(16), indexing="ij")
    bump = np.stack([0.5 + 0.35 * np.sin(2 * np.pi * x / 8), 0.5 + 0.35 * np.cos(2 * np.pi * y / 5),
                     np.full(x.shape, 0.8)], -1).astype(np.float32)
    out.textures = list(out.textures) + [bump]
    nm = -np.ones(out.materials.shape[0], np.int32)
    nm[[0, 2, 3, 6]] = len(out.textures) - 1
    out.normal_maps = nm
    return out

