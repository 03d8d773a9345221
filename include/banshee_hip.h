/*
 * banshee_hip.h -- C ABI of the MI355X (gfx950) global-illumination hot path.
 *
 * Plain C: opaque handles, caller-owned buffers, int status codes, no exceptions
 * cross this boundary. Device buffers are hipMalloc'd memory (from bsh_device_alloc,
 * torch, or the caller's own allocator); "stream" is a hipStream_t (NULL = the
 * legacy default stream).
 *
 * Each entry point names the reference interface it replaces (paths relative to
 * FireRays/Banshee/ of the reference). The reference's plug-in points are C++
 * virtual interfaces; the reference side binds this ABI through two thin adapters,
 * a HipAccelerator : Intersectable (per-ray callers) and a HipImageRenderer : Renderer
 * (whole frames).
 *
 * Threading: one context per device; a scene is immutable after bsh_scene_commit;
 * ray queries on different streams may run concurrently from different threads;
 * bsh_render / bsh_bake_sh / bsh_render_guided / bsh_bake_sh_guided / bsh_batch_run share
 * the context's wavefront buffers and serialise on a per-context lock (they return after
 * their work has completed).
 */
#ifndef BANSHEE_HIP_H
#define BANSHEE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int bsh_status;
enum
{
    BSH_OK = 0,
    BSH_ERR_INVALID = 1,     /* bad argument or scene (the reference throws std::runtime_error, arealight.cpp:16-19) */
    BSH_ERR_HIP = 2,         /* a HIP runtime call failed */
    BSH_ERR_UNSUPPORTED = 3, /* configuration this build does not implement */
    BSH_ERR_NO_DEVICE = 4,   /* no usable gfx950 device */
    BSH_ERR_STATE = 5        /* call order violated (e.g. render before commit) */
};

typedef struct bsh_context bsh_context;
typedef struct bsh_scene bsh_scene;

/* Ray, 32 B: the reference `ray` (math/ray.h:42-62) minus its unused id.
   Valid interval (tmin, tmax]: a hit is accepted iff tmin < t <= tmax (primitive/mesh.cpp:33). */
typedef struct
{
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} bsh_ray;

/* Closest hit, 16 B: global primitive index (bundle-concatenation order,
   accelerator/bvh.cpp:26-51; -1 = miss), distance and the barycentrics (b1, b2) of
   Mesh::IntersectFace (mesh.cpp:7-45): p = (1-b1-b2) p1 + b1 p2 + b2 p3. */
typedef struct
{
    int32_t prim;
    float t, b1, b2;
} bsh_hit;

/* Bundle = one reference Mesh (primitive/mesh.h:48-181): a contiguous triangle range
   with its own vertex and normal ranges. Only identity transforms are supported:
   callers bake ShapeBundle::SetTransform into the vertices. */
typedef struct
{
    int32_t tri_begin, tri_count;
    int32_t vert_begin, vert_count;
    int32_t norm_begin, norm_count;
} bsh_bundle;

/* Material row. type 0: SimpleMaterial(Lambert(kd=color)) (bsdf/lambert.h:45-135);
   type 1: SimpleMaterial(Microfacet(eta, ks=color, FresnelDielectric, GgxDistribution(alpha)))
   (bsdf/microfacet.h:53-150, microfacet_distribution.h:196-286, fresnel.h:53-91);
   type 2: Emissive(e=color) (material/emissive.h:45-99);
   type 3: SimpleMaterial(OrenNayar(kd=color, kr=alpha)) (bsdf/orennayar.h:47-190);
   type 4: SimpleMaterial(PerfectReflect(eta, ks=color, pad[0] != 0 ? FresnelDielectric : none))
   (bsdf/perfect_reflect.h:52-130); type 5: the same with PerfectRefract (perfect_refract.h:44-150);
   type 6: Glass(eta, ks=color) (material/glass.h:45-110);
   type 7: MixedMaterial(eta) with AddBsdf of the single-BSDF rows alpha, pad[0], pad[1] in that
   order (-1: none; rows of types 0, 1, 3, 4, 5; at least one reflecting)
   (material/mixedmaterial.h:45-174). */
typedef struct
{
    float type, color[3], eta, alpha, pad[2];
} bsh_material;

enum
{
    BSH_LAMBERT = 0,
    BSH_MICROFACET = 1,
    BSH_EMISSIVE = 2,
    BSH_ORENNAYAR = 3,
    BSH_PERFECT_REFLECT = 4,
    BSH_PERFECT_REFRACT = 5,
    BSH_GLASS = 6,
    BSH_MIXED = 7
};

/* AreaLight(shapeidx, bundle, material) (light/arealight.cpp:11-23). As in the
   reference, the bundle's light pointer is the LAST AreaLight created on it; the
   BSDF-sampled MIS branch of DiTracer::GetDi matches hits by that pointer
   (tracer/ditracer.cpp:202). */
typedef struct
{
    int32_t bundle, shape;
} bsh_area_light;

/* PerscpectiveCamera(eye, at, up, zcap=(znear, zfar), fovy, aspect)
   (camera/perspective_camera.cpp:5-23). */
typedef struct
{
    float eye[3], at[3], up[3];
    float znear, zfar, fovy, aspect;
} bsh_camera;

typedef struct
{
    int32_t max_leaf;   /* triangles per leaf, 1..8 (the reference uses 1, bvh.cpp:74) */
    int32_t sah_bins;   /* bins per axis for the binned SAH, 4..64 */
    int32_t threads;    /* host build threads, 0 = hardware concurrency */
    int32_t width;      /* node width: 2 (64-B binary nodes) or 4 (128-B 4-wide nodes collapsed from the
                           binary SAH tree); 0 = 4. A tree too deep for the 4-wide stack stays binary. */
    int32_t quantized;  /* width 4 only: 0 (default) or 1 stores the tree in 64-B nodes with 8-bit
                           child boxes (conservative, so closest hits are unchanged), -1 in 128-B
                           float nodes. A scene whose coordinates leave the exponent range the
                           quantised boxes allow keeps float nodes. The layout built is reported as
                           node bytes by bsh_scene_info and bsh_bvh_build_host (64 for binary or
                           quantised, 128 for float 4-wide). */
} bsh_bvh_params;

/* Frame parameters for bsh_render. The estimator is GiTracer(max_depth)
   (tracer/gitracer.cpp:13-140) with DiTracer::GetDi NEE (ditracer.cpp:57-239), driven by
   ImageRenderer::Render (renderer/imagerenderer.cpp:9-59) with the image sampler
   <img_sampler>(img_grid) and light and BSDF samplers <light_sampler>(1), <brdf_sampler>(1),
   all fed by the keyed sample stream. Camera sample g = (y*width + x)*spp + s, where spp is
   the image sampler's num_samples(): img_grid^2 for the stratified, regular and CMJ samplers,
   img_grid for RandomSampler, the next power of two >= img_grid for SobolSampler. */
typedef struct
{
    int32_t width, height;         /* full image plane (ImagePlane::resolution) */
    int32_t crop_x, crop_y;        /* rendered window; crop_w = crop_h = 0 means the full frame */
    int32_t crop_w, crop_h;
    int32_t img_grid;              /* spp = img_grid^2 */
    int32_t spp_begin, spp_count;  /* render samples [spp_begin, spp_begin+spp_count) of each pixel; count 0 = all */
    int32_t max_depth;             /* GiTracer maxdepth_ */
    uint32_t seed;
    int32_t chunk_spp;             /* reserved, must be 0 (BSH_ERR_INVALID otherwise); the paths in flight per
                                      wavefront chunk are bsh_tuning.paths_per_chunk */
    float bg[3];                   /* World::bgcolor_ */
    int32_t flags;                 /* BSH_RENDER_* */
    int32_t img_sampler;           /* BSH_SAMPLER_* of the image sampler, built with argument img_grid */
    int32_t light_sampler;         /* BSH_SAMPLER_* of the light sampler (argument 1); Sobol is not supported */
    int32_t brdf_sampler;          /* BSH_SAMPLER_* of the BSDF sampler (argument 1); Sobol is not supported */
} bsh_render_params;

/* Sampler kinds (the reference's sampler directory). 0 keeps the StratifiedSampler of earlier versions. */
enum
{
    BSH_SAMPLER_STRATIFIED = 0, /* StratifiedSampler(g), sampler/stratified_sampler.cpp:7-24 */
    BSH_SAMPLER_REGULAR = 1,    /* RegularSampler(g): cell centres, sampler/regular_sampler.h:46-72 */
    BSH_SAMPLER_RANDOM = 2,     /* RandomSampler(n), sampler/random_sampler.h:48-62 */
    BSH_SAMPLER_CMJ = 3,        /* CmjSampler(g): Kensler's correlated multi-jitter, sampler/cmj_sampler.cpp:3-73 */
    BSH_SAMPLER_SOBOL = 4       /* SobolSampler(n): scrambled (0,2)-sequence, reset per pixel,
                                   sampler/sobol_sampler.cpp:5-48 */
};

enum
{
    BSH_RENDER_ACCUMULATE = 1,  /* add into accum instead of overwriting it */
    BSH_RENDER_NO_QUIRKS = 2,   /* quirk-free estimator: no MINPDF cut or clamp, MIS in solid angle with GetPdf(view, dir),
                                   light hits matched by triangle, emitters end paths after bounce 0, a fresh BSDF
                                   sample per bounce (the unbiased judge of the guided tracer) */
    BSH_RENDER_COUNT = 4        /* count BVH node visits and triangle tests (slower; for roofline accounting) */
};

/* Per-kernel timings and ray counts of the last bsh_render / bsh_bake_sh / bsh_batch_run
   on a context (profiling). */
typedef struct
{
    uint64_t rays_closest, rays_any;   /* BVH queries issued (Intersectable::Intersect x2) */
    uint64_t camera_samples;
    uint64_t node_visits, tri_tests;   /* totals over all queries, BSH_RENDER_COUNT only */
    double trace_ms, shade_ms, total_ms; /* summed HIP-event durations */
    uint32_t trace_launches, shade_launches;
    uint32_t guide_launches;           /* guide network inference launches (guided render or bake) */
    uint64_t guide_rows;               /* path vertices run through the guide network */
    double guide_ms;                   /* summed duration of the inference launches */
    uint32_t graph_replayed;           /* bsh_batch_run with use_graph: 1 when the run replayed the context's
                                          previous graph (same jobs' shapes, tuning and buffers), 0 when it
                                          captured one */
} bsh_render_stats;

/* ---------------- context ---------------- */
bsh_status bsh_context_create(int device, bsh_context** out);
bsh_status bsh_context_destroy(bsh_context* ctx);
const char* bsh_last_error(void);  /* thread-local message for the last failing call */

/* Device memory helpers (hipMalloc/hipMemcpy wrappers) for hosts without a HIP runtime binding. */
bsh_status bsh_device_alloc(bsh_context* ctx, size_t bytes, void** dptr);
bsh_status bsh_device_free(bsh_context* ctx, void* dptr);
bsh_status bsh_copy_to_device(bsh_context* ctx, void* dst, const void* src, size_t bytes);
bsh_status bsh_copy_to_host(bsh_context* ctx, void* dst, const void* src, size_t bytes);
bsh_status bsh_device_sync(bsh_context* ctx);

/* ---------------- scene (World, world/world.h:51-88) ---------------- */
/* Replaces building World::shapebundles_ from Mesh ctors (mesh.cpp:126-214): the arrays are
   copied. Host pointers. tri_vidx/tri_nidx are global indices; tri_material indexes the
   material table. */
bsh_status bsh_scene_create(bsh_context* ctx, const float* verts, size_t nverts, const float* normals,
                            size_t nnormals, const int32_t* tri_vidx, const int32_t* tri_nidx,
                            const int32_t* tri_material, size_t ntris, const bsh_bundle* bundles,
                            size_t nbundles, bsh_scene** out);
/* Asset import, host only (no device needed): replaces AssimpAssetImporter::Import
   (import/assimp_assetimporter.cpp:20-146) for Wavefront OBJ/MTL files. What its onmaterial_,
   onprimitive_ and onlight_ callbacks would receive comes back as the arrays of
   bsh_scene_create / bsh_scene_set_materials / bsh_scene_set_lights:
   - materials: Assimp's default material (Lambert 0.6) first, then the MTL's in file order; a
     material named "light" or with a nonzero Ke is Emissive(0.6 (60, 55, 40)), any other
     Lambert(Kd) (assimp_assetimporter.cpp:74-85);
   - one bundle per non-empty run of faces between `o` / `usemtl` statements, in file order
     (Assimp's OBJ loader), its distinct (position, normal) pairs as its vertices;
   - triangles: quads split at their concave corner if any, else (a b c), (a c d); larger convex
     polygons fanned; faces without normals get their triangle's normal;
   - one AreaLight(shape, bundle) per triangle of every emissive bundle (:137-145).
   Texture maps are not applied (ignored_maps counts the materials that name one). Call with every
   array NULL to fill *counts, then again with arrays of at least those sizes (read from *counts). */
typedef struct
{
    size_t nverts, nnormals, ntris, nbundles, nmaterials, nlights;
    size_t ignored_maps;
} bsh_asset_counts;
bsh_status bsh_import_obj(const char* path, bsh_asset_counts* counts, float* verts, float* normals,
                          int32_t* tri_vidx, int32_t* tri_nidx, int32_t* tri_material, bsh_bundle* bundles,
                          bsh_material* materials, bsh_area_light* lights);
/* World::materials_ */
bsh_status bsh_scene_set_materials(bsh_scene* s, const bsh_material* mats, size_t n);
/* ShapeBundle::SetTransform(m, minv) of every bundle (primitive/shapebundle.h:139,261-265):
   n x 32 floats, row-major m (4x4) then row-major minv, n = the scene's bundle count; NULL
   restores identities. bsh_scene_commit bakes them into world space as the reference's meshes
   transform at use: points by transform_point(p, m), normals by transform_normal(n, minv) =
   minv^T n (math/mathutils.h:129-151); an area light's sample pdf keeps the untransformed
   triangle's area (mesh.cpp:121) and its GetPdf the transformed one's (mesh.cpp:304-321). The
   reference intersects in object space (mesh.cpp:216-261); the baked world-space test agrees
   with it up to rounding. */
bsh_status bsh_scene_set_transforms(bsh_scene* s, const float* m_minv, size_t n);
/* Texture coordinates (Mesh's uvs_ and Face::ti0..ti2, primitive/mesh.h:57-63): nuvs (u, v) pairs
   and three indices into them per triangle, in triangle order. A hit then gets FillHit's UV frame
   and coordinates (mesh.cpp:72-94); without them its uv is (0, 0) and its frame the orthovector
   one. With bundle transforms, the UV frame is taken from the world-space triangle (the
   reference maps its object-space tangents by minv^T, the same only for rotations). */
bsh_status bsh_scene_set_uvs(bsh_scene* s, const float* uvs, size_t nuvs, const int32_t* tri_uv);
/* Texture maps (what the BSDF constructors name as kdmap / ksmap / krmap): n maps of width x
   height RGB texels, row 0 at v = 0, each starting at float `offset` of `texels` (nfloats
   floats); read at a hit's (u, v) with bilinear filtering and periodic wrap (the texture system's
   defaults, texture/texturesystem.h:35). Host pointers, copied. */
typedef struct
{
    int32_t width, height;
    uint64_t offset;
} bsh_texture;
bsh_status bsh_scene_set_textures(bsh_scene* s, const float* texels, size_t nfloats, const bsh_texture* maps, size_t n);
/* Per material row: {colour map, kr map}, map indices or -1. The colour map replaces the row's
   kd (Lambert, OrenNayar) or ks (Microfacet, PerfectReflect, PerfectRefract, Glass; GET_VALUE,
   bsdf/bsdf.h:42); the kr map OrenNayar's kr (its first channel, orennayar.h:113). Emissive and
   MixedMaterial rows take none (a mixture's components carry their own). Normal maps
   (bsdf/normal_mapping.h) are not supported. */
bsh_status bsh_scene_set_material_maps(bsh_scene* s, const int32_t* maps, size_t nrows);
/* NormalMapping(bsdf, nmap) (bsdf/normal_mapping.h:40-90) around the single BSDF of material
   rows 0, 1, 3, 4, 5: per row a map index or -1. Bsdf::MapNormal (bsdf/bsdf.h:97-113) reads the
   map's closest texel at the hit's (u, v) as a tangent-space normal 2 texel - 1 in the frame of
   dpdu, dpdv and n. As in the reference, Sample and Evaluate rewrite the hit's normal in place
   and GetPdf maps a copy, so within one light's GetDi the BSDF sample sees the normal mapped
   twice when the light sample was unoccluded, once otherwise; the wavefront traces both BSDF
   rays and keeps the one the shadow ray selects (at most 16 lights at bounce 0). The quirk-free
   estimator maps the normal once for every use. NULL clears them. */
bsh_status bsh_scene_set_normal_maps(bsh_scene* s, const int32_t* maps, size_t nrows);
/* World::lights_ (AreaLight only) */
bsh_status bsh_scene_set_lights(bsh_scene* s, const bsh_area_light* lights, size_t n);

/* The reference's other lights (light/pointlight.h, directional_light.h, environment_light.h,
   environment_light_is.h). They follow the area lights in World::lights_, in array order: the
   light list GiTracer loops over is the area lights of bsh_scene_set_lights, then these. */
enum
{
    BSH_LIGHT_POINT = 1,          /* PointLight(p = v, e): GetSample d = p - x, pdf 1, e / |d|^2 */
    BSH_LIGHT_DIRECTIONAL = 2,    /* DirectionalLight(d = v, e): v is normalised as the ctor does */
    BSH_LIGHT_ENVIRONMENT = 3,    /* EnvironmentLight(texsys, map, scale, gamma): cosine-sampled */
    BSH_LIGHT_ENVIRONMENT_IS = 4  /* EnvironmentLightIs(texsys, map, scale, gamma): sampled by a
                                     512 x 256 Distribution2D of the map's luminance x sin(theta) */
};
typedef struct
{
    int32_t type;       /* BSH_LIGHT_* */
    float v[3];         /* point: position; directional: direction; environment: unused */
    float e[3];         /* point, directional: emission; environment: unused */
    float scale, gamma; /* environment: radiance scale and the map's gamma (reference defaults 1, 2.2) */
} bsh_light;
bsh_status bsh_scene_set_other_lights(bsh_scene* s, const bsh_light* lights, size_t n);
/* The lat-long map the environment lights read (the texture they name): height rows of width
   RGB floats, host memory, copied. Row 0 is theta = 0 (+y); column 0 is phi = 0 (+x), phi
   growing toward +z (math/mathutils.h:68-74). It is read with bilinear filtering and periodic
   wrap, texel centres at half-integer coordinates (TextureSystem defaults). */
bsh_status bsh_scene_set_env_map(bsh_scene* s, const float* texels, int width, int height);
/* World::camera_ */
bsh_status bsh_scene_set_camera(bsh_scene* s, const bsh_camera* cam);
/* World::Commit (world/world.cpp:5-16): host SAH build, flatten, upload. params may be NULL. */
bsh_status bsh_scene_commit(bsh_scene* s, const bsh_bvh_params* params);
bsh_status bsh_scene_destroy(bsh_scene* s);
/* The host half of bsh_scene_commit on its own (no device needed): builds the BVH of ntris
   triangles (tri_verts: 9 floats each) and, when nodes is not NULL, copies the flattened nodes
   (GpuNode 64 B at *width 2; at *width 4 GpuNode4 128 B or GpuNodeQ 64 B, per *node_bytes;
   nodes_cap in nodes) and the leaf order (tri_order: ntris ints). Call with nodes = NULL first
   to learn *nnodes. */
bsh_status bsh_bvh_build_host(const float* tri_verts, size_t ntris, const bsh_bvh_params* params, void* nodes,
                              size_t nodes_cap, int32_t* tri_order, size_t* nnodes, int32_t* width, int32_t* depth,
                              int32_t* node_bytes);
/* Node count, triangle count, device bytes, build milliseconds, max depth, BVH width built and
   bytes per node (64: binary or quantised 4-wide, 128: float 4-wide). Any output may be NULL. */
bsh_status bsh_scene_info(const bsh_scene* s, uint64_t* nodes, uint64_t* tris, uint64_t* bytes, double* build_ms,
                          int32_t* depth, int32_t* width, int32_t* node_bytes);

/* ---------------- ray queries (Intersectable, accelerator/intersectable.h:41-49) ---------------- */
/* Intersectable::Intersect(ray const&, Hit&) -> Bvh::Intersect closest hit (bvh.cpp:499-573).
   Device pointers. */
bsh_status bsh_intersect_closest(bsh_scene* s, const bsh_ray* rays, size_t n, bsh_hit* hits, void* stream);
/* Intersectable::Intersect(ray const&) -> Bvh::Intersect any hit (bvh.cpp:575-647).
   occluded[i] = 1 if any triangle has tmin < t <= tmax. Device pointers. */
bsh_status bsh_intersect_any(bsh_scene* s, const bsh_ray* rays, size_t n, int32_t* occluded, void* stream);
/* Counting mode: same queries, also writing per-ray node visits and triangle tests
   (visits[2*i], visits[2*i+1]); used for the algorithmic-bytes roofline. */
bsh_status bsh_intersect_count(bsh_scene* s, const bsh_ray* rays, size_t n, int any, bsh_hit* hits,
                               uint32_t* visits, void* stream);

/* ---------------- frame render (Renderer::Render, renderer/renderer.h:49) ---------------- */
/* Replaces ImageRenderer/MtImageRenderer::Render(world) (imagerenderer.cpp:9-59,
   mt_imagerenderer.cpp:16-117) for a GiTracer. accum: device float[crop_w*crop_h*8]:
   (sum r, g, b, n, sum r^2, g^2, b^2, 0) per pixel, row-major, NOT y-flipped
   (FileImagePlane flips only when writing a file, fileimageplane.cpp:29-34).
   samples: optional device float[crop_w*crop_h*spp_count*3] receiving every camera
   sample's returned radiance (GiTracer::GetLi's return value) in (pixel, s) order. */
bsh_status bsh_render(bsh_scene* s, const bsh_render_params* p, float* accum, float* samples, void* stream);
/* Stats of the last render / bake / batch on the scene's context. */
bsh_status bsh_render_stats_get(const bsh_scene* s, bsh_render_stats* out);

/* ---------------- spherical harmonics (math/sh.h, math/shproject.h, tracer/shtracer.h) ---------------- */
/* Probe baking (no reference counterpart; it composes GiTracer::GetLi with ShEvaluate,
   sh.cpp:97-153): for every probe, rays_per_probe stratified directions over the sphere
   (uniform pdf 1/(4 pi)), each estimated with the same GiTracer(max_depth) as bsh_render,
   projected as c_i = (4 pi / N) sum_j L(w_j) Y_i(w_j). Probe ray j of probe p is keyed
   as sample g = p*N + j. */
typedef struct
{
    int32_t rays_per_probe;   /* N, a perfect square (stratified sqrt(N) x sqrt(N) grid) */
    int32_t lmax;             /* 0..4; the reference's order-3 probes use lmax = 2 (9 coefficients) */
    int32_t max_depth;
    uint32_t seed;
    int32_t chunk_probes;     /* reserved, must be 0 (BSH_ERR_INVALID otherwise); see bsh_tuning.paths_per_chunk */
    float tmin;               /* probe ray tmin */
    float bg[3];
    int32_t flags;            /* BSH_RENDER_COUNT and BSH_RENDER_NO_QUIRKS honoured */
} bsh_probe_params;

/* probes: device float[nprobes*3]; coeffs: device float[nprobes*(lmax+1)^2*3] (coefficient-major,
   RGB inner); radiance: optional device float[nprobes*N*3] receiving every probe ray's GetLi. */
bsh_status bsh_bake_sh(bsh_scene* s, const float* probes, size_t nprobes, const bsh_probe_params* p, float* coeffs,
                       float* radiance, void* stream);
/* ShProjectEnvironmentMap(envmap, W, H, lmax, coeffs) (math/shproject.cpp:73-130): Riemann sum over a
   lat-long map with the reference's y-up direction. env: device float[H*W*3]; coeffs: device
   float[(lmax+1)^2*3], overwritten. */
bsh_status bsh_sh_project_env(bsh_context* ctx, const float* env, int width, int height, int lmax, float* coeffs,
                              void* stream);
/* bsh_sh_project_env without any allocation. No counterpart in the reference: it is for callers
   that cannot allocate device memory (code under stream capture, tight serving loops). The sin/cos
   tables and the block partials live in a 4 MiB device array of the library (one per device,
   held until the process ends), so 2(W+H) + ceil(W*H/256)*3(lmax+1)^2 floats (tables rounded up
   to 64 floats) must fit in 1,048,576 (e.g. 2048x1024 at lmax 4, 4096x2048 at lmax 2); larger
   maps return BSH_ERR_INVALID. The tables are copied up through a stack buffer, with the stream
   synchronised after each 8 KB piece; the kernels are then queued on `stream` without waiting.
   Calls on one device share the array: calls that overlap on one device (other streams or
   threads) race. Same results as bsh_sh_project_env, bit for bit. */
bsh_status bsh_sh_project_env_noalloc(bsh_context* ctx, const float* env, int width, int height, int lmax,
                                      float* coeffs, void* stream);
/* ShTracer::GetE (tracer/shtracer.cpp:20-36): irradiance sum_i Y_i(n) ShConvolveCosTheta(c)_i 0.7/pi
   (math/sh.cpp:160-180). coeffs device float[(lmax+1)^2*3]; normals/out device float[n*3]. */
bsh_status bsh_sh_irradiance(bsh_context* ctx, int lmax, const float* coeffs, const float* normals, size_t n,
                             float* out, void* stream);

/* ---------------- batching (no reference counterpart) ---------------- */
/* Launch-shape parameters of the wavefront: paths in flight per chunk (queue capacity)
   and persistent trace workgroups per CU (1..5; 0 = automatic: 3 when the run's chunks are
   dealt over two wavefront lanes, else 5); use_graph captures
   the whole batch as one hipGraph and launches it once. refill_lanes: on 4-wide BVHs each trace lane holds its own
   ray and a wave claims new rays once this many of its 64 lanes are idle (1..64; 0 = the
   default, 56); -1 traces in whole 64-ray batches per wave instead (also the form used when
   a chunk holds a binary BVH). claim_rays: queue entries a trace wave claims per atomic on
   the launch's work counter, keeping what its idle lanes do not take for its next refills
   (64..4096; 0 = automatic: exactly the idle lanes when every job of the run shares one
   scene, 128 when the run spans several, whose small scenes refill often enough for the
   shared counter to limit). Results do not depend on refill_lanes or claim_rays. */
typedef struct
{
    int32_t paths_per_chunk;
    int32_t trace_blocks_per_cu;
    int32_t use_graph;
    int32_t refill_lanes;
    int32_t claim_rays;
    int32_t lanes;     /* wavefront lanes: 0 = automatic (two when the run has chunks to alternate and writes
                          no training records), 1 = one lane on the caller's stream, 2 = two lanes */
} bsh_tuning;

enum { BSH_JOB_RENDER = 0, BSH_JOB_BAKE_SH = 1 };

/* One entry of a batch: exactly the arguments of bsh_render (kind BSH_JOB_RENDER) or of
   bsh_bake_sh (kind BSH_JOB_BAKE_SH), on any committed scene of the context. */
typedef struct
{
    bsh_scene* scene;
    int32_t kind;
    int32_t reserved;
    bsh_render_params render;
    float* accum;
    float* samples;
    bsh_probe_params probe;
    const float* probes;
    size_t nprobes;
    float* coeffs;
    float* radiance;
} bsh_job;

/* Runs every job in one wavefront: paths of all jobs share the chunks, queues and
   persistent trace kernel (a chunk mixes scenes), optionally captured as one hipGraph.
   Each job's output is bit-identical to running it alone. tuning NULL = context tuning. */
bsh_status bsh_batch_run(bsh_context* ctx, const bsh_job* jobs, size_t njobs, const bsh_tuning* tuning,
                         void* stream);
bsh_status bsh_context_set_tuning(bsh_context* ctx, const bsh_tuning* t);
bsh_status bsh_context_get_tuning(const bsh_context* ctx, bsh_tuning* t);
bsh_status bsh_context_stats_get(const bsh_context* ctx, bsh_render_stats* out);
/* Per bounce launch of the last bsh_render / bsh_bake_sh / bsh_batch_run / guided run on a context
   (profiling): entry b sums the b-th trace launch of every chunk. Launch 0 traces the camera (or
   probe) rays; launch b >= 1 the extension and MIS rays (closest hit) and the shadow rays (any
   hit) that shade pass b-1 emitted. */
typedef struct
{
    uint64_t rays_closest, rays_any;
    uint64_t nodes_closest, tris_closest;   /* node visits and triangle tests, BSH_RENDER_COUNT only */
    uint64_t nodes_any, tris_any;
    double trace_ms;                        /* summed HIP-event durations of these launches (0 in graph mode) */
} bsh_launch_stats;
/* Copies min(cap, launches) entries to out and the number of launches of the last run to *n. */
bsh_status bsh_context_launch_stats(const bsh_context* ctx, bsh_launch_stats* out, size_t cap, size_t* n);
/* Autotuner: times the batch under every candidate (paths_per_chunk x trace_blocks_per_cu),
   ncand each, and stores the fastest as the context tuning (also returned in best).
   The jobs' outputs are overwritten. NULL candidate arrays use the built-in sweep. */
bsh_status bsh_autotune(bsh_context* ctx, const bsh_job* jobs, size_t njobs, const int32_t* paths_per_chunk,
                        size_t n_paths, const int32_t* blocks_per_cu, size_t n_blocks, bsh_tuning* best,
                        double* best_ms);

/* ---------------- learned path guiding (no reference counterpart) ---------------- */
/* A vMF-mixture proposal (8 lobes) predicted by an MLP of position, normal and roughness
   (7 inputs, 6-band frequency encoding -> 96, 4 x 64 ReLU, 48 outputs; FP16 weights,
   FP32 accumulation on MFMA), trained online with Adam on the tracer's own radiance
   samples by minimising -sum_s (L_s |cos_s| / p_s) log q(w_s). */
typedef struct bsh_guide bsh_guide;
typedef struct
{
    uint32_t seed;        /* He-uniform initialisation seed */
    float lr, beta1, beta2, eps;   /* Adam */
    float bsdf_fraction;  /* one-sample MIS: probability of sampling the BSDF instead of the guide, in (0, 1] */
    int32_t guided_bounces; /* path vertices 0 .. guided_bounces-1 use the guide, later ones the BSDF alone
                               (and skip the network); 0 = every vertex */
    int32_t reserved;
} bsh_guide_params;

size_t bsh_guide_num_params(void);
bsh_status bsh_guide_create(bsh_context* ctx, const bsh_guide_params* p, bsh_guide** out);
bsh_status bsh_guide_destroy(bsh_guide* g);
/* FP32 master parameters (bsh_guide_num_params floats, layout W1[96][64], b1, W2[64][64], b2,
   W3, b3, W4, b4, W5[64][48], b5); set resets the Adam state. Host pointers. */
bsh_status bsh_guide_get_params(const bsh_guide* g, float* host);
bsh_status bsh_guide_set_params(bsh_guide* g, const float* host);
/* Training state for checkpoint and resume (no reference counterpart: the reference keeps no
   state beyond its image plane): the FP32 master parameters, Adam's first and second moments
   (bsh_guide_num_params floats each) and its step count. set_state restores all of them (and
   the FP16 copies derived from the parameters), so training resumed from a saved state
   continues bit for bit as it would have without the interruption. Host pointers. */
bsh_status bsh_guide_get_state(const bsh_guide* g, float* params, float* m1, float* m2, int64_t* step);
bsh_status bsh_guide_set_state(bsh_guide* g, const float* params, const float* m1, const float* m2, int64_t step);
/* The one-sample MIS BSDF fraction used by later guided render and bake calls, in (0, 1]. */
bsh_status bsh_guide_set_bsdf_fraction(bsh_guide* g, float bsdf_fraction);
/* Network outputs for n rows of raw inputs (8 floats per row: position in [0,1]^3,
   roughness, normal, unused). Device pointers; out n x 48. */
bsh_status bsh_guide_forward(bsh_guide* g, const float* inputs, size_t n, float* out, void* stream);
/* Training records: 12 floats per row (position, roughness, normal, unused, direction,
   weight). bsh_render_guided writes weight = lum(L_in) |n.w| / p_mix. Device records; grad
   host float[num_params]; loss host (weighted mean). */
bsh_status bsh_guide_gradient(bsh_guide* g, const float* records, size_t n, float* grad, float* loss);
/* FP32 mode of bsh_guide_forward and bsh_guide_gradient, for verification: the same network,
   loss and backward from the FP32 master parameters with FP32 features, activations and
   gradients and plain FMA arithmetic (no MFMA, no FP16 rounding), every sum in a fixed
   order. Same arguments; at most 2^22 rows (BSH_ERR_UNSUPPORTED beyond). Not a fast path:
   work buffers are allocated per call and the call returns after the work is done. */
bsh_status bsh_guide_forward_fp32(bsh_guide* g, const float* inputs, size_t n, float* out, void* stream);
bsh_status bsh_guide_gradient_fp32(bsh_guide* g, const float* records, size_t n, float* grad, float* loss);
/* One Adam step on the records. loss may be NULL. */
bsh_status bsh_guide_train_step(bsh_guide* g, const float* records, size_t n, float* loss, void* stream);
/* `steps` Adam steps on batches of up to `batch` records, e.g. the records one guided pass
   wrote: with stride = max(1, n / batch), step s uses records s mod stride, + stride, ...
   so every batch spans the whole array. loss (may be NULL) receives the last step's loss. */
bsh_status bsh_guide_train(bsh_guide* g, const float* records, size_t n, int steps, size_t batch, float* loss,
                           void* stream);
/* Renders like bsh_render (flags must include BSH_RENDER_NO_QUIRKS) with the continuation
   direction drawn from the one-sample MIS mixture of BSDF and guide (guide may be NULL: BSDF
   only, the plain quirk-free tracer), and appends training records of the first four path
   vertices to `records` (device float[cap*12]) in (path, vertex) order, the same on every
   run; records beyond cap are dropped and *nrecords receives the count written. */
bsh_status bsh_render_guided(bsh_scene* s, const bsh_render_params* p, bsh_guide* guide, float* accum,
                             float* records, size_t cap, size_t* nrecords, void* stream);
/* Probe baking like bsh_bake_sh with the quirk-free estimator (p->flags must include
   BSH_RENDER_NO_QUIRKS), guided at every vertex of a probe path. The probe itself is vertex
   -1: the network is queried at the probe position (roughness 0, normal 0), and probe ray j
   is, with probability bsdf_fraction, the stratified uniform direction of cell j, else a
   sample of the probe's mixture; its projection weight is w_j = 1 / (4 pi p_mix(w_j)),
   p_mix = bsdf_fraction / (4 pi) + (1 - bsdf_fraction) q, and c_i = (4 pi / N)
   sum_j w_j L(w_j) Y_i(w_j) stays unbiased. Surface vertices continue as in
   bsh_render_guided. guide may be NULL (the plain quirk-free bake, w_j = 1). radiance as in
   bsh_bake_sh (L(w_j), unweighted); ray_dirs: optional device float[nprobes*N*4] receiving
   every probe ray's (direction, w_j). records, cap and nrecords as in bsh_render_guided; on
   a guided bake record slot 0 of a path is the probe (normal 0, the record's 8th float 0,
   weight lum(L) / p_mix: no cosine factor), then its first three surface vertices. */
bsh_status bsh_bake_sh_guided(bsh_scene* s, const float* probes, size_t nprobes, const bsh_probe_params* p,
                              bsh_guide* guide, float* coeffs, float* radiance, float* ray_dirs, float* records,
                              size_t cap, size_t* nrecords, void* stream);

/* ---------------- diagnostics (exercise device code paths in isolation) ---------------- */
/* One v_mfma_f32_16x16x16_f16 tile: d = a b for 16x16 row-major FP16 a, b (bit patterns). */
bsh_status bsh_diag_mfma(bsh_context* ctx, const uint16_t* a, const uint16_t* b, float* d);
/* Guide distribution on given network outputs (host pointers): outputs n x 48, uniforms u n x 3,
   probe directions n x 3 -> out n x 8 = (vMF-mixture sample xyz, pdf at the sample, pdf at
   dirs, sum of lobe weights, 0, 0). */
bsh_status bsh_diag_guide(bsh_context* ctx, size_t n, const float* outputs, const float* u, const float* dirs,
                          float* out);
/* Device ShEvaluate (math/sh.cpp:97-153) on n unit directions. Host pointers: dirs n x 3,
   out n x (lmax+1)^2. */
bsh_status bsh_diag_sh_eval(bsh_context* ctx, int lmax, size_t n, const float* dirs, float* out);
/* Device evaluation of Material::Sample / Evaluate / GetPdf for one material row, same
   argument layout as the probes in tests: frames n x 12 (n, ng, dpdu, dpdv), view n x 3,
   u n x 2, ea/eb n x 3; out n x 12. Host pointers. */
bsh_status bsh_diag_bsdf(bsh_context* ctx, const bsh_material* mat, size_t n, const float* frames,
                         const float* view, const float* u, const float* ea, const float* eb, float* out);
/* Device evaluation of AreaLight::GetSample / GetPdf for scene light `light`. Host pointers.
   p n x 3, u n x 2, w n x 3; out n x 8 (le3, pdf, d3, getpdf). */
bsh_status bsh_diag_light(bsh_scene* s, int light, size_t n, const float* p, const float* u, const float* w,
                          float* out);
/* Device camera rays for image-plane samples in [0,1]^2 (PerscpectiveCamera::GenerateRay,
   perspective_camera.cpp:25-38). Host pointers: samples n x 2, rays n x 8. */
bsh_status bsh_diag_camera(bsh_context* ctx, const bsh_camera* cam, size_t n, const float* samples, float* rays);
/* Device keyed sample stream values ks_u32(ks_base(seed, camsample), stream, counter). Host pointers. */
bsh_status bsh_diag_keyed(bsh_context* ctx, uint32_t seed, size_t n, const uint32_t* camsample,
                          const uint32_t* stream, const uint32_t* counter, uint32_t* out);
/* Device sampler patterns (sampler/stratified_sampler.cpp, regular_sampler.h, random_sampler.h,
   cmj_sampler.cpp, sobol_sampler.cpp) on the keyed stream, as bsh_render forms them. image != 0:
   the image samples of a one-row image of n / spp pixels, camera sample g = p*spp + s, for the
   sampler BSH_SAMPLER_<kind>(arg); image == 0: the first n samples of a light sampler
   <kind>(1) in camera sample 0. out: host float[n*2]. */
bsh_status bsh_diag_samples(bsh_context* ctx, int kind, int arg, int image, uint32_t seed, size_t n, float* out);
/* Device Material::Sample / Evaluate / GetPdf of material row `material` of a committed scene
   (any type; MixedMaterial reads its component rows), the rand() draws of Glass and
   MixedMaterial for case i keyed by ks_base(seed, i) from counter 0. Host pointers, as
   bsh_diag_bsdf: out n x 12 = Sample f(3), pdf, wo(3), type, Evaluate(ea, eb)(3), GetPdf(ea, eb). */
bsh_status bsh_diag_material(bsh_scene* s, int material, uint32_t seed, size_t n, const float* frames,
                             const float* view, const float* u, const float* ea, const float* eb, float* out);
/* Device Light::GetSample / GetPdf / GetLe of light `light` of the light list (area lights, then
   the other lights) at points p (n x 3) with shading normals nrm (n x 3), samples u (n x 2) and
   pdf/GetLe directions w (n x 3); out: n x 16 = le(3), pdf, d(3), GetPdf(w), GetLe(w)(3), 0 x 5.
   Host pointers. */
bsh_status bsh_diag_light_full(bsh_scene* s, int light, size_t n, const float* p, const float* nrm, const float* u,
                               const float* w, float* out);

/* Memory ceiling for the roofline (no reference counterpart): times one streaming read of `reps`
   passes over `bytes` of device memory at `buf` (a multiple of 16, at least 4 KiB) after one
   untimed pass, and returns the duration in *ms. per_xcd = 0: every byte is read once per pass
   (HBM for buffers beyond the 256 MB Infinity Cache, the Infinity Cache for buffers inside it);
   per_xcd = 1: each of the 8 XCDs reads the whole buffer every pass (its L2, for buffers that fit
   one XCD's 4 MB L2), so a pass moves 8 x bytes. per_xcd | 2: non-temporal loads (the fastest
   form from HBM); per_xcd | 4: 8 workgroups per CU instead of 32 (the faster form from L2).
   Eight 16-B loads in flight per lane. */
bsh_status bsh_diag_stream_read(bsh_context* ctx, const void* buf, size_t bytes, int reps, int per_xcd, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* BANSHEE_HIP_H */
