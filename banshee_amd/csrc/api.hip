// C ABI: context, device memory helpers, scene construction and World::Commit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"

namespace bsh {

static thread_local std::string tl_error;

void set_error(const std::string& msg) { tl_error = msg; }

bsh_status fail(bsh_status code, const std::string& msg)
{
    tl_error = msg;
    return code;
}

bsh_status hip_fail(hipError_t e, const char* what)
{
    tl_error = std::string(what) + ": " + hipGetErrorString(e);
    return BSH_ERR_HIP;
}

// Every launch-shape parameter a tuning may carry, checked before any work is enqueued:
// trace workgroups per CU are bounded by the trace kernel's LDS stack; a claim must cover a
// whole wave's idle lanes (64), or a wave's reserve would hand out entries past its claim.
bsh_status validate_tuning(const bsh_tuning& t, const char* who)
{
    const std::string w(who);
    if (t.paths_per_chunk < 0) return fail(BSH_ERR_INVALID, w + ": paths_per_chunk must be >= 0");
    if (t.trace_blocks_per_cu < 0 || t.trace_blocks_per_cu > 5)
        return fail(BSH_ERR_INVALID, w + ": trace_blocks_per_cu must be 0..5 (LDS limit)");
    if (t.refill_lanes < -1 || t.refill_lanes > 64) return fail(BSH_ERR_INVALID, w + ": refill_lanes must be -1..64");
    if (t.claim_rays != 0 && (t.claim_rays < 64 || t.claim_rays > 4096))
        return fail(BSH_ERR_INVALID, w + ": claim_rays must be 0 or 64..4096");
    if (t.use_graph != 0 && t.use_graph != 1) return fail(BSH_ERR_INVALID, w + ": use_graph must be 0 or 1");
    if (t.lanes < 0 || t.lanes > bsh_context::kLanes) return fail(BSH_ERR_INVALID, w + ": lanes must be 0, 1 or 2");
    return BSH_OK;
}

CameraDev make_camera(const bsh_camera& c)
{
    // PerscpectiveCamera ctor, perspective_camera.cpp:5-23
    f3 eye = mk(c.eye[0], c.eye[1], c.eye[2]);
    f3 at = mk(c.at[0], c.at[1], c.at[2]);
    f3 up = mk(c.up[0], c.up[1], c.up[2]);
    f3 fwd = normalize(at - eye);
    f3 right = cross(fwd, normalize(up));
    f3 up2 = cross(right, fwd);
    CameraDev d;
    d.eye[0] = eye.x; d.eye[1] = eye.y; d.eye[2] = eye.z;
    d.fwd[0] = fwd.x; d.fwd[1] = fwd.y; d.fwd[2] = fwd.z;
    d.right[0] = right.x; d.right[1] = right.y; d.right[2] = right.z;
    d.up[0] = up2.x; d.up[1] = up2.y; d.up[2] = up2.z;
    d.znear = c.znear;
    d.zfar = c.zfar;
    d.dimx = 2.f * tanf(c.fovy * 0.5f) * c.znear;
    d.dimy = d.dimx / c.aspect;
    return d;
}

}  // namespace bsh

using namespace bsh;

namespace bsh {
// EnvironmentLightIs's ctor (environment_light_is.cpp:13-37) over the scene's map: the 512 x 256
// luminance x sin(theta) image, sampled at uv = (j / 512, i / 256) with the texture filter, then
// Distribution2D's tables (math/distribution1d.cpp:7-31, distribution2d.cpp:5-27) in the same
// float operations. out: marginal funcvals (256), cdf (257), conditional funcvals (256 x 512),
// cdfs (256 x 513), funcints (256); returns the marginal funcint.
static void dist1d_build(const float* vals, int n, float* val_out, float* cdf, float& funcint)
{
    for (int i = 0; i < n; ++i) val_out[i] = vals[i];
    cdf[0] = 0.f;
    for (int i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + vals[i - 1] / (float)n;
    funcint = cdf[n];
    for (int i = 0; i < n + 1; ++i) cdf[i] /= funcint;
}

float build_env_distribution(const float* tex, int w, int h, float* out)
{
    const int n = kEnvDistW, m = kEnvDistH;
    std::vector<float> img((size_t)n * m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j)
        {
            float u = (float)j / (float)n, v = (float)i / (float)m;
            float sintheta = sinf(kPi * (float)((float)i + 0.5f) / (float)m);
            f3 c = env_bilinear(tex, w, h, u, v);
            float luminance = 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
            img[(size_t)i * n + j] = luminance * sintheta;
        }
    float* mval = out;
    float* mcdf = mval + m;
    float* cval = mcdf + m + 1;
    float* ccdf = cval + (size_t)m * n;
    float* cint = ccdf + (size_t)m * (n + 1);
    for (int i = 0; i < m; ++i)
        dist1d_build(img.data() + (size_t)i * n, n, cval + (size_t)i * n, ccdf + (size_t)i * (n + 1), cint[i]);
    float mint = 0.f;
    dist1d_build(cint, m, mval, mcdf, mint);
    return mint;
}
}  // namespace bsh

extern "C" {

const char* bsh_last_error(void) { return tl_error.c_str(); }

bsh_status bsh_context_create(int device, bsh_context** out)
{
    if (!out) return fail(BSH_ERR_INVALID, "bsh_context_create: out is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= device || device < 0)
        return fail(BSH_ERR_NO_DEVICE, "bsh_context_create: no HIP device " + std::to_string(device));
    BSH_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    BSH_HIP(hipGetDeviceProperties(&prop, device));
    std::string arch = prop.gcnArchName;
    if (arch.rfind("gfx950", 0) != 0)
        return fail(BSH_ERR_NO_DEVICE, "bsh_context_create: device is " + arch + ", this build targets gfx950");
    bsh_status sv = env_scratch_init(device);
    if (sv != BSH_OK) return sv;
    bsh_context* c = new bsh_context();
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    c->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
    *out = c;
    return BSH_OK;
}

bsh_status bsh_context_destroy(bsh_context* ctx)
{
    delete ctx;
    return BSH_OK;
}

bsh_status bsh_device_alloc(bsh_context* ctx, size_t bytes, void** dptr)
{
    if (!ctx || !dptr) return fail(BSH_ERR_INVALID, "bsh_device_alloc: NULL argument");
    BSH_HIP(hipSetDevice(ctx->device));
    BSH_HIP(hipMalloc(dptr, std::max<size_t>(bytes, 1)));
    return BSH_OK;
}

bsh_status bsh_device_free(bsh_context* ctx, void* dptr)
{
    if (!ctx) return fail(BSH_ERR_INVALID, "bsh_device_free: NULL context");
    if (dptr) BSH_HIP(hipFree(dptr));
    return BSH_OK;
}

bsh_status bsh_copy_to_device(bsh_context* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx) return fail(BSH_ERR_INVALID, "bsh_copy_to_device: NULL context");
    if (bytes) BSH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return BSH_OK;
}

bsh_status bsh_copy_to_host(bsh_context* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx) return fail(BSH_ERR_INVALID, "bsh_copy_to_host: NULL context");
    if (bytes) BSH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return BSH_OK;
}

bsh_status bsh_device_sync(bsh_context* ctx)
{
    if (!ctx) return fail(BSH_ERR_INVALID, "bsh_device_sync: NULL context");
    BSH_HIP(hipDeviceSynchronize());
    return BSH_OK;
}

bsh_status bsh_scene_create(bsh_context* ctx, const float* verts, size_t nverts, const float* normals,
                            size_t nnormals, const int32_t* tri_vidx, const int32_t* tri_nidx,
                            const int32_t* tri_material, size_t ntris, const bsh_bundle* bundles, size_t nbundles,
                            bsh_scene** out)
{
    if (!ctx || !out) return fail(BSH_ERR_INVALID, "bsh_scene_create: NULL context or out");
    if ((nverts && !verts) || (nnormals && !normals) || (ntris && (!tri_vidx || !tri_nidx || !tri_material)) ||
        (nbundles && !bundles))
        return fail(BSH_ERR_INVALID, "bsh_scene_create: NULL array with nonzero count");
    if (ntris >= (size_t)1 << 29) return fail(BSH_ERR_UNSUPPORTED, "bsh_scene_create: more than 2^29 triangles");
    for (size_t i = 0; i < 3 * ntris; ++i)
    {
        if (tri_vidx[i] < 0 || (size_t)tri_vidx[i] >= nverts)
            return fail(BSH_ERR_INVALID, "bsh_scene_create: vertex index out of range at triangle " +
                                             std::to_string(i / 3));
        if (tri_nidx[i] < 0 || (size_t)tri_nidx[i] >= nnormals)
            return fail(BSH_ERR_INVALID, "bsh_scene_create: normal index out of range at triangle " +
                                             std::to_string(i / 3));
    }
    size_t covered = 0;
    for (size_t b = 0; b < nbundles; ++b)
    {
        const bsh_bundle& bb = bundles[b];
        if (bb.tri_begin != (int32_t)covered || bb.tri_count < 0)
            return fail(BSH_ERR_INVALID, "bsh_scene_create: bundles must tile the triangle range in order");
        covered += (size_t)bb.tri_count;
    }
    if (covered != ntris) return fail(BSH_ERR_INVALID, "bsh_scene_create: bundles do not cover all triangles");
    for (size_t i = 0; i < 3 * nverts; ++i)
        if (!std::isfinite(verts[i])) return fail(BSH_ERR_INVALID, "bsh_scene_create: non-finite vertex");
    bsh_scene* s = new bsh_scene();
    s->ctx = ctx;
    s->verts.assign(verts, verts + 3 * nverts);
    s->normals.assign(normals, normals + 3 * nnormals);
    s->tri_v.assign(tri_vidx, tri_vidx + 3 * ntris);
    s->tri_n.assign(tri_nidx, tri_nidx + 3 * ntris);
    s->tri_m.assign(tri_material, tri_material + ntris);
    s->bundles.assign(bundles, bundles + nbundles);
    *out = s;
    return BSH_OK;
}

bsh_status bsh_scene_set_materials(bsh_scene* s, const bsh_material* mats, size_t n)
{
    if (!s || (n && !mats)) return fail(BSH_ERR_INVALID, "bsh_scene_set_materials: NULL argument");
    for (size_t i = 0; i < n; ++i)
    {
        int t = (int)mats[i].type;
        if (t < BSH_LAMBERT || t > BSH_MIXED || (float)t != mats[i].type)
            return fail(BSH_ERR_INVALID, "bsh_scene_set_materials: unknown material type at " + std::to_string(i));
    }
    s->mats.assign(mats, mats + n);
    s->committed = false;
    return BSH_OK;
}

bsh_status bsh_scene_set_lights(bsh_scene* s, const bsh_area_light* lights, size_t n)
{
    if (!s || (n && !lights)) return fail(BSH_ERR_INVALID, "bsh_scene_set_lights: NULL argument");
    s->lights.assign(lights, lights + n);
    s->committed = false;
    return BSH_OK;
}

bsh_status bsh_scene_set_uvs(bsh_scene* s, const float* uvs, size_t nuvs, const int32_t* tri_uv)
{
    if (!s) return fail(BSH_ERR_INVALID, "bsh_scene_set_uvs: NULL scene");
    if (!uvs || !tri_uv)
    {
        s->uvs.clear();
        s->tri_uv.clear();
        s->committed = false;
        return BSH_OK;
    }
    const size_t nt = s->tri_m.size();
    for (size_t i = 0; i < 3 * nt; ++i)
        if (tri_uv[i] < 0 || (size_t)tri_uv[i] >= nuvs)
            return fail(BSH_ERR_INVALID, "bsh_scene_set_uvs: uv index out of range at triangle " + std::to_string(i / 3));
    for (size_t i = 0; i < 2 * nuvs; ++i)
        if (!std::isfinite(uvs[i])) return fail(BSH_ERR_INVALID, "bsh_scene_set_uvs: non-finite coordinate");
    s->uvs.assign(uvs, uvs + 2 * nuvs);
    s->tri_uv.assign(tri_uv, tri_uv + 3 * nt);
    s->committed = false;
    return BSH_OK;
}

bsh_status bsh_scene_set_textures(bsh_scene* s, const float* texels, size_t nfloats, const bsh_texture* maps, size_t n)
{
    if (!s || (n && (!texels || !maps))) return fail(BSH_ERR_INVALID, "bsh_scene_set_textures: NULL argument");
    for (size_t k = 0; k < n; ++k)
    {
        const bsh_texture& d = maps[k];
        if (d.width <= 0 || d.height <= 0 || d.offset + 3 * (uint64_t)d.width * (uint64_t)d.height > nfloats)
            return fail(BSH_ERR_INVALID, "bsh_scene_set_textures: map " + std::to_string(k) + " lies outside the texels");
    }
    for (size_t i = 0; i < nfloats; ++i)
        if (!std::isfinite(texels[i])) return fail(BSH_ERR_INVALID, "bsh_scene_set_textures: non-finite texel");
    s->tex_texels.assign(texels, texels + nfloats);
    s->tex_desc.assign(maps, maps + n);
    s->committed = false;
    return BSH_OK;
}

bsh_status bsh_scene_set_material_maps(bsh_scene* s, const int32_t* maps, size_t nrows)
{
    if (!s) return fail(BSH_ERR_INVALID, "bsh_scene_set_material_maps: NULL scene");
    if (!maps)
    {
        s->mat_maps.clear();
        s->committed = false;
        return BSH_OK;
    }
    s->mat_maps.assign(maps, maps + 2 * nrows);
    s->committed = false;
    return BSH_OK;
}

bsh_status bsh_scene_set_normal_maps(bsh_scene* s, const int32_t* maps, size_t nrows)
{
    if (!s) return fail(BSH_ERR_INVALID, "bsh_scene_set_normal_maps: NULL scene");
    if (!maps) s->nmaps.clear();
    else s->nmaps.assign(maps, maps + nrows);
    s->committed = false;
    return BSH_OK;
}

bsh_status bsh_scene_set_transforms(bsh_scene* s, const float* m_minv, size_t n)
{
    if (!s) return fail(BSH_ERR_INVALID, "bsh_scene_set_transforms: NULL scene");
    if (!m_minv)
    {
        s->xforms.clear();
        s->committed = false;
        return BSH_OK;
    }
    if (n != s->bundles.size())
        return fail(BSH_ERR_INVALID, "bsh_scene_set_transforms: one transform pair per bundle (" +
                                         std::to_string(s->bundles.size()) + ")");
    for (size_t i = 0; i < 32 * n; ++i)
        if (!std::isfinite(m_minv[i])) return fail(BSH_ERR_INVALID, "bsh_scene_set_transforms: non-finite entry");
    s->xforms.assign(m_minv, m_minv + 32 * n);
    s->committed = false;
    return BSH_OK;
}

bsh_status bsh_scene_set_other_lights(bsh_scene* s, const bsh_light* lights, size_t n)
{
    if (!s || (n && !lights)) return fail(BSH_ERR_INVALID, "bsh_scene_set_other_lights: NULL argument");
    for (size_t l = 0; l < n; ++l)
        if (lights[l].type < BSH_LIGHT_POINT || lights[l].type > BSH_LIGHT_ENVIRONMENT_IS)
            return fail(BSH_ERR_INVALID, "bsh_scene_set_other_lights: light " + std::to_string(l) + " has an unknown type");
    s->others.assign(lights, lights + n);
    s->committed = false;
    return BSH_OK;
}

bsh_status bsh_scene_set_env_map(bsh_scene* s, const float* texels, int width, int height)
{
    if (!s || !texels) return fail(BSH_ERR_INVALID, "bsh_scene_set_env_map: NULL argument");
    if (width <= 0 || height <= 0 || (size_t)width * height > ((size_t)1 << 26))
        return fail(BSH_ERR_INVALID, "bsh_scene_set_env_map: bad size");
    s->env_tex.assign(texels, texels + 3 * (size_t)width * height);
    for (float v : s->env_tex)
        if (!std::isfinite(v)) return fail(BSH_ERR_INVALID, "bsh_scene_set_env_map: non-finite texel");
    s->env_w = width;
    s->env_h = height;
    s->committed = false;
    return BSH_OK;
}

bsh_status bsh_scene_set_camera(bsh_scene* s, const bsh_camera* cam)
{
    if (!s || !cam) return fail(BSH_ERR_INVALID, "bsh_scene_set_camera: NULL argument");
    s->cam = *cam;
    s->has_cam = true;
    return BSH_OK;
}

bsh_status bsh_scene_commit(bsh_scene* s, const bsh_bvh_params* params)
{
    if (!s) return fail(BSH_ERR_INVALID, "bsh_scene_commit: NULL scene");
    BSH_HIP(hipSetDevice(s->ctx->device));
    const size_t nt = s->tri_m.size();
    // bundle transforms, baked into world-space copies: p' = transform_point(p, m), n' =
    // transform_normal(n, minv) = minv^T n, unnormalised (math/mathutils.h:129-151; FillHit
    // interpolates and then normalises, mesh.cpp:67-69)
    std::vector<float> wverts, wnormals;
    const bool xf = !s->xforms.empty();
    if (xf)
    {
        wverts = s->verts;
        wnormals = s->normals;
        for (size_t b = 0; b < s->bundles.size(); ++b)
        {
            const float* M = &s->xforms[32 * b];
            const float* Mi = M + 16;
            const bsh_bundle& B = s->bundles[b];
            for (int32_t k = B.vert_begin; k < B.vert_begin + B.vert_count; ++k)
            {
                const float* p = &s->verts[3 * (size_t)k];
                for (int r = 0; r < 3; ++r)
                {
                    float acc = 0.f;
                    for (int c = 0; c < 3; ++c) acc += M[4 * r + c] * p[c];
                    wverts[3 * (size_t)k + r] = acc + M[4 * r + 3];
                }
            }
            for (int32_t k = B.norm_begin; k < B.norm_begin + B.norm_count; ++k)
            {
                const float* n = &s->normals[3 * (size_t)k];
                for (int r = 0; r < 3; ++r)
                {
                    float acc = 0.f;
                    for (int c = 0; c < 3; ++c) acc += Mi[4 * c + r] * n[c];
                    wnormals[3 * (size_t)k + r] = acc;
                }
            }
        }
    }
    const std::vector<float>& VV = xf ? wverts : s->verts;
    const std::vector<float>& NN = xf ? wnormals : s->normals;
    // materials referenced by triangles must exist
    for (size_t i = 0; i < nt; ++i)
        if (s->tri_m[i] < 0 || (size_t)s->tri_m[i] >= s->mats.size())
            return fail(BSH_ERR_INVALID, "bsh_scene_commit: material index out of range at triangle " +
                                             std::to_string(i));
    // material rows: types 0-7; a MixedMaterial's components are rows of single BSDFs
    s->ext_mats = 0;
    for (size_t k = 0; k < s->mats.size(); ++k)
    {
        const bsh_material& m = s->mats[k];
        const int t = (int)m.type;
        if (!(m.type == (float)t) || t < BSH_LAMBERT || t > BSH_MIXED)
            return fail(BSH_ERR_INVALID, "bsh_scene_commit: material " + std::to_string(k) + " has an unknown type");
        if (t >= BSH_ORENNAYAR) s->ext_mats = 1;
        if (t == BSH_MIXED)
        {
            int nr = 0, ntr = 0;
            for (float fid : {m.alpha, m.pad[0], m.pad[1]})
            {
                const int id = (int)fid;
                if (id < 0) continue;
                if ((size_t)id >= s->mats.size())
                    return fail(BSH_ERR_INVALID, "bsh_scene_commit: mixed material " + std::to_string(k) +
                                                     " names a component row out of range");
                const int ct = (int)s->mats[id].type;
                if (ct != BSH_LAMBERT && ct != BSH_MICROFACET && ct != BSH_ORENNAYAR && ct != BSH_PERFECT_REFLECT &&
                    ct != BSH_PERFECT_REFRACT)
                    return fail(BSH_ERR_INVALID, "bsh_scene_commit: mixed material " + std::to_string(k) +
                                                     ": a component must be a single BSDF (types 0, 1, 3, 4, 5)");
                (bsdf_flags(ct) & BSDF_REFLECTION ? nr : ntr)++;
            }
            if (nr == 0)
                return fail(BSH_ERR_INVALID, "bsh_scene_commit: mixed material " + std::to_string(k) +
                                                 " needs a reflecting component (MixedMaterial::Sample asserts one)");
        }
    }
    {
        std::vector<float4> mt(2 * std::max<size_t>(s->mats.size(), 1), make_float4(0.f, 0.f, 0.f, 0.f));
        for (size_t k = 0; k < s->mats.size(); ++k)
        {
            const bsh_material& m = s->mats[k];
            mt[2 * k] = make_float4(m.type, m.color[0], m.color[1], m.color[2]);
            mt[2 * k + 1] = make_float4(m.eta, m.alpha, m.pad[0], m.pad[1]);
        }
        BSH_HIP(s->d_mats.alloc(mt.size()));
        BSH_HIP(hipMemcpy(s->d_mats.p, mt.data(), mt.size() * sizeof(float4), hipMemcpyHostToDevice));
    }
    // texture coordinates, maps and the rows that read them
    if (!s->tri_uv.empty() && s->tri_uv.size() != 3 * nt)
        return fail(BSH_ERR_STATE, "bsh_scene_commit: texture coordinates were set for another triangle count");
    if (!s->mat_maps.empty())
    {
        if (s->mat_maps.size() != 2 * s->mats.size())
            return fail(BSH_ERR_INVALID, "bsh_scene_commit: material maps need one pair per material row");
        for (size_t k = 0; k < s->mat_maps.size(); ++k)
        {
            const int32_t id = s->mat_maps[k];
            if (id < -1 || (id >= 0 && (size_t)id >= s->tex_desc.size()))
                return fail(BSH_ERR_INVALID, "bsh_scene_commit: material " + std::to_string(k / 2) +
                                                 " names a texture map out of range");
            const int t = (int)s->mats[k / 2].type;
            if (id >= 0 && (t == BSH_EMISSIVE || t == BSH_MIXED || (k % 2 == 1 && t != BSH_ORENNAYAR)))
                return fail(BSH_ERR_INVALID, "bsh_scene_commit: material " + std::to_string(k / 2) +
                                                 " takes no such map (colour maps: rows 0, 1, 3-6; kr maps: OrenNayar)");
        }
    }
    if (!s->nmaps.empty())
    {
        // NormalMapping wraps one BSDF (bsdf/normal_mapping.h:40-56): rows 0, 1, 3, 4, 5
        if (s->nmaps.size() != s->mats.size())
            return fail(BSH_ERR_INVALID, "bsh_scene_commit: normal maps need one entry per material row");
        for (size_t k = 0; k < s->nmaps.size(); ++k)
        {
            const int32_t id = s->nmaps[k];
            if (id < -1 || (id >= 0 && (size_t)id >= s->tex_desc.size()))
                return fail(BSH_ERR_INVALID, "bsh_scene_commit: material " + std::to_string(k) +
                                                 " names a normal map out of range");
            const int t = (int)s->mats[k].type;
            if (id >= 0 && (t == BSH_EMISSIVE || t == BSH_GLASS || t == BSH_MIXED))
                return fail(BSH_ERR_INVALID, "bsh_scene_commit: material " + std::to_string(k) +
                                                 " takes no normal map (rows 0, 1, 3, 4, 5: one BSDF)");
        }
    }
    if (!s->tri_uv.empty() || !s->mat_maps.empty() || !s->nmaps.empty())
        s->ext_mats = 1;   // read in the kLights shade variants
    {
        std::vector<float4> uv(2 * std::max<size_t>(s->tri_uv.empty() ? 0 : nt, 1), make_float4(0.f, 0.f, 0.f, 0.f));
        if (!s->tri_uv.empty())
            for (size_t i = 0; i < nt; ++i)
            {
                const float* a = &s->uvs[2 * (size_t)s->tri_uv[3 * i]];
                const float* b = &s->uvs[2 * (size_t)s->tri_uv[3 * i + 1]];
                const float* c = &s->uvs[2 * (size_t)s->tri_uv[3 * i + 2]];
                uv[2 * i] = make_float4(a[0], a[1], b[0], b[1]);
                uv[2 * i + 1] = make_float4(c[0], c[1], 0.f, 0.f);
            }
        BSH_HIP(s->d_tri_uv.alloc(uv.size()));
        BSH_HIP(hipMemcpy(s->d_tri_uv.p, uv.data(), uv.size() * sizeof(float4), hipMemcpyHostToDevice));
        BSH_HIP(s->d_tex.alloc(std::max<size_t>(s->tex_texels.size(), 1)));
        if (!s->tex_texels.empty())
            BSH_HIP(hipMemcpy(s->d_tex.p, s->tex_texels.data(), s->tex_texels.size() * 4, hipMemcpyHostToDevice));
        std::vector<int4> td(std::max<size_t>(s->tex_desc.size(), 1), make_int4(0, 0, 0, 0));
        for (size_t k = 0; k < s->tex_desc.size(); ++k)
            td[k] = make_int4(s->tex_desc[k].width, s->tex_desc[k].height, (int)(uint32_t)(s->tex_desc[k].offset & 0xffffffffu),
                              (int)(uint32_t)(s->tex_desc[k].offset >> 32));
        BSH_HIP(s->d_tex_desc.alloc(td.size()));
        BSH_HIP(hipMemcpy(s->d_tex_desc.p, td.data(), td.size() * sizeof(int4), hipMemcpyHostToDevice));
        std::vector<int2> mm(std::max<size_t>(s->mats.size(), 1), make_int2(-1, -1));
        for (size_t k = 0; k < s->mat_maps.size() / 2; ++k) mm[k] = make_int2(s->mat_maps[2 * k], s->mat_maps[2 * k + 1]);
        BSH_HIP(s->d_mat_maps.alloc(mm.size()));
        BSH_HIP(hipMemcpy(s->d_mat_maps.p, mm.data(), mm.size() * sizeof(int2), hipMemcpyHostToDevice));
        std::vector<int> nm(std::max<size_t>(s->mats.size(), 1), -1);
        for (size_t k = 0; k < s->nmaps.size(); ++k) nm[k] = s->nmaps[k];
        BSH_HIP(s->d_nmaps.alloc(nm.size()));
        BSH_HIP(hipMemcpy(s->d_nmaps.p, nm.data(), nm.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    // lights: validate exactly as AreaLight's ctor does (arealight.cpp:16-19)
    std::vector<int32_t> bundle_light(s->bundles.size(), -1);
    s->h_lights.clear();
    for (size_t l = 0; l < s->lights.size(); ++l)
    {
        const bsh_area_light& L = s->lights[l];
        if (L.bundle < 0 || (size_t)L.bundle >= s->bundles.size() || L.shape < 0 ||
            L.shape >= s->bundles[L.bundle].tri_count)
            return fail(BSH_ERR_INVALID, "bsh_scene_commit: light " + std::to_string(l) + " out of range");
        int tri = s->bundles[L.bundle].tri_begin + L.shape;
        const bsh_material& m = s->mats[s->tri_m[tri]];
        if ((int)m.type != BSH_EMISSIVE)
            return fail(BSH_ERR_INVALID, "AreaLight: The material passed is not an emissive one");
        bundle_light[L.bundle] = (int32_t)l;  // bundle_.arealight_ = this (last one wins)
        const float* V = VV.data();
        const float* V0 = s->verts.data();   // object space
        const int32_t* tv = &s->tri_v[3 * tri];
        f3 p1 = mk(V[3 * tv[0]], V[3 * tv[0] + 1], V[3 * tv[0] + 2]);
        f3 p2 = mk(V[3 * tv[1]], V[3 * tv[1] + 1], V[3 * tv[1] + 2]);
        f3 p3 = mk(V[3 * tv[2]], V[3 * tv[2] + 1], V[3 * tv[2] + 2]);
        f3 o1 = mk(V0[3 * tv[0]], V0[3 * tv[0] + 1], V0[3 * tv[0] + 2]);
        f3 o2 = mk(V0[3 * tv[1]], V0[3 * tv[1] + 1], V0[3 * tv[1] + 2]);
        f3 o3 = mk(V0[3 * tv[2]], V0[3 * tv[2] + 1], V0[3 * tv[2] + 2]);
        LightDev d;
        d.tri = tri;
        // GetShapeSurfaceArea takes the transformed vertices (mesh.cpp:304-321); FillSample's area
        // pdf the untransformed ones (mesh.cpp:121): they differ under a scaling transform
        d.area = sqrtf(fabsf(sqnorm(cross(p3 - p1, p3 - p2)))) * 0.5f;
        d.pdf = 1.f / (sqrtf(fabsf(sqnorm(cross(o3 - o1, o3 - o2)))) * 0.5f);
        d.type = LIGHT_AREA;
        d.e = make_float4(m.color[0], m.color[1], m.color[2], 0.f);
        const int32_t* tn = &s->tri_n[3 * tri];
        const float* N = NN.data();
        for (int k = 0; k < 3; ++k)
        {
            d.p[k] = make_float4(V[3 * tv[k]], V[3 * tv[k] + 1], V[3 * tv[k] + 2], 0.f);
            d.n[k] = make_float4(N[3 * tn[k]], N[3 * tn[k] + 1], N[3 * tn[k] + 2], 0.f);
        }
        s->h_lights.push_back(d);
    }
    // the other lights follow, as World::lights_ holds them after the area lights
    bool has_env = false, has_env_is = false;
    for (size_t l = 0; l < s->others.size(); ++l)
    {
        const bsh_light& O = s->others[l];
        LightDev d{};
        d.tri = -1;
        d.type = O.type;
        if (O.type == BSH_LIGHT_POINT || O.type == BSH_LIGHT_DIRECTIONAL)
        {
            f3 v = mk(O.v[0], O.v[1], O.v[2]);
            if (O.type == BSH_LIGHT_DIRECTIONAL) v = normalize(v);   // DirectionalLight ctor: d_(normalize(d))
            d.p[0] = make_float4(v.x, v.y, v.z, 0.f);
            d.e = make_float4(O.e[0], O.e[1], O.e[2], 0.f);
        }
        else
        {
            if (!(O.gamma != 0.f) || !std::isfinite(O.scale))
                return fail(BSH_ERR_INVALID, "bsh_scene_commit: environment light " + std::to_string(l) +
                                                 " needs a finite scale and a nonzero gamma");
            d.e = make_float4(O.scale, 1.f / O.gamma, 0.f, 0.f);   // invgamma_(1.f / gamma)
            has_env = true;
            has_env_is = has_env_is || O.type == BSH_LIGHT_ENVIRONMENT_IS;
        }
        s->h_lights.push_back(d);
    }
    if (has_env && s->env_tex.empty())
        return fail(BSH_ERR_STATE, "bsh_scene_commit: environment lights need bsh_scene_set_env_map");
    s->env = EnvDev{};
    if (has_env)
    {
        const size_t ntex = s->env_tex.size();
        const size_t nd = has_env_is ? (size_t)kEnvDistH * (2 + kEnvDistW + (kEnvDistW + 1)) + kEnvDistH + 1 : 0;
        std::vector<float> buf(ntex + nd);
        std::copy(s->env_tex.begin(), s->env_tex.end(), buf.begin());
        float mint = 0.f;
        if (has_env_is) mint = build_env_distribution(s->env_tex.data(), s->env_w, s->env_h, buf.data() + ntex);
        BSH_HIP(s->d_env.alloc(buf.size()));
        BSH_HIP(hipMemcpy(s->d_env.p, buf.data(), buf.size() * sizeof(float), hipMemcpyHostToDevice));
        s->env.tex = s->d_env.p;
        s->env.w = s->env_w;
        s->env.h = s->env_h;
        if (has_env_is)
        {
            float* q = s->d_env.p + ntex;
            s->env.mval = q;
            s->env.mcdf = q + kEnvDistH;
            s->env.cval = q + 2 * kEnvDistH + 1;
            s->env.ccdf = s->env.cval + (size_t)kEnvDistH * kEnvDistW;
            s->env.cint = s->env.ccdf + (size_t)kEnvDistH * (kEnvDistW + 1);
            s->env.mint = mint;
        }
    }
    // BVH
    std::vector<float> tv(9 * nt);
    for (size_t i = 0; i < nt; ++i)
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) tv[9 * i + 3 * k + c] = VV[3 * s->tri_v[3 * i + k] + c];
    BuildParams bp;
    if (params)
    {
        if (params->max_leaf > 0) bp.max_leaf = params->max_leaf;
        if (params->sah_bins > 0) bp.bins = params->sah_bins;
        bp.threads = params->threads;
        if (params->width != 0 && params->width != 2 && params->width != 4)
            return fail(BSH_ERR_INVALID, "bsh_scene_commit: BVH width must be 0, 2 or 4");
        if (params->width != 0) bp.width = params->width;
        if (params->quantized < -1 || params->quantized > 1)
            return fail(BSH_ERR_INVALID, "bsh_scene_commit: quantized must be -1, 0 or 1");
        bp.quantized = params->quantized;
    }
    BuildOutput bo;
    build_bvh(tv.data(), nt, bp, bo);
    std::vector<float4> tris(3 * std::max<size_t>(nt, 1));
    for (size_t k = 0; k < nt; ++k)
    {
        int prim = bo.tri_order[k];
        const float* p = &tv[9 * prim];
        f3 p1 = mk(p[0], p[1], p[2]), p2 = mk(p[3], p[4], p[5]), p3 = mk(p[6], p[7], p[8]);
        f3 e1 = p2 - p1, e2 = p3 - p1;   // mesh.cpp:13-14
        float pb;
        std::memcpy(&pb, &prim, 4);
        tris[3 * k + 0] = make_float4(p1.x, p1.y, p1.z, pb);
        tris[3 * k + 1] = make_float4(e1.x, e1.y, e1.z, 0.f);
        tris[3 * k + 2] = make_float4(e2.x, e2.y, e2.z, 0.f);
    }
    // per-primitive shading records: vertices, normals and material copied in (kShadeF4 float4)
    std::vector<float4> stri((size_t)kShadeF4 * std::max<size_t>(nt, 1), make_float4(0.f, 0.f, 0.f, 0.f));
    {
        auto v4 = [&](const std::vector<float>& a, int32_t k, float w) {
            return make_float4(a[3 * (size_t)k], a[3 * (size_t)k + 1], a[3 * (size_t)k + 2], w);
        };
        size_t b = 0;
        for (size_t i = 0; i < nt; ++i)
        {
            while (b + 1 < s->bundles.size() && (int32_t)i >= s->bundles[b].tri_begin + s->bundles[b].tri_count) ++b;
            const int32_t bl = s->bundles.empty() ? -1 : bundle_light[b];
            const bsh_material& m = s->mats[s->tri_m[i]];
            float4* r = &stri[(size_t)kShadeF4 * i];
            int32_t mi = s->tri_m[i];
            float mw, bw;
            std::memcpy(&mw, &mi, 4);
            std::memcpy(&bw, &bl, 4);
            r[0] = v4(VV, s->tri_v[3 * i], mw);
            r[1] = v4(VV, s->tri_v[3 * i + 1], bw);
            r[2] = v4(VV, s->tri_v[3 * i + 2], 0.f);
            r[3] = v4(NN, s->tri_n[3 * i], 0.f);
            r[4] = v4(NN, s->tri_n[3 * i + 1], 0.f);
            r[5] = v4(NN, s->tri_n[3 * i + 2], 0.f);
            r[6] = make_float4(m.type, m.color[0], m.color[1], m.color[2]);
            r[7] = make_float4(m.eta, m.alpha, m.pad[0], m.pad[1]);
        }
    }
    const size_t node_f4 = bo.width != 4 ? bo.nodes.size() * 4 : bo.quantized ? bo.nodesq.size() * 4 : bo.nodes4.size() * 8;
    const void* node_src = bo.width != 4 ? (const void*)bo.nodes.data()
                           : bo.quantized ? (const void*)bo.nodesq.data() : (const void*)bo.nodes4.data();
    BSH_HIP(s->d_nodes.alloc(node_f4));
    BSH_HIP(hipMemcpy(s->d_nodes.p, node_src, node_f4 * sizeof(float4), hipMemcpyHostToDevice));
    BSH_HIP(s->d_tris.alloc(tris.size()));
    BSH_HIP(hipMemcpy(s->d_tris.p, tris.data(), tris.size() * sizeof(float4), hipMemcpyHostToDevice));
    BSH_HIP(s->d_stri.alloc(stri.size()));
    BSH_HIP(hipMemcpy(s->d_stri.p, stri.data(), stri.size() * sizeof(float4), hipMemcpyHostToDevice));
    BSH_HIP(s->d_lights.alloc(std::max<size_t>(s->h_lights.size(), 1)));
    if (!s->h_lights.empty())
        BSH_HIP(hipMemcpy(s->d_lights.p, s->h_lights.data(), s->h_lights.size() * sizeof(LightDev),
                          hipMemcpyHostToDevice));
    // the light table's power CDF (quirk-free light picks): power ~ area x luminance(e), summed in
    // double; lights of no power are never picked unless every light has none (then uniform)
    {
        const size_t nl = s->h_lights.size();
        std::vector<double> pw(nl);
        double tot = 0.0;
        double area_tot = 0.0;
        size_t area_pos = 0;
        for (size_t l = 0; l < nl; ++l)
        {
            const LightDev& d = s->h_lights[l];
            if (d.type != LIGHT_AREA) continue;
            pw[l] = (double)d.area * std::max(0.0, 0.2126 * d.e.x + 0.7152 * d.e.y + 0.0722 * d.e.z);
            area_tot += pw[l];
            area_pos += pw[l] > 0.0 ? 1 : 0;
        }
        // point, directional and environment lights have no power comparable to an area light's;
        // each is picked like an average area light (any positive share keeps the estimator
        // unbiased, the pick's 1 / probability weighing it)
        for (size_t l = 0; l < nl; ++l)
            if (s->h_lights[l].type != LIGHT_AREA) pw[l] = area_pos ? area_tot / (double)area_pos : 1.0;
        for (size_t l = 0; l < nl; ++l) tot += pw[l];
        std::vector<float> cdf(std::max<size_t>(nl, 1), 1.f);
        double run = 0.0;
        for (size_t l = 0; l < nl; ++l)
        {
            run += tot > 0.0 ? pw[l] : 1.0;
            cdf[l] = l + 1 == nl ? 1.f : (float)(run / (tot > 0.0 ? tot : (double)nl));
        }
        BSH_HIP(s->d_light_cdf.alloc(cdf.size()));
        BSH_HIP(hipMemcpy(s->d_light_cdf.p, cdf.data(), cdf.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    for (int k = 0; k < 3; ++k)
    {
        s->box_lo[k] = std::numeric_limits<float>::max();
        s->box_hi[k] = -std::numeric_limits<float>::max();
    }
    for (size_t i = 0; i < VV.size() / 3; ++i)
        for (int k = 0; k < 3; ++k)
        {
            s->box_lo[k] = std::min(s->box_lo[k], VV[3 * i + k]);
            s->box_hi[k] = std::max(s->box_hi[k], VV[3 * i + k]);
        }
    s->nnodes = bo.width == 4 ? bo.nodes4.size() : bo.nodes.size();
    s->width = bo.width;
    s->quantized = bo.width == 4 && bo.quantized;
    s->max_depth = bo.max_depth;
    s->build_ms = bo.build_ms;
    s->committed = true;
    return BSH_OK;
}

bsh_status bsh_scene_destroy(bsh_scene* s)
{
    if (s)
    {
        (void)hipSetDevice(s->ctx->device);
        delete s;
    }
    return BSH_OK;
}

bsh_status bsh_bvh_build_host(const float* tri_verts, size_t ntris, const bsh_bvh_params* params, void* nodes,
                              size_t nodes_cap, int32_t* tri_order, size_t* nnodes, int32_t* width, int32_t* depth,
                              int32_t* node_bytes)
{
    if (!tri_verts || ntris == 0 || !nnodes) return fail(BSH_ERR_INVALID, "bsh_bvh_build_host: bad argument");
    BuildParams bp;
    if (params)
    {
        if (params->max_leaf > 0) bp.max_leaf = params->max_leaf;
        if (params->sah_bins > 0) bp.bins = params->sah_bins;
        bp.threads = params->threads;
        if (params->width != 0 && params->width != 2 && params->width != 4)
            return fail(BSH_ERR_INVALID, "bsh_bvh_build_host: BVH width must be 0, 2 or 4");
        if (params->width != 0) bp.width = params->width;
        if (params->quantized < -1 || params->quantized > 1)
            return fail(BSH_ERR_INVALID, "bsh_bvh_build_host: quantized must be -1, 0 or 1");
        bp.quantized = params->quantized;
    }
    BuildOutput bo;
    build_bvh(tri_verts, ntris, bp, bo);
    const bool q = bo.width == 4 && bo.quantized;
    const size_t n = bo.width == 4 ? bo.nodes4.size() : bo.nodes.size();
    const size_t nb = bo.width != 4 ? sizeof(GpuNode) : q ? sizeof(GpuNodeQ) : sizeof(GpuNode4);
    *nnodes = n;
    if (width) *width = bo.width;
    if (node_bytes) *node_bytes = (int32_t)nb;
    if (depth) *depth = bo.max_depth;
    if (nodes)
    {
        if (nodes_cap < n) return fail(BSH_ERR_INVALID, "bsh_bvh_build_host: node buffer too small");
        std::memcpy(nodes, bo.width != 4 ? (const void*)bo.nodes.data()
                           : q ? (const void*)bo.nodesq.data() : (const void*)bo.nodes4.data(), n * nb);
    }
    if (tri_order) std::memcpy(tri_order, bo.tri_order.data(), ntris * sizeof(int32_t));
    return BSH_OK;
}

bsh_status bsh_scene_info(const bsh_scene* s, uint64_t* nodes, uint64_t* tris, uint64_t* bytes, double* build_ms,
                          int32_t* depth, int32_t* width, int32_t* node_bytes)
{
    if (!s) return fail(BSH_ERR_INVALID, "bsh_scene_info: NULL scene");
    if (nodes) *nodes = s->nnodes;
    if (tris) *tris = s->tri_m.size();
    if (bytes)
        *bytes = s->nnodes * (s->width != 4 ? sizeof(GpuNode) : s->quantized ? sizeof(GpuNodeQ) : sizeof(GpuNode4)) +
                 s->tri_m.size() * (48 + 32) +
                 (s->verts.size() + s->normals.size()) / 3 * 16;
    if (build_ms) *build_ms = s->build_ms;
    if (depth) *depth = s->max_depth;
    if (width) *width = s->width;
    if (node_bytes) *node_bytes = s->width != 4 ? (int32_t)sizeof(GpuNode) : s->quantized ? (int32_t)sizeof(GpuNodeQ)
                                                                                       : (int32_t)sizeof(GpuNode4);
    return BSH_OK;
}

bsh_status bsh_intersect_closest(bsh_scene* s, const bsh_ray* rays, size_t n, bsh_hit* hits, void* stream)
{
    return launch_intersect(s, rays, n, 0, hits, nullptr, nullptr, (hipStream_t)stream);
}

bsh_status bsh_intersect_any(bsh_scene* s, const bsh_ray* rays, size_t n, int32_t* occluded, void* stream)
{
    return launch_intersect(s, rays, n, 1, nullptr, occluded, nullptr, (hipStream_t)stream);
}

bsh_status bsh_intersect_count(bsh_scene* s, const bsh_ray* rays, size_t n, int any, bsh_hit* hits, uint32_t* visits,
                               void* stream)
{
    return launch_intersect(s, rays, n, any ? 1 : 0, hits, nullptr, visits, (hipStream_t)stream);
}

bsh_status bsh_render(bsh_scene* s, const bsh_render_params* p, float* accum, float* samples, void* stream)
{
    JobDev J;
    bsh_status r = make_render_job(s, p, accum, samples, J);
    if (r != BSH_OK) return r;
    std::vector<JobDev> jobs(1, J);
    return run_jobs(s->ctx, jobs, (p->flags & BSH_RENDER_COUNT) != 0, nullptr, 0, (hipStream_t)stream);
}

bsh_status bsh_bake_sh(bsh_scene* s, const float* probes, size_t nprobes, const bsh_probe_params* p, float* coeffs,
                       float* radiance, void* stream)
{
    JobDev J;
    bsh_status r = make_probe_job(s, probes, nprobes, p, coeffs, radiance, J);
    if (r != BSH_OK) return r;
    if (nprobes == 0) return BSH_OK;
    std::vector<JobDev> jobs(1, J);
    return run_jobs(s->ctx, jobs, (p->flags & BSH_RENDER_COUNT) != 0, nullptr, 0, (hipStream_t)stream);
}

// Runs one quirk-free job (a camera frame or a probe set) whose continuation samples the
// BSDF/guide mixture when guide is set, appending training records when records is set.
static bsh_status run_guided_job(bsh_scene* s, JobDev& J, bool count, bsh_guide* guide, float* records, size_t cap,
                                 size_t* nrecords, hipStream_t st)
{
    if (guide)
    {
        J.flags |= JOB_GUIDED;
        J.bsdf_frac = guide_bsdf_fraction(guide);
        J.guide_bounces = guide_bounces(guide);
    }
    if (records && cap) J.flags |= JOB_RECORD;
    GuideRun gr;
    gr.guide = guide;
    gr.records = records;
    gr.cap = cap;
    gr.bounces = guide ? guide_bounces(guide) : 0;
    std::vector<JobDev> jobs(1, J);
    bsh_status r = run_jobs(s->ctx, jobs, count, nullptr, 0, st, &gr);
    if (nrecords) *nrecords = gr.written;
    return r;
}

bsh_status bsh_render_guided(bsh_scene* s, const bsh_render_params* p, bsh_guide* guide, float* accum,
                             float* records, size_t cap, size_t* nrecords, void* stream)
{
    if (nrecords) *nrecords = 0;
    if (!p || !(p->flags & BSH_RENDER_NO_QUIRKS))
        return fail(BSH_ERR_INVALID, "bsh_render_guided: needs BSH_RENDER_NO_QUIRKS (the guided estimator is quirk-free)");
    if (cap && !records) return fail(BSH_ERR_INVALID, "bsh_render_guided: NULL records with nonzero capacity");
    JobDev J;
    bsh_status r = make_render_job(s, p, accum, nullptr, J);
    if (r != BSH_OK) return r;
    return run_guided_job(s, J, (p->flags & BSH_RENDER_COUNT) != 0, guide, records, cap, nrecords, (hipStream_t)stream);
}

bsh_status bsh_bake_sh_guided(bsh_scene* s, const float* probes, size_t nprobes, const bsh_probe_params* p,
                              bsh_guide* guide, float* coeffs, float* radiance, float* ray_dirs, float* records,
                              size_t cap, size_t* nrecords, void* stream)
{
    if (nrecords) *nrecords = 0;
    if (!p || !(p->flags & BSH_RENDER_NO_QUIRKS))
        return fail(BSH_ERR_INVALID, "bsh_bake_sh_guided: needs BSH_RENDER_NO_QUIRKS (the guided estimator is quirk-free)");
    if (cap && !records) return fail(BSH_ERR_INVALID, "bsh_bake_sh_guided: NULL records with nonzero capacity");
    JobDev J;
    bsh_status r = make_probe_job(s, probes, nprobes, p, coeffs, radiance, J);
    if (r != BSH_OK) return r;
    if (nprobes == 0) return BSH_OK;
    J.ray_dirs = ray_dirs;
    if (guide) J.flags |= JOB_PROBE_DIR;
    return run_guided_job(s, J, (p->flags & BSH_RENDER_COUNT) != 0, guide, records, cap, nrecords, (hipStream_t)stream);
}

bsh_status bsh_render_stats_get(const bsh_scene* s, bsh_render_stats* out)
{
    if (!s || !out) return fail(BSH_ERR_INVALID, "bsh_render_stats_get: NULL argument");
    *out = s->ctx->stats;
    return BSH_OK;
}

bsh_status bsh_context_stats_get(const bsh_context* ctx, bsh_render_stats* out)
{
    if (!ctx || !out) return fail(BSH_ERR_INVALID, "bsh_context_stats_get: NULL argument");
    *out = ctx->stats;
    return BSH_OK;
}

bsh_status bsh_context_launch_stats(const bsh_context* ctx, bsh_launch_stats* out, size_t cap, size_t* n)
{
    if (!ctx || !n || (cap && !out)) return fail(BSH_ERR_INVALID, "bsh_context_launch_stats: NULL argument");
    *n = ctx->launch_stats.size();
    for (size_t k = 0; k < std::min(cap, ctx->launch_stats.size()); ++k) out[k] = ctx->launch_stats[k];
    return BSH_OK;
}

bsh_status bsh_context_set_tuning(bsh_context* ctx, const bsh_tuning* t)
{
    if (!ctx || !t) return fail(BSH_ERR_INVALID, "bsh_context_set_tuning: NULL argument");
    bsh_status r = validate_tuning(*t, "bsh_context_set_tuning");
    if (r != BSH_OK) return r;
    ctx->tuning = *t;
    return BSH_OK;
}

bsh_status bsh_context_get_tuning(const bsh_context* ctx, bsh_tuning* t)
{
    if (!ctx || !t) return fail(BSH_ERR_INVALID, "bsh_context_get_tuning: NULL argument");
    *t = ctx->tuning;
    return BSH_OK;
}

static bsh_status build_jobs(bsh_context* ctx, const bsh_job* jobs, size_t njobs, std::vector<JobDev>& out, bool& count)
{
    out.clear();
    count = false;
    for (size_t k = 0; k < njobs; ++k)
    {
        const bsh_job& j = jobs[k];
        if (!j.scene || j.scene->ctx != ctx)
            return fail(BSH_ERR_INVALID, "bsh_batch_run: job " + std::to_string(k) + " has no scene of this context");
        JobDev J;
        bsh_status r;
        if (j.kind == BSH_JOB_RENDER)
        {
            r = make_render_job(j.scene, &j.render, j.accum, j.samples, J);
            count = count || (j.render.flags & BSH_RENDER_COUNT);
        }
        else if (j.kind == BSH_JOB_BAKE_SH)
        {
            r = make_probe_job(j.scene, j.probes, j.nprobes, &j.probe, j.coeffs, j.radiance, J);
            if (r == BSH_OK && j.nprobes == 0) continue;
        }
        else
        {
            return fail(BSH_ERR_INVALID, "bsh_batch_run: unknown job kind");
        }
        if (r != BSH_OK) return r;
        out.push_back(J);
    }
    return BSH_OK;
}

bsh_status bsh_batch_run(bsh_context* ctx, const bsh_job* jobs, size_t njobs, const bsh_tuning* tuning, void* stream)
{
    if (!ctx || (njobs && !jobs)) return fail(BSH_ERR_INVALID, "bsh_batch_run: NULL argument");
    std::vector<JobDev> hj;
    bool count = false;
    bsh_status r = build_jobs(ctx, jobs, njobs, hj, count);
    if (r != BSH_OK) return r;
    const bsh_tuning& t = tuning ? *tuning : ctx->tuning;
    r = validate_tuning(t, "bsh_batch_run");
    if (r != BSH_OK) return r;
    return run_jobs(ctx, hj, count, &t, t.use_graph, (hipStream_t)stream);
}

bsh_status bsh_autotune(bsh_context* ctx, const bsh_job* jobs, size_t njobs, const int32_t* paths_per_chunk,
                        size_t n_paths, const int32_t* blocks_per_cu, size_t n_blocks, bsh_tuning* best,
                        double* best_ms)
{
    if (!ctx || (njobs && !jobs)) return fail(BSH_ERR_INVALID, "bsh_autotune: NULL argument");
    static const int32_t kPaths[] = {1 << 21, 1 << 22, 1 << 23, 1 << 24};
    static const int32_t kBlocks[] = {2, 3, 4, 5};
    if (!paths_per_chunk || n_paths == 0) { paths_per_chunk = kPaths; n_paths = 4; }
    if (!blocks_per_cu || n_blocks == 0) { blocks_per_cu = kBlocks; n_blocks = 4; }
    std::vector<JobDev> hj;
    bool count = false;
    bsh_status r = build_jobs(ctx, jobs, njobs, hj, count);
    if (r != BSH_OK) return r;
    bsh_tuning bt = ctx->tuning;
    double bms = 1e300;
    for (size_t a = 0; a < n_paths; ++a)
        for (size_t b = 0; b < n_blocks; ++b)
        {
            if (blocks_per_cu[b] < 1 || blocks_per_cu[b] > 5 || paths_per_chunk[a] < 1)
                return fail(BSH_ERR_INVALID, "bsh_autotune: candidate out of range");
            bsh_tuning t{paths_per_chunk[a], blocks_per_cu[b], 0, ctx->tuning.refill_lanes, ctx->tuning.claim_rays,
                          ctx->tuning.lanes};
            r = validate_tuning(t, "bsh_autotune");
            if (r != BSH_OK) return r;
            std::vector<JobDev> copy = hj;
            r = run_jobs(ctx, copy, false, &t, 0, nullptr);
            if (r != BSH_OK) return r;
            double ms = ctx->stats.total_ms;
            if (ms < bms)
            {
                bms = ms;
                bt = t;
            }
        }
    bt.use_graph = ctx->tuning.use_graph;
    ctx->tuning = bt;
    if (best) *best = bt;
    if (best_ms) *best_ms = bms;
    return BSH_OK;
}

}  // extern "C"
