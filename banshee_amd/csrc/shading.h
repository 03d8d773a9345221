// Scene-dependent shading helpers shared by the render and diagnostic kernels.
#pragma once

#include "material_dev.h"
#include "scene_dev.h"

namespace bsh {

// A primitive's shading record: its three vertices, three normals and its material, copied
// from the scene's arrays at commit. The loads of one record are independent of each other,
// so a hit reaches everything shading reads after one level of loads.
__device__ __forceinline__ void tri_verts(const SceneDev& s, int prim, f3& p1, f3& p2, f3& p3, f3& n1, f3& n2, f3& n3,
                                          Mat& m)
{
    const float4* r = s.stri + (size_t)kShadeF4 * prim;
    const float4 a = r[0], b = r[1], c = r[2], d = r[3], e = r[4], f = r[5], g = r[6], h = r[7];
    p1 = ld3(a);
    p2 = ld3(b);
    p3 = ld3(c);
    n1 = ld3(d);
    n2 = ld3(e);
    n3 = ld3(f);
    m.type = (int)g.x;
    m.c = mk(g.y, g.z, g.w);
    m.eta = h.x;
    m.alpha = h.y;
    m.x = h.z;
    m.y = h.w;
}

// The area light of the primitive's bundle, -1 when it has none.
__device__ __forceinline__ int tri_bundle_light(const SceneDev& s, int prim)
{
    return __float_as_int(s.stri[(size_t)kShadeF4 * prim + 1].w);
}

// Quirk-free light pick: the first light whose cumulative power share exceeds u in [0, 1)
// (binary search over the scene's CDF); w receives 1 / its probability.
__device__ __forceinline__ int pick_light(const SceneDev& s, float u, float& w)
{
    int lo = 0, hi = s.nlights - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        if (s.light_cdf[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    const float p = s.light_cdf[lo] - (lo > 0 ? s.light_cdf[lo - 1] : 0.f);
    w = 1.f / p;
    return lo;
}

// Light::GetPdf of the non-area lights at a point with shading normal n, direction w:
// point and directional lights 0 (light.h, pointlight.h, directional_light.h);
// EnvironmentLight n.normalize(w) / pi (environment_light.cpp:64-70); EnvironmentLightIs the
// distribution's density in solid angle (environment_light_is.cpp:94-107).
__device__ __forceinline__ float other_light_pdf(const SceneDev& s, const LightDev& L, f3 n, f3 w)
{
    if (L.type == LIGHT_ENV) return dot(n, normalize(w)) * (1.f / kPi);
    if (L.type == LIGHT_ENV_IS)
    {
        float u, v, theta;
        dir_to_uv(w, u, v, &theta);
        float sintheta = sinf(theta);
        return sintheta == 0.f ? 0.f : dist2d_pdf(s.env, u, v) / (2.f * kPi * kPi * sintheta);
    }
    return 0.f;
}

// Light::GetSample of the non-area lights (pointlight.cpp, directional_light.cpp,
// environment_light.cpp:9-35, environment_light_is.cpp:49-73) at a point p with shading
// normal n and sample (ux, uy): d is the (long) vector to the light, pdf its density.
__device__ __forceinline__ f3 other_light_sample(const SceneDev& s, const LightDev& L, f3 p, f3 n, float ux, float uy,
                                                 f3& d, float& pdf)
{
    if (L.type == LIGHT_POINT)
    {
        d = ld3(L.p[0]) - p;
        pdf = 1.f;
        float d2inv = 1.f / sqnorm(d);
        return ld3(L.e) * d2inv;
    }
    if (L.type == LIGHT_DIRECTIONAL)
    {
        d = -ld3(L.p[0]) * 100000000.f;
        pdf = 1.f;
        return ld3(L.e);
    }
    float u, v;
    if (L.type == LIGHT_ENV)
    {
        d = map_to_hemisphere_cos(n, ux, uy);
        pdf = dot(n, d) * (1.f / kPi);
        dir_to_uv(d, u, v);
        d = d * 10000000.f;
    }
    else
    {
        float dpdf = 0.f;
        dist2d_sample(s.env, ux, uy, u, v, dpdf);
        float theta = v * kPi;
        float phi = u * 2.f * kPi;
        float costheta = cosf(theta), sintheta = sinf(theta);
        float sinphi = sinf(phi), cosphi = cosf(phi);
        d = 100000000.f * normalize(mk(sintheta * cosphi, costheta, sintheta * sinphi));
        pdf = sintheta == 0.f ? 0.f : dpdf / (2.f * kPi * kPi * sintheta);
    }
    return env_radiance(s.env, L.e.x, L.e.y, u, v);
}

// Light::GetLe of a ray that left the scene in direction d: environment lights only
// (environment_light.cpp:37-50); every other light returns 0.
__device__ __forceinline__ f3 other_light_le(const SceneDev& s, const LightDev& L, f3 d)
{
    if (L.type != LIGHT_ENV && L.type != LIGHT_ENV_IS) return mk(0.f, 0.f, 0.f);
    float u, v;
    dir_to_uv(d, u, v);
    return env_radiance(s.env, L.e.x, L.e.y, u, v);
}

// AreaLight::GetPdf -> Mesh::GetPdfOnShape (arealight.cpp:55-58, mesh.cpp:354-378)
__device__ __forceinline__ float light_pdf(const LightDev* Lp, f3 p, f3 w)
{
    const LightDev& L = *Lp;
    const f3 p1 = ld3(L.p[0]), p2 = ld3(L.p[1]), p3 = ld3(L.p[2]);
    const f3 n1 = ld3(L.n[0]), n2 = ld3(L.n[1]), n3 = ld3(L.n[2]);
    float t, a, b;
    if (!intersect_face(p1, p2 - p1, p3 - p1, p, w, 0.001f, 100000.f, t, a, b)) return 0.f;
    Frame h = fill_hit(p1, p2, p3, n1, n2, n3, a, b);
    if (dot(-w, h.n) > 0.f)
    {
        f3 d = p - h.p;
        return sqnorm(d) / (dot(normalize(d), h.n) * L.area);
    }
    return 0.f;
}

// AreaLight::GetSample -> Mesh::GetSampleOnShape/FillSample (arealight.cpp:25-53, mesh.cpp:99-123,323-336)
__device__ __forceinline__ f3 light_sample(const LightDev* Lp, f3 p, float ux, float uy, f3& d, float& pdf,
                                          float* cos_l = nullptr)
{
    const LightDev& L = *Lp;
    const f3 p1 = ld3(L.p[0]), p2 = ld3(L.p[1]), p3 = ld3(L.p[2]);
    const f3 n1 = ld3(L.n[0]), n2 = ld3(L.n[1]), n3 = ld3(L.n[2]);
    float a = sqrtf(ux) * (1.f - uy);
    float b = sqrtf(ux) * uy;
    float w = 1.f - a - b;
    f3 sp = w * p1 + a * p2 + b * p3;
    f3 sn = normalize(w * n1 + a * n2 + b * n3);
    d = sp - p;
    float cl = dot(sn, -normalize(d));
    if (cos_l) *cos_l = cl;
    if (L.pdf > 0.f && cl > 0.f)
    {
        pdf = L.pdf;
        float d2inv = 1.f / sqnorm(d);
        return emissive_le(ld3(L.e), sn, -normalize(d)) * d2inv;
    }
    pdf = 0.f;
    return f3{0.f, 0.f, 0.f};
}

}  // namespace bsh
