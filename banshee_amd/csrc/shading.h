// Scene-dependent shading helpers shared by the render and diagnostic kernels.
#pragma once

#include "scene_dev.h"

namespace bsh {

__device__ __forceinline__ Mat load_mat(const SceneDev& s, int m)
{
    float4 a = s.mats[2 * m], b = s.mats[2 * m + 1];
    Mat r;
    r.type = (int)a.x;
    r.c = mk(a.y, a.z, a.w);
    r.eta = b.x;
    r.alpha = b.y;
    return r;
}

__device__ __forceinline__ void tri_verts(const SceneDev& s, int prim, f3& p1, f3& p2, f3& p3, f3& n1, f3& n2, f3& n3,
                                          int& mat, int& blight)
{
    int4 a = s.shade[2 * prim], b = s.shade[2 * prim + 1];
    p1 = ld3(s.verts[a.x]);
    p2 = ld3(s.verts[a.y]);
    p3 = ld3(s.verts[a.z]);
    n1 = ld3(s.normals[b.x]);
    n2 = ld3(s.normals[b.y]);
    n3 = ld3(s.normals[b.z]);
    mat = a.w;
    blight = b.w;
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
