// Device-resident scene and the shared traversal routine.
#pragma once

#include <hip/hip_runtime.h>

#include "bsh_math.h"
#include "bvh_build.h"

namespace bsh {

// One area light (AreaLight, light/arealight.h:43-81): emissive triangle `tri`,
// emission e of its material, and the triangle's area and area pdf exactly as
// Mesh::GetShapeSurfaceArea / FillSample compute them (mesh.cpp:304-321, 121).
struct LightDev
{
    int tri;
    float area;
    float pdf;
    float pad;
    float4 e;
};

struct SceneDev
{
    const float4* nodes;      // GpuNode as 4 x float4
    const float4* tris;       // leaf order: (p1, prim bits), (e1, 0), (e2, 0)
    const int4* shade;        // per prim: (v0, v1, v2, material), (n0, n1, n2, bundle light)
    const float4* verts;
    const float4* normals;
    const float4* mats;       // per material: (type, c.xyz), (eta, alpha, 0, 0)
    const LightDev* lights;
    int nlights;
    int ntris;
};

struct TraceResult
{
    int prim;
    float t, b1, b2;
};

struct TraceCounts
{
    uint32_t nodes, tris;
};

// Stack-based BVH2 traversal with both child boxes in the parent, near child first,
// leaves intersected as soon as their box is hit, and a per-lane LDS stack of
// kMaxStack entries laid out [entry][lane] (stride = blockDim.x) so a wave's pushes
// and pops hit 64 distinct banks. Closest hit (any == false) replaces
// Bvh::Intersect(ray, Hit&) (bvh.cpp:499-573): accept tmin < t <= current t; any hit
// replaces Bvh::Intersect(ray) (bvh.cpp:575-647) and returns at the first accept.
template <bool kAny, bool kCount>
__device__ __forceinline__ bool traverse(const SceneDev& s, f3 o, f3 d, float tmin, float tmax, uint32_t* stk,
                                         int stride, TraceResult& res, TraceCounts& cnt)
{
    // Slabs are evaluated as b*inv - o*inv with one FMA; a zero direction component
    // would give inf - inf = NaN there, so the box test uses |d| >= 1e-20 (the
    // triangle test keeps the exact direction).
    const float dx = fabsf(d.x) < 1e-20f ? copysignf(1e-20f, d.x) : d.x;
    const float dy = fabsf(d.y) < 1e-20f ? copysignf(1e-20f, d.y) : d.y;
    const float dz = fabsf(d.z) < 1e-20f ? copysignf(1e-20f, d.z) : d.z;
    const float ix = 1.f / dx, iy = 1.f / dy, iz = 1.f / dz;
    const float ox = o.x * ix, oy = o.y * iy, oz = o.z * iz;
    float tcur = tmax;
    int prim = -1;
    float hb1 = 0.f, hb2 = 0.f;
    uint32_t node = 0;
    int sp = 0;
    for (;;)
    {
        const float4 a = s.nodes[4 * node + 0];
        const float4 b = s.nodes[4 * node + 1];
        const float4 c = s.nodes[4 * node + 2];
        const int4 r = reinterpret_cast<const int4*>(s.nodes)[4 * node + 3];
        if (kCount) cnt.nodes++;
        // child 0
        float lx0 = __builtin_fmaf(a.x, ix, -ox), hx0 = __builtin_fmaf(a.y, ix, -ox);
        float ly0 = __builtin_fmaf(a.z, iy, -oy), hy0 = __builtin_fmaf(a.w, iy, -oy);
        float lz0 = __builtin_fmaf(c.x, iz, -oz), hz0 = __builtin_fmaf(c.y, iz, -oz);
        float tn0 = fmaxf(fmaxf(fminf(lx0, hx0), fminf(ly0, hy0)), fmaxf(fminf(lz0, hz0), tmin));
        float tf0 = fminf(fminf(fmaxf(lx0, hx0), fmaxf(ly0, hy0)), fminf(fmaxf(lz0, hz0), tcur));
        // child 1
        float lx1 = __builtin_fmaf(b.x, ix, -ox), hx1 = __builtin_fmaf(b.y, ix, -ox);
        float ly1 = __builtin_fmaf(b.z, iy, -oy), hy1 = __builtin_fmaf(b.w, iy, -oy);
        float lz1 = __builtin_fmaf(c.z, iz, -oz), hz1 = __builtin_fmaf(c.w, iz, -oz);
        float tn1 = fmaxf(fmaxf(fminf(lx1, hx1), fminf(ly1, hy1)), fmaxf(fminf(lz1, hz1), tmin));
        float tf1 = fminf(fminf(fmaxf(lx1, hx1), fmaxf(ly1, hy1)), fminf(fmaxf(lz1, hz1), tcur));
        bool h0 = (tn0 <= tf0) && (r.z >= 0);
        bool h1 = (tn1 <= tf1) && (r.w >= 0);
        // leaves
#pragma unroll
        for (int k = 0; k < 2; ++k)
        {
            bool hk = k == 0 ? h0 : h1;
            int cnt_k = k == 0 ? r.z : r.w;
            if (hk && cnt_k > 0)
            {
                int first = k == 0 ? r.x : r.y;
                for (int q = first; q < first + cnt_k; ++q)
                {
                    const float4 v0 = s.tris[3 * q + 0];
                    const float4 e1 = s.tris[3 * q + 1];
                    const float4 e2 = s.tris[3 * q + 2];
                    if (kCount) cnt.tris++;
                    float t, u, v;
                    if (intersect_face(ld3(v0), ld3(e1), ld3(e2), o, d, tmin, kAny ? tmax : tcur, t, u, v))
                    {
                        prim = __builtin_bit_cast(int, v0.w);
                        if (kAny)
                        {
                            res.prim = prim;
                            res.t = t;
                            res.b1 = u;
                            res.b2 = v;
                            return true;
                        }
                        tcur = t;
                        hb1 = u;
                        hb2 = v;
                    }
                }
                if (k == 0) h0 = false; else h1 = false;
            }
        }
        if (h0 && h1)
        {
            bool first0 = tn0 <= tn1;
            uint32_t nearc = first0 ? (uint32_t)r.x : (uint32_t)r.y;
            uint32_t farc = first0 ? (uint32_t)r.y : (uint32_t)r.x;
            // The builder caps internal depth at kMaxStack - 2, so the stack never fills;
            // were it to, the far child is dropped rather than written out of bounds.
            if (sp < kMaxStack) stk[(sp++) * stride] = farc;
            node = nearc;
        }
        else if (h0)
        {
            node = (uint32_t)r.x;
        }
        else if (h1)
        {
            node = (uint32_t)r.y;
        }
        else
        {
            if (sp == 0) break;
            node = stk[(--sp) * stride];
        }
    }
    res.prim = prim;
    res.t = tcur;
    res.b1 = hb1;
    res.b2 = hb2;
    return prim >= 0;
}

}  // namespace bsh
