// Device-resident scene and the shared traversal routine.
#pragma once

#include <hip/hip_runtime.h>

#include "bsh_math.h"
#include "bvh_build.h"
#include "light_dev.h"

namespace bsh {

// Device-memory views for hot loads and the traversal stack: through a global (or
// LDS) address-space pointer the compiler emits global (or LDS) instructions; through a
// generic pointer it emits flat ones, which count against both the memory and the LDS wait
// counters, so the stack's LDS accesses would wait on outstanding node fetches.
typedef float nf4 __attribute__((ext_vector_type(4)));
typedef int ni4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const nf4 gnf4;
typedef __attribute__((address_space(1))) const ni4 gni4;
__device__ __forceinline__ float4 gload(const float4* p, size_t i)
{
    const nf4 v = ((const gnf4*)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 gload_i(const float4* p, size_t i)
{
    const ni4 v = ((const gni4*)p)[i];
    return make_int4(v.x, v.y, v.z, v.w);
}
typedef __attribute__((address_space(1))) uint32_t gu32;
typedef __attribute__((address_space(3))) uint32_t lu32;

// One area light (AreaLight, light/arealight.h:43-81): emissive triangle `tri`,
// emission e of its material, and the triangle's area and area pdf exactly as
// Mesh::GetShapeSurfaceArea / FillSample compute them (mesh.cpp:304-321, 121).
struct LightDev
{
    int tri;                  // area lights: the emissive triangle; other lights: -1
    float area;
    float pdf;
    int type;                 // LIGHT_* (light_dev.h)
    // area, point and directional lights: emission; environment lights: (scale, 1 / gamma)
    float4 e;
    // the light triangle's vertices and vertex normals, copied from the scene's arrays so a
    // light sample or pdf reads them with one load instead of going through the triangle's
    // index record
    // area lights: triangle vertices and normals; point light: p[0] = position; directional
    // light: p[0] = its normalised direction
    float4 p[3];
    float4 n[3];
};

constexpr int kShadeF4 = 8;   // float4 per primitive shading record

struct SceneDev
{
    const float4* nodes;      // GpuNode as 4 x float4
    const float4* tris;       // leaf order: (p1, prim bits), (e1, 0), (e2, 0)
    const float4* stri;       // per prim, kShadeF4 float4: p1 (w: material index), p2 (w: bundle
                              // light), p3, n1, n2, n3, its material (type, c.xyz), (eta, alpha, 0, 0)
    const LightDev* lights;
    const float* light_cdf;   // per light: cumulative share of the emitted power (area x luminance), last = 1
    EnvDev env;               // environment map of the environment lights (tex NULL when none)
    const float4* mats;       // material rows, 2 float4 each: (type, c), (eta, alpha, x, y)
    const float4* tri_uv;     // per primitive: (u1, v1, u2, v2), (u3, v3, 0, 0); NULL without UVs
    const float* tex;         // texture maps' RGB texels
    const int4* tex_desc;     // per map: width, height, first float (low, high 32 bits)
    const int2* mat_maps;     // per material row: colour map, kr map (-1: none); NULL when none
    const int* nmaps;         // per material row: normal map (-1: none); NULL when no row has one
    int nlights;
    int narea;                // lights [0, narea) are area lights, the rest point/directional/environment
    int ext_mats;             // 1 when a material row is of type 3-7 (material_dev.h)
    int ntris;
    int width;                // BVH layout: 2 (GpuNode, 4 float4) or 4 (GpuNode4, 8 float4)
    int qnodes;               // width 4: 1 when the nodes are GpuNodeQ (4 float4)
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
// leaves intersected as soon as their box is hit, and a per-lane stack of kMaxStack
// entries: the first kLds in LDS laid out [entry][lane] (stride = blockDim.x) so a wave's
// pushes and pops hit 64 distinct banks, the rest (deep entries, rarely touched) in the
// lane's own global slots `ovf` (stride ostride). A smaller LDS part lets more waves share
// a CU. Closest hit (any == false) replaces Bvh::Intersect(ray, Hit&) (bvh.cpp:499-573):
// accept tmin < t <= current t; any hit replaces Bvh::Intersect(ray) (bvh.cpp:575-647) and
// returns at the first accept.
template <bool kAny, bool kCount, int kLds = kMaxStack>
__device__ __forceinline__ bool traverse(const SceneDev& s, f3 o, f3 d, float tmin, float tmax, uint32_t* stk,
                                         int stride, TraceResult& res, TraceCounts& cnt, uint32_t* ovf = nullptr,
                                         int ostride = 0)
{
    static_assert(kLds >= 1 && kLds <= kMaxStack, "LDS stack part out of range");
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
    // the stack is in LDS and its overflow in device memory (see gload)
    lu32* const lstk = (lu32*)stk;
    gu32* const govf = (gu32*)ovf;
    for (;;)
    {
        const float4 a = gload(s.nodes, 4 * (size_t)node + 0);
        const float4 b = gload(s.nodes, 4 * (size_t)node + 1);
        const float4 c = gload(s.nodes, 4 * (size_t)node + 2);
        const int4 r = gload_i(s.nodes, 4 * (size_t)node + 3);
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
                    const float4 v0 = gload(s.tris, 3 * (size_t)q + 0);
                    const float4 e1 = gload(s.tris, 3 * (size_t)q + 1);
                    const float4 e2 = gload(s.tris, 3 * (size_t)q + 2);
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
            if (sp < kLds) lstk[sp * stride] = farc;
            else if (kLds < kMaxStack && sp < kMaxStack) govf[(sp - kLds) * ostride] = farc;
            if (sp < kMaxStack) ++sp;
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
            --sp;
            node = (kLds == kMaxStack || sp < kLds) ? lstk[sp * stride] : govf[(sp - kLds) * ostride];
        }
    }
    res.prim = prim;
    res.t = tcur;
    res.b1 = hb1;
    res.b2 = hb2;
    return prim >= 0;
}

// Stack push/pop over the LDS part (kLds entries) and the global overflow slots.
// The pointer types are template parameters so a kernel can pass address-space-qualified
// pointers (LDS for stk, global for ovf) and get LDS and global instructions instead of flat.
template <int kLds, int kCap, class SP, class OP>
__device__ __forceinline__ void stack_push(SP stk, int stride, OP ovf, int ostride, int& sp, uint32_t v)
{
    if (sp < kLds) stk[sp * stride] = v;
    else if (kLds < kCap && sp < kCap) ovf[(sp - kLds) * ostride] = v;
    if (sp < kCap) ++sp;
}

template <int kLds, int kCap, class SP, class OP>
__device__ __forceinline__ uint32_t stack_pop(SP stk, int stride, OP ovf, int ostride, int& sp)
{
    --sp;
    return (kLds >= kCap || sp < kLds) ? stk[sp * stride] : ovf[(sp - kLds) * ostride];
}

// One 4-wide node's slab tests: entry distance tn[c], hit[c] (the box overlaps [tmin, tcur]
// and the child is not empty), ref[c] and count cn[c], from a GpuNode4 (kQ false: 8 float4
// per node) or a GpuNodeQ (kQ true: 4 float4; a plane origin + q 2^e is met at
// q (2^e / d) + (origin / d - o / d), one conversion and one fma; 2^e / d is ldexp(1 / d, e),
// exact unless it leaves the normal range). For the quantised node the plane met first on each
// axis, the lower one when the direction is positive (1 / d keeps the sign of d, which is
// never 0 here), is chosen per plane word, four children at once; its distance equals the min
// of the two planes' distances for a box with lo <= hi, since the rounding of fma is monotone.
// (For the float node choosing the rows costs as many address instructions as it saves.)
template <bool kQ>
__device__ __forceinline__ void node4_test(const float4* nodes, uint32_t node, float ix, float iy, float iz, float ox,
                                           float oy, float oz, float tmin, float tcur, float (&tn)[4], bool (&hit)[4],
                                           int (&ref)[4], int (&cn)[4])
{
    if (kQ)
    {
        const size_t nb = 4 * (size_t)node;
        const float4 W0 = gload(nodes, nb);
        const int4 I1 = gload_i(nodes, nb + 1), I2 = gload_i(nodes, nb + 2);
        const uint4 W1 = make_uint4((uint32_t)I1.x, (uint32_t)I1.y, (uint32_t)I1.z, (uint32_t)I1.w);
        const uint4 W2 = make_uint4((uint32_t)I2.x, (uint32_t)I2.y, (uint32_t)I2.z, (uint32_t)I2.w);
        const int4 REF = gload_i(nodes, nb + 3);
        // 2^e / d = ldexp(1 / d, e) from the signed exponent bytes (one extract and one
        // v_ldexp per axis; equal to the product with 2^e in every case)
        const uint32_t ex = __float_as_uint(W0.w);
        const float sx = __builtin_amdgcn_ldexpf(ix, (int)(int8_t)(ex & 0xffu));
        const float sy = __builtin_amdgcn_ldexpf(iy, (int)(int8_t)((ex >> 8) & 0xffu));
        const float sz = __builtin_amdgcn_ldexpf(iz, (int)(int8_t)((ex >> 16) & 0xffu));
        const float bx = __builtin_fmaf(W0.x, ix, -ox), by = __builtin_fmaf(W0.y, iy, -oy),
                    bz = __builtin_fmaf(W0.z, iz, -oz);
        ref[0] = REF.x; ref[1] = REF.y; ref[2] = REF.z; ref[3] = REF.w;
        // per axis the entry plane is the lower one when the direction is positive: choosing
        // the plane words once per node replaces a min and a max per axis and child (for a
        // box with lo <= hi the same values; an empty child's inverted box is rejected by
        // its count either way)
        const uint32_t nx = ix >= 0.f ? W1.x : W1.y, fx = ix >= 0.f ? W1.y : W1.x;
        const uint32_t ny = iy >= 0.f ? W1.z : W1.w, fy = iy >= 0.f ? W1.w : W1.z;
        const uint32_t nz = iz >= 0.f ? W2.x : W2.y, fz = iz >= 0.f ? W2.y : W2.x;
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
            const int sh = 8 * c;
            const float x0 = __builtin_fmaf((float)((nx >> sh) & 0xffu), sx, bx);
            const float x1 = __builtin_fmaf((float)((fx >> sh) & 0xffu), sx, bx);
            const float y0 = __builtin_fmaf((float)((ny >> sh) & 0xffu), sy, by);
            const float y1 = __builtin_fmaf((float)((fy >> sh) & 0xffu), sy, by);
            const float z0 = __builtin_fmaf((float)((nz >> sh) & 0xffu), sz, bz);
            const float z1 = __builtin_fmaf((float)((fz >> sh) & 0xffu), sz, bz);
            cn[c] = (int)(int8_t)(uint8_t)((W2.z >> sh) & 0xffu);
            tn[c] = fmaxf(fmaxf(x0, y0), fmaxf(z0, tmin));
            const float tf = fminf(fminf(x1, y1), fminf(z1, tcur));
            hit[c] = (tn[c] <= tf) && (cn[c] >= 0);
        }
    }
    else
    {
        const size_t nb = 8 * (size_t)node;
        const float4 LX = gload(nodes, nb), HX = gload(nodes, nb + 1), LY = gload(nodes, nb + 2),
                     HY = gload(nodes, nb + 3), LZ = gload(nodes, nb + 4), HZ = gload(nodes, nb + 5);
        const int4 REF = gload_i(nodes, nb + 6);
        const int4 CNT = gload_i(nodes, nb + 7);
        const float lx[4] = {LX.x, LX.y, LX.z, LX.w}, hx[4] = {HX.x, HX.y, HX.z, HX.w};
        const float ly[4] = {LY.x, LY.y, LY.z, LY.w}, hy[4] = {HY.x, HY.y, HY.z, HY.w};
        const float lz[4] = {LZ.x, LZ.y, LZ.z, LZ.w}, hz[4] = {HZ.x, HZ.y, HZ.z, HZ.w};
        ref[0] = REF.x; ref[1] = REF.y; ref[2] = REF.z; ref[3] = REF.w;
        cn[0] = CNT.x; cn[1] = CNT.y; cn[2] = CNT.z; cn[3] = CNT.w;
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
            const float x0 = __builtin_fmaf(lx[c], ix, -ox), x1 = __builtin_fmaf(hx[c], ix, -ox);
            const float y0 = __builtin_fmaf(ly[c], iy, -oy), y1 = __builtin_fmaf(hy[c], iy, -oy);
            const float z0 = __builtin_fmaf(lz[c], iz, -oz), z1 = __builtin_fmaf(hz[c], iz, -oz);
            tn[c] = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
            const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tcur));
            hit[c] = (tn[c] <= tf) && (cn[c] >= 0);
        }
    }
}

// 4-wide traversal over GpuNode4 or GpuNodeQ: one visit loads the node (128 or 64 B), tests four boxes,
// intersects leaf children at once, then visits the hit internal children nearest first
// (a 5-comparator sorting network on their entry distances; the far ones are pushed,
// farthest first). Same accept rules as traverse(). The builder bounds the stack depth
// by kMaxStack4 (at most three pushes per level).
template <bool kAny, bool kCount, int kLds>
__device__ __forceinline__ bool traverse4(const SceneDev& s, f3 o, f3 d, float tmin, float tmax, uint32_t* stk,
                                          int stride, TraceResult& res, TraceCounts& cnt, uint32_t* ovf,
                                          int ostride)
{
    const float dx = fabsf(d.x) < 1e-20f ? copysignf(1e-20f, d.x) : d.x;
    const float dy = fabsf(d.y) < 1e-20f ? copysignf(1e-20f, d.y) : d.y;
    const float dz = fabsf(d.z) < 1e-20f ? copysignf(1e-20f, d.z) : d.z;
    const float ix = 1.f / dx, iy = 1.f / dy, iz = 1.f / dz;
    const float ox = o.x * ix, oy = o.y * iy, oz = o.z * iz;
    const float kInf = __builtin_inff();
    float tcur = tmax;
    int prim = -1;
    float hb1 = 0.f, hb2 = 0.f;
    uint32_t node = 0;
    int sp = 0;
    for (;;)
    {
        if (kCount) cnt.nodes++;
        int ref[4], cn[4];
        float tn[4];
        bool hit[4];
        if (s.qnodes) node4_test<true>(s.nodes, node, ix, iy, iz, ox, oy, oz, tmin, tcur, tn, hit, ref, cn);
        else node4_test<false>(s.nodes, node, ix, iy, iz, ox, oy, oz, tmin, tcur, tn, hit, ref, cn);
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
            if (hit[c] && cn[c] > 0)
            {
                for (int q = ref[c]; q < ref[c] + cn[c]; ++q)
                {
                    const float4 v0 = gload(s.tris, 3 * (size_t)q + 0);
                    const float4 e1 = gload(s.tris, 3 * (size_t)q + 1);
                    const float4 e2 = gload(s.tris, 3 * (size_t)q + 2);
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
                hit[c] = false;
            }
        }
        float key[4];
        uint32_t id[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
            key[c] = (hit[c] && tn[c] <= tcur) ? tn[c] : kInf;
            id[c] = (uint32_t)ref[c];
        }
#define BSH_CSWAP(a, b)                                       \
    if (key[b] < key[a])                                      \
    {                                                         \
        float tk = key[a]; key[a] = key[b]; key[b] = tk;      \
        uint32_t ti = id[a]; id[a] = id[b]; id[b] = ti;       \
    }
        BSH_CSWAP(0, 1)
        BSH_CSWAP(2, 3)
        BSH_CSWAP(0, 2)
        BSH_CSWAP(1, 3)
        BSH_CSWAP(1, 2)
#undef BSH_CSWAP
        const int n = (key[0] < kInf) + (key[1] < kInf) + (key[2] < kInf) + (key[3] < kInf);
        if (n == 0)
        {
            if (sp == 0) break;
            node = stack_pop<kLds, kMaxStack4>((lu32*)stk, stride, (gu32*)ovf, ostride, sp);
            continue;
        }
#pragma unroll
        for (int k = 3; k >= 1; --k)
            if (k < n) stack_push<kLds, kMaxStack4>((lu32*)stk, stride, (gu32*)ovf, ostride, sp, id[k]);
        node = id[0];
    }
    res.prim = prim;
    res.t = tcur;
    res.b1 = hb1;
    res.b2 = hb2;
    return prim >= 0;
}

}  // namespace bsh
