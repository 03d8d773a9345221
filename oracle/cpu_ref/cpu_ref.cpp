/*
 * TEST INFRASTRUCTURE ONLY -- the CPU restatement of the reference hot path.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load this
 * library, and only as a checker or as the timed CPU baseline; the product never links
 * or calls it. It restates, from the reference's sources (paths relative to
 * FireRays/Banshee/), the arithmetic of:
 *   Bvh build and traversal ........ accelerator/bvh.cpp:16-355 (build), 499-647 (queries)
 *   Mesh::IntersectFace/FillHit .... primitive/mesh.cpp:7-123, 304-378
 *   Lambert / Microfacet / GGX ..... bsdf/lambert.h:66-127, microfacet.h:80-150,
 *                                    microfacet_distribution.h:205-283, fresnel.h:59-90
 *   Emissive, AreaLight ............ material/emissive.h:55-95, light/arealight.cpp:25-58
 *   GiTracer / DiTracer ............ tracer/gitracer.cpp:13-140, tracer/ditracer.cpp:57-239
 *   PerscpectiveCamera, samplers ... camera/perspective_camera.cpp:5-38,
 *                                    sampler/stratified_sampler.cpp:7-24, rng/mcrng.h:45-76
 *   ImageRenderer::RenderTile ...... renderer/imagerenderer.cpp:61-101
 *   SH basis, convolution, projection math/sh.cpp:11-180, math/shproject.cpp:73-130
 * in the same float operation order (built with -ffp-contract=off). glibc's rand() is
 * restated from its published TYPE_3 additive-feedback algorithm for the McRng sequence
 * check; renders use the keyed sample stream (oracle/keyed_stream.h).
 *
 * Parity pins: every function is checked against the compiled reference's fixtures
 * (tests/golden) and, where the mathematics fixes the answer, against closed forms
 * (tests/test_oracle_pins.py).
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "../keyed_stream.h"

namespace {

const float PI = 3.14159265358979323846f;
const float MINPDF = 0.05f;
const float MAXRADIANCE = 4.f;

struct V3
{
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
inline V3 mulv(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 muls(V3 a, float c) { return V3{a.x * c, a.y * c, a.z * c}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }          // float3.h:96-99
inline float sqn(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }                  // float3.h:50
inline V3 cross(V3 a, V3 b)                                                            // float3.h:108-114
{
    return V3{a.y * b.z - b.y * a.z, b.x * a.z - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 nrm(V3 v)                                                                    // float3.h:51,58
{
    float s = sqn(v);
    if (s > 0.f) return muls(v, 1.f / std::sqrt(s));
    return v;
}
inline float clampf(float x, float a, float b) { return x < a ? a : (x > b ? b : x); } // mathutils.h:90-93
inline V3 at(const float* p, int i) { return V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

V3 orthovector(V3 n)                                                                   // mathutils.h:216-228
{
    V3 p;
    if (std::fabs(n.z) > 0.707106781186547524401f)
    {
        float k = std::sqrt(n.y * n.y + n.z * n.z);
        p = V3{0.f, -n.z / k, n.y / k};
    }
    else
    {
        float k = std::sqrt(n.x * n.x + n.y * n.y);
        p = V3{-n.y / k, n.x / k, 0.f};
    }
    return p;
}

V3 map_to_hemisphere(V3 n, float sx, float sy, float e)                                // mathutils.h:231-244
{
    V3 u = orthovector(n);
    V3 v = cross(u, n);
    u = cross(n, v);
    float sinpsi = sinf(2 * PI * sx);
    float cospsi = cosf(2 * PI * sx);
    float costheta = powf(1.f - sy, 1.f / (e + 1.f));
    float sintheta = std::sqrt(1.f - costheta * costheta);
    return nrm(add(add(muls(muls(u, sintheta), cospsi), muls(muls(v, sintheta), sinpsi)), muls(n, costheta)));
}

inline float power_heuristic(float f, float g) { return (f * f) / (f * f + g * g); }  // mathutils.h:302-307

/* ---------------- scene ---------------- */
struct Hit
{
    float t;
    V3 p, n, ng, dpdu, dpdv;
    int m;
    int prim;
    int bundle;
};

struct Mat
{
    int type;
    V3 c;
    float eta, alpha;
};

struct Light
{
    int bundle, tri;
    float area;
};

struct Box
{
    V3 lo, hi;
};
inline Box box_empty()
{
    float M = std::numeric_limits<float>::max();
    return Box{V3{M, M, M}, V3{-M, -M, -M}};
}
inline void grow(Box& b, V3 p)                                                          // bbox.h:104-114
{
    b.lo = V3{std::min(b.lo.x, p.x), std::min(b.lo.y, p.y), std::min(b.lo.z, p.z)};
    b.hi = V3{std::max(b.hi.x, p.x), std::max(b.hi.y, p.y), std::max(b.hi.z, p.z)};
}
inline void grow(Box& b, const Box& o)
{
    b.lo = V3{std::min(b.lo.x, o.lo.x), std::min(b.lo.y, o.lo.y), std::min(b.lo.z, o.lo.z)};
    b.hi = V3{std::max(b.hi.x, o.hi.x), std::max(b.hi.y, o.hi.y), std::max(b.hi.z, o.hi.z)};
}
inline float comp(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
inline V3 center(const Box& b) { return muls(add(b.hi, b.lo), 0.5f); }                 // bbox.h:60
inline V3 extents(const Box& b) { return sub(b.hi, b.lo); }
inline float area(const Box& b)                                                         // bbox.h:236-240
{
    V3 e = extents(b);
    return 2.f * (e.x * e.y + e.x * e.z + e.y * e.z);
}
inline int maxdim(const Box& b)                                                         // bbox.h:223-234
{
    V3 e = extents(b);
    if (e.x >= e.y && e.x >= e.z) return 0;
    if (e.y >= e.x && e.y >= e.z) return 1;
    if (e.z >= e.x && e.z >= e.y) return 2;
    return 0;
}

struct Node
{
    Box box;
    int leaf;
    int lc, rc;       // internal
    int start, num;   // leaf
};

struct Scene
{
    std::vector<float> V, N;
    std::vector<int> tv, tn, tm, tb;   // per triangle: vertex idx, normal idx, material, bundle
    std::vector<int> bundle_light;     // last AreaLight created on the bundle, -1 if none
    std::vector<Mat> mats;
    std::vector<Light> lights;
    float cam[13];
    std::vector<Node> nodes;
    std::vector<int> primids;
    int root = -1;
};

/* ---- BVH build: Bvh::BuildImpl/BuildNode/FindSahSplit, bvh.cpp:68-355, 357-496 (usesah = true) ---- */
struct Req
{
    int start, num;
    int* slot;
    Box bounds, cbounds;
    int level;
};

struct Builder
{
    Scene& s;
    std::vector<Box> bounds;
    std::vector<V3> cen;
    explicit Builder(Scene& sc) : s(sc) {}

    bool sah(const Req& r, int& dim, float& split)
    {
        const int B = 64;
        int splitidx = -1;
        float best = std::numeric_limits<float>::max();
        dim = 0;
        V3 ext = extents(r.cbounds);
        if (sqn(ext) == 0.f) return false;
        struct Bin { Box b; int n; };
        static thread_local Bin bins[3][64];
        float invarea = 1.f / area(r.bounds);
        V3 rootmin = r.cbounds.lo;
        for (int a = 0; a < 3; ++a)
        {
            float rminc = comp(rootmin, a), rng = comp(ext, a), inv = 1.f / rng;
            if (rng == 0.f) continue;
            for (int i = 0; i < B; ++i) { bins[a][i].n = 0; bins[a][i].b = box_empty(); }
            for (int i = r.start; i < r.start + r.num; ++i)
            {
                int idx = s.primids[i];
                int bi = (int)std::min<float>(B * ((comp(cen[idx], a) - rminc) * inv), B - 1);
                ++bins[a][bi].n;
                grow(bins[a][bi].b, bounds[idx]);
            }
            Box right[B - 1];
            Box rb = box_empty();
            for (int i = B - 1; i > 0; --i)
            {
                grow(rb, bins[a][i].b);
                right[i - 1] = rb;
            }
            Box lb = box_empty();
            int lcnt = 0, rcnt = r.num;
            for (int i = 0; i < B - 1; ++i)
            {
                grow(lb, bins[a][i].b);
                lcnt += bins[a][i].n;
                rcnt -= bins[a][i].n;
                float c = 10.f + (lcnt * area(lb) + rcnt * area(right[i])) * invarea;
                if (c < best)
                {
                    dim = a;
                    splitidx = i;
                    best = c;
                }
            }
        }
        if (splitidx == -1) return false;
        split = comp(rootmin, dim) + (splitidx + 1) * (comp(ext, dim) / B);
        return true;
    }

    void node(const Req& r)
    {
        int id = (int)s.nodes.size();
        s.nodes.push_back(Node());
        s.nodes[id].box = r.bounds;
        if (r.slot == nullptr) s.root = id;
        if (r.num < 2)
        {
            s.nodes[id].leaf = 1;
            s.nodes[id].start = r.start;
            s.nodes[id].num = r.num;
            return;
        }
        s.nodes[id].leaf = 0;
        int axis = maxdim(r.cbounds);
        float border = comp(center(r.cbounds), axis);
        if (r.level < 10)
        {
            int d;
            float sp;
            if (sah(r, d, sp) && !(sp != sp))
            {
                axis = d;
                border = sp;
            }
        }
        Box lb = box_empty(), rb = box_empty(), lcb = box_empty(), rcb = box_empty();
        int splitidx = r.start;
        bool near2far = (r.num + r.start) & 1;
        int* pi = s.primids.data();
        if (comp(extents(r.cbounds), axis) > 0.f)
        {
            int first = r.start, last = r.start + r.num;
            // in-place partition with the direction flip of bvh.cpp:109-184
            auto goes_left = [&](int p) {
                float c = comp(cen[p], axis);
                return near2far ? (c < border) : (c >= border);
            };
            for (;;)
            {
                while (first != last && goes_left(pi[first]))
                {
                    grow(lb, bounds[pi[first]]);
                    grow(lcb, cen[pi[first]]);
                    ++first;
                }
                if (first == last--) break;
                grow(rb, bounds[pi[first]]);
                grow(rcb, cen[pi[first]]);
                while (first != last && !goes_left(pi[last]))
                {
                    grow(rb, bounds[pi[last]]);
                    grow(rcb, cen[pi[last]]);
                    --last;
                }
                if (first == last) break;
                grow(lb, bounds[pi[last]]);
                grow(lcb, cen[pi[last]]);
                std::swap(pi[first++], pi[last]);
            }
            splitidx = first;
        }
        if (splitidx == r.start || splitidx == r.start + r.num)
        {
            splitidx = r.start + (r.num >> 1);
            for (int i = r.start; i < splitidx; ++i) { grow(lb, bounds[pi[i]]); grow(lcb, cen[pi[i]]); }
            for (int i = splitidx; i < r.start + r.num; ++i) { grow(rb, bounds[pi[i]]); grow(rcb, cen[pi[i]]); }
        }
        Req L{r.start, splitidx - r.start, &s.nodes[id].lc, lb, lcb, r.level + 1};
        Req R{splitidx, r.num - (splitidx - r.start), &s.nodes[id].rc, rb, rcb, r.level + 1};
        int lid = (int)s.nodes.size();
        node(L);
        s.nodes[id].lc = lid;
        int rid = (int)s.nodes.size();
        node(R);
        s.nodes[id].rc = rid;
    }

    void build()
    {
        size_t n = s.tm.size();
        bounds.resize(n);
        cen.resize(n);
        s.primids.resize(n);
        Box cb = box_empty(), tb = box_empty();
        for (size_t i = 0; i < n; ++i)
        {
            V3 p1 = at(s.V.data(), s.tv[3 * i]), p2 = at(s.V.data(), s.tv[3 * i + 1]), p3 = at(s.V.data(), s.tv[3 * i + 2]);
            Box b{V3{std::min(p1.x, p2.x), std::min(p1.y, p2.y), std::min(p1.z, p2.z)},
                  V3{std::max(p1.x, p2.x), std::max(p1.y, p2.y), std::max(p1.z, p2.z)}};   // bbox(p1, p2)
            grow(b, p3);
            bounds[i] = b;
            cen[i] = center(b);
            grow(cb, cen[i]);
            grow(tb, b);
            s.primids[i] = (int)i;
        }
        s.nodes.clear();
        s.nodes.reserve(2 * n);
        if (n) node(Req{0, (int)n, nullptr, tb, cb, 0});
    }
};

/* ---- queries ---- */
// Mesh::IntersectFace, mesh.cpp:7-45
inline bool isect_face(const Scene& s, int prim, V3 o, V3 d, float tmin, float tmax, float& t, float& a, float& b)
{
    V3 p1 = at(s.V.data(), s.tv[3 * prim]), p2 = at(s.V.data(), s.tv[3 * prim + 1]), p3 = at(s.V.data(), s.tv[3 * prim + 2]);
    V3 e1 = sub(p2, p1), e2 = sub(p3, p1);
    V3 s1 = cross(d, e2);
    float invd = 1.f / dot(s1, e1);
    V3 dd = sub(o, p1);
    float b1 = dot(dd, s1) * invd;
    if (b1 < 0.f || b1 > 1.0) return false;
    V3 s2 = cross(dd, e1);
    float b2 = dot(d, s2) * invd;
    if (b2 < 0.f || b1 + b2 > 1.f) return false;
    float tmp = dot(e2, s2) * invd;
    if (tmp > tmin && tmp <= tmax)
    {
        t = tmp;
        a = b1;
        b = b2;
        return true;
    }
    return false;
}

// Mesh::FillHit, mesh.cpp:47-97 (identity transform, no UVs)
void fill_hit(const Scene& s, int prim, float t, float a, float b, Hit& h)
{
    const float* V = s.V.data();
    const float* N = s.N.data();
    V3 p1 = at(V, s.tv[3 * prim]), p2 = at(V, s.tv[3 * prim + 1]), p3 = at(V, s.tv[3 * prim + 2]);
    V3 n1 = at(N, s.tn[3 * prim]), n2 = at(N, s.tn[3 * prim + 1]), n3 = at(N, s.tn[3 * prim + 2]);
    float w = 1.f - a - b;
    h.t = t;
    h.p = add(add(muls(p1, w), muls(p2, a)), muls(p3, b));
    h.n = nrm(add(add(muls(n1, w), muls(n2, a)), muls(n3, b)));
    h.ng = nrm(cross(sub(p3, p1), sub(p2, p1)));
    if (dot(h.n, h.ng) < 0) h.ng = neg(h.ng);
    h.dpdu = orthovector(h.n);
    h.dpdv = cross(h.n, h.dpdu);
    h.m = s.tm[prim];
    h.prim = prim;
    h.bundle = s.tb[prim];
}

// intersects(r, invrd, box, dirneg, maxt), bbox.h:166-188
inline bool slab(V3 o, V3 inv, const int dn[3], const Box& b, float tmin_r, float maxt)
{
    const V3* bb = &b.lo;
    float tmin = (bb[dn[0]].x - o.x) * inv.x;
    float tmax = (bb[1 - dn[0]].x - o.x) * inv.x;
    float tymin = (bb[dn[1]].y - o.y) * inv.y;
    float tymax = (bb[1 - dn[1]].y - o.y) * inv.y;
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = (bb[dn[2]].z - o.z) * inv.z;
    float tzmax = (bb[1 - dn[2]].z - o.z) * inv.z;
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    return (tmin < maxt) && (tmax > tmin_r);
}

// Bvh::Intersect closest (bvh.cpp:499-573) / any (bvh.cpp:575-647): left child first,
// right pushed when both are hit, root box never tested.
bool bvh_query(const Scene& s, V3 o, V3 d, float tmin, float tmax, bool any, int& prim, float& tout, float& ao,
               float& bo)
{
    if (s.root < 0) return false;
    V3 inv = V3{1.f / d.x, 1.f / d.y, 1.f / d.z};
    int dn[3] = {d.x < 0.f ? 1 : 0, d.y < 0.f ? 1 : 0, d.z < 0.f ? 1 : 0};
    int stack[128];
    int sp = 0;
    int nd = s.root;
    float tcur = tmax;
    bool hit = false;
    for (;;)
    {
        const Node& n = s.nodes[nd];
        if (n.leaf)
        {
            for (int i = n.start; i < n.start + n.num; ++i)
            {
                int p = s.primids[i];
                float t, a, b;
                if (isect_face(s, p, o, d, tmin, any ? tmax : tcur, t, a, b))
                {
                    if (any) return true;
                    hit = true;
                    tcur = t;
                    prim = p;
                    ao = a;
                    bo = b;
                }
            }
        }
        else
        {
            float mt = any ? tmax : tcur;
            bool l = slab(o, inv, dn, s.nodes[n.lc].box, tmin, mt);
            bool r = slab(o, inv, dn, s.nodes[n.rc].box, tmin, mt);
            if (l)
            {
                if (r) stack[sp++] = n.rc;
                nd = n.lc;
                continue;
            }
            else if (r)
            {
                nd = n.rc;
                continue;
            }
        }
        if (sp == 0) break;
        nd = stack[--sp];
    }
    if (hit) tout = tcur;
    return hit;
}

/* ---------------- materials ---------------- */
float ggx_D(float al, V3 m, V3 n)                                  // microfacet_distribution.h:205-220
{
    float ndotm = dot(m, n);
    if (ndotm <= 0.f) return 0.f;
    float ndotm2 = ndotm * ndotm;
    float sinmn = std::sqrt(1.f - clampf(ndotm * ndotm, 0.f, 1.f));
    float tanmn = sinmn / ndotm;
    float a2 = al * al;
    return (a2 / (PI * ndotm2 * ndotm2 * (a2 + tanmn * tanmn) * (a2 + tanmn * tanmn)));
}
float ggx_pdf(float al, V3 n, V3 wi, V3 wo)                        // microfacet_distribution.h:247-257
{
    V3 m = nrm(add(wi, wo));
    float mpdf = ggx_D(al, m, n) * std::abs(dot(n, m));
    return mpdf / (4.f * dot(wo, m));
}
float ggx_G1(float al, V3 v, V3 m, V3 n)                           // microfacet_distribution.h:259-276
{
    float ndotv = dot(n, v);
    float mdotv = dot(m, v);
    if (ndotv * mdotv <= 0.f) return 0.f;
    float sinnv = std::sqrt(1.f - clampf(ndotv * ndotv, 0.f, 1.f));
    float tannv = sinnv / ndotv;
    float a2 = al * al;
    return 2.f / (1.f + std::sqrt(1.f + a2 * tannv * tannv));
}
float fresnel(float etai, float etat, float ndotwi)                 // fresnel.h:59-90
{
    float ei = etai, et = etat;
    if (ndotwi < 0)
    {
        ei = etat;
        et = etai;
        ndotwi = -ndotwi;
    }
    float sint = (ei / et) * sqrtf(std::max(0.f, 1.f - ndotwi * ndotwi));
    if (sint > 1.f) return 1.f;
    float ndotwt = sqrtf(std::max(0.f, 1.f - sint * sint));
    float rparl = ((et * ndotwi) - (ei * ndotwt)) / ((et * ndotwi) + (ei * ndotwt));
    float rperp = ((ei * ndotwi) - (et * ndotwt)) / ((ei * ndotwi) + (et * ndotwt));
    return (rparl * rparl + rperp * rperp) * 0.5f;
}
V3 le_emissive(V3 e, V3 n, V3 wo)                                   // emissive.h:84-95
{
    float c = dot(n, wo);
    if (c > 0.f) return muls(e, c);
    return V3{0, 0, 0};
}
V3 mf_eval(const Mat& m, const Hit& h, V3 wi, V3 wo)                // microfacet.h:100-131
{
    float same = dot(wi, h.n) * dot(wo, h.n);
    if (same < 0.f) return V3{0, 0, 0};
    V3 n = dot(h.n, wi) < 0.f ? neg(h.n) : h.n;
    float co = dot(n, wo), ci = dot(n, wi);
    if (ci == 0.f || co == 0.f) return V3{0, 0, 0};
    V3 wh = nrm(add(wi, wo));
    float F = fresnel(1.f, m.eta, dot(wi, wh));
    float G = ggx_G1(m.alpha, wi, wh, n) * ggx_G1(m.alpha, wo, wh, n);
    return muls(m.c, ggx_D(m.alpha, wh, n) * G * F / (4.f * ci * co));
}
V3 mat_eval(const Mat& m, const Hit& h, V3 wi, V3 wo)
{
    if (m.type == 0)                                                  // lambert.h:88-106
    {
        float same = dot(wi, h.ng) * dot(wo, h.ng);
        if (same > 0.f) return muls(m.c, 1.f / PI);
        return V3{0, 0, 0};
    }
    if (m.type == 1) return mf_eval(m, h, wi, wo);
    return le_emissive(m.c, h.n, wi);                                  // emissive.h:64-69
}
float mat_pdf(const Mat& m, const Hit& h, V3 wi, V3 wo)
{
    if (m.type == 0)                                                  // lambert.h:108-126
    {
        float same = dot(wi, h.ng) * dot(wo, h.ng);
        if (same > 0.f)
        {
            V3 n = dot(wi, h.ng) >= 0.f ? h.n : neg(h.n);
            return dot(n, wo) * (1.f / PI);
        }
        return 0.f;
    }
    if (m.type == 1)                                                  // microfacet.h:134-150
    {
        float same = dot(wi, h.n) * dot(wo, h.n);
        if (same < 0.f) return 0.f;
        V3 n = dot(h.n, wi) < 0 ? neg(h.n) : h.n;
        return ggx_pdf(m.alpha, n, wi, wo);
    }
    return 0.f;
}
V3 mat_sample(const Mat& m, const Hit& h, float ux, float uy, V3 wi, V3& wo, float& pdf, int& type)
{
    if (m.type == 0)                                                  // lambert.h:66-85
    {
        type = 1 | 4;
        V3 n = dot(wi, h.ng) >= 0.f ? h.n : neg(h.n);
        wo = map_to_hemisphere(n, ux, uy, 1.f);
        float invpi = 1.f / PI;
        pdf = dot(n, wo) * invpi;
        return muls(m.c, invpi);
    }
    if (m.type == 1)                                                  // microfacet.h:80-97, distribution :223-245
    {
        type = 1 | 16;
        V3 n = h.n, du = h.dpdu, dv = h.dpdv;
        if (dot(h.n, wi) < 0.f)
        {
            n = neg(n);
            du = neg(du);
            dv = neg(dv);
        }
        float temp = std::atan(m.alpha * std::sqrt(ux) / std::sqrt(1.f - ux));
        float theta = (float)((temp >= 0) ? temp : (temp + 2 * PI));
        float ct = std::cos(theta);
        float st = std::sqrt(1.f - clampf(ct * ct, 0.f, 1.f));
        float cp = std::cos(2.f * PI * uy);
        float spn = std::sin(2.f * PI * uy);
        V3 wh = nrm(add(add(muls(muls(du, st), cp), muls(muls(dv, st), spn)), muls(n, ct)));
        wo = add(neg(wi), muls(wh, 2.f * dot(wi, wh)));
        pdf = ggx_pdf(m.alpha, n, wi, wo);
        return mf_eval(m, h, wi, wo);
    }
    type = 0;                                                         // emissive.h:55-61 (wo untouched:
    wo = V3{0, 0, 0};                                                 // the caller's float3() is zero)
    pdf = 0.f;
    return le_emissive(m.c, h.n, wi);
}

/* ---------------- area light ---------------- */
float tri_area(const Scene& s, int prim)                            // mesh.cpp:304-321
{
    V3 p1 = at(s.V.data(), s.tv[3 * prim]), p2 = at(s.V.data(), s.tv[3 * prim + 1]), p3 = at(s.V.data(), s.tv[3 * prim + 2]);
    return sqrtf(std::fabs(sqn(cross(sub(p3, p1), sub(p3, p2))))) * 0.5f;
}
// AreaLight::GetSample (arealight.cpp:25-53) -> Mesh::GetSampleOnShape/FillSample (mesh.cpp:99-123, 323-336)
V3 light_sample(const Scene& s, const Light& L, V3 p, float ux, float uy, V3& d, float& pdf)
{
    int prim = L.tri;
    float a = sqrtf(ux) * (1.f - uy), b = sqrtf(ux) * uy;
    const float* Vv = s.V.data();
    const float* Nn = s.N.data();
    V3 p1 = at(Vv, s.tv[3 * prim]), p2 = at(Vv, s.tv[3 * prim + 1]), p3 = at(Vv, s.tv[3 * prim + 2]);
    V3 n1 = at(Nn, s.tn[3 * prim]), n2 = at(Nn, s.tn[3 * prim + 1]), n3 = at(Nn, s.tn[3 * prim + 2]);
    float w = 1.f - a - b;
    V3 sp = add(add(muls(p1, w), muls(p2, a)), muls(p3, b));
    V3 sn = nrm(add(add(muls(n1, w), muls(n2, a)), muls(n3, b)));
    float spdf = 1.f / (sqrtf(std::fabs(sqn(cross(sub(p3, p1), sub(p3, p2))))) * 0.5f);
    d = sub(sp, p);
    if (spdf > 0.f && dot(sn, neg(nrm(d))) > 0.f)
    {
        pdf = spdf;
        float d2inv = 1.f / sqn(d);
        return muls(le_emissive(s.mats[s.tm[prim]].c, sn, neg(nrm(d))), d2inv);
    }
    pdf = 0.f;
    return V3{0, 0, 0};
}
// AreaLight::GetPdf -> Mesh::GetPdfOnShape, mesh.cpp:354-378
float light_pdf(const Scene& s, const Light& L, V3 p, V3 w)
{
    float t, a, b;
    if (!isect_face(s, L.tri, p, w, 0.001f, 100000.f, t, a, b)) return 0.f;
    Hit h;
    fill_hit(s, L.tri, t, a, b, h);
    if (dot(neg(w), h.n) > 0.f)
    {
        V3 d = sub(p, h.p);
        return sqn(d) / (dot(nrm(d), h.n) * L.area);
    }
    return 0.f;
}

/* ---------------- keyed sample draws ---------------- */
struct Draws
{
    uint32_t base;
    uint32_t ctr[8];
    void sample2d(uint32_t stream, float& x, float& y)
    {
        // float2(NextFloat(), NextFloat()) evaluated right to left by g++: first draw -> .y
        y = ks_float(ks_u32(base, stream, ctr[stream]++));
        x = ks_float(ks_u32(base, stream, ctr[stream]++));
    }
    int rnd() { return ks_rand(base, ctr[KS_RAND]++); }
};

/* ---------------- GI estimator ---------------- */
inline bool world_closest(const Scene& s, V3 o, V3 d, float tmin, float tmax, Hit& h, uint64_t* cnt)
{
    if (cnt) cnt[0]++;
    int prim = -1;
    float t = 0, a = 0, b = 0;
    if (!bvh_query(s, o, d, tmin, tmax, false, prim, t, a, b)) return false;
    fill_hit(s, prim, t, a, b, h);
    return true;
}
inline bool world_any(const Scene& s, V3 o, V3 d, float tmin, float tmax, uint64_t* cnt)
{
    if (cnt) cnt[1]++;
    int prim;
    float t, a, b;
    return bvh_query(s, o, d, tmin, tmax, true, prim, t, a, b);
}

// DiTracer::GetDi with one light and one BSDF sample (ditracer.cpp:57-239)
V3 get_di(const Scene& s, int li, Draws& D, V3 wo, const Hit& hit, uint64_t* cnt)
{
    V3 radiance = V3{0, 0, 0};
    const Light& L = s.lights[li];
    const Mat& mat = s.mats[hit.m];
    float lx, ly, bx, by;
    D.sample2d(KS_LIGHT, lx, ly);
    D.sample2d(KS_BRDF, bx, by);
    V3 ld;
    float lightpdf = 0.f, bsdfpdf = 0.f;
    V3 le = light_sample(s, L, hit.p, lx, ly, ld, lightpdf);
    if (lightpdf > MINPDF && sqn(le) > 0.f)
    {
        V3 wi = nrm(ld);
        float dist = sqrtf(sqn(ld));
        bool occluded = world_any(s, hit.p, wi, 0.01f, dist - 0.01f, cnt);
        if (!occluded)
        {
            V3 f = mat_eval(mat, hit, wi, wo);
            bsdfpdf = mat_pdf(mat, hit, wi, wo);
            float w = power_heuristic(lightpdf, bsdfpdf);
            radiance = add(radiance, muls(muls(muls(mulv(le, f), std::fabs(dot(hit.n, wi))), w), 1.f / lightpdf));
        }
    }
    int type = 0;
    V3 wi;
    V3 f = mat_sample(mat, hit, bx, by, wo, wi, bsdfpdf, type);
    wi = nrm(wi);
    if (sqn(f) > 0.f && bsdfpdf > MINPDF)
    {
        lightpdf = light_pdf(s, L, hit.p, wi);
        if (lightpdf < MINPDF) return muls(radiance, 1.f / 1);
        float w = power_heuristic(bsdfpdf, lightpdf);
        Hit sh;
        V3 le2 = V3{0, 0, 0};
        if (world_closest(s, hit.p, wi, 0.01f, 10000000.f, sh, cnt))
        {
            if (s.bundle_light[sh.bundle] == li)
            {
                V3 d = sub(sh.p, hit.p);
                if (dot(sh.n, neg(wi)) > 0.f)
                {
                    float d2inv = 1.f / sqn(d);
                    const Mat& lm = s.mats[sh.m];
                    V3 e = lm.type == 2 ? le_emissive(lm.c, sh.n, neg(wi)) : V3{0, 0, 0};
                    le2 = muls(e, d2inv);
                }
            }
        }
        if (sqn(le2) > 0.f)
            radiance = add(radiance, muls(muls(muls(mulv(le2, f), std::fabs(dot(hit.n, wi))), w), 1.f / bsdfpdf));
    }
    return muls(radiance, 1.f / 1);
}

// GiTracer::GetLi with numsamples = 1 (gitracer.cpp:13-140)
V3 get_li(const Scene& s, V3 o, V3 d, float tmin, float tmax, Draws& D, int maxdepth, V3 bg, uint64_t* cnt)
{
    V3 radiance = V3{0, 0, 0};
    Hit hp;
    if (!world_closest(s, o, d, tmin, tmax, hp, cnt)) return add(radiance, bg);
    float ux, uy;
    D.sample2d(KS_BRDF, ux, uy);
    V3 rd = d;
    Hit hit = hp;
    V3 thr = V3{1.f, 1.f, 1.f};
    int nl = (int)s.lights.size();
    for (int bounce = 0; bounce < maxdepth; ++bounce)
    {
        const Mat& mat = s.mats[hit.m];
        V3 wo = neg(rd);
        if (bounce == 0 && mat.type == 2)
        {
            radiance = add(radiance, le_emissive(mat.c, hit.n, wo));
            break;
        }
        if (nl > 0)
        {
            if (bounce != 0)
            {
                int idx = (int)((unsigned)D.rnd() % (unsigned)nl);
                radiance = add(radiance, muls(mulv(thr, get_di(s, idx, D, wo, hit, cnt)), (float)nl));
            }
            else
            {
                for (int i = 0; i < nl; ++i) radiance = add(radiance, mulv(thr, get_di(s, i, D, wo, hit, cnt)));
            }
        }
        int type = 0;
        float pdf = 0.f;
        V3 wi;
        V3 f = mat_sample(mat, hit, ux, uy, wo, wi, pdf, type);
        if (sqn(f) == 0.f || pdf < MINPDF) break;
        V3 no = hit.p;
        rd = nrm(wi);
        thr = mulv(thr, muls(f, std::fabs(dot(hit.n, wi)) / pdf));
        if (bounce > 3)
        {
            float rnd = (float)D.rnd() / 2147483647;
            float lum = 0.2126f * thr.x + 0.7152f * thr.y + 0.0722f * thr.z;
            float q = std::min(0.5f, lum);
            if (rnd > q) break;
            thr = muls(thr, 1.f / q);
        }
        if (bounce + 1 == maxdepth) break;   // the reference traces this ray but never reads its hit
        if (!world_closest(s, no, rd, 0.01f, 1000000.f, hit, cnt)) break;
    }
    V3 r = muls(radiance, 1.f / 1);
    return V3{clampf(r.x, 0.f, MAXRADIANCE), clampf(r.y, 0.f, MAXRADIANCE), clampf(r.z, 0.f, MAXRADIANCE)};
}

/* An estimator of the integral the quirk-free GPU tracer estimates, sharing none of its light
   sampling, MIS or light pdfs: paths scatter by BSDF sampling alone and collect the emitted
   radiance GetLe = e max(n.wo, 0) (emissive.h:84-95) wherever they land. Emitters end paths.
   Vertices 0 .. maxdepth-1 scatter, so the emitters it reaches are those the GPU tracer's
   next-event estimation at vertices 0 .. maxdepth-1 reaches. The continuation draws a fresh BSDF
   sample per vertex, weighs it by Evaluate(view, dir) |n.dir| / GetPdf(view, dir), and applies
   Russian roulette after vertex 3 as GiTracer does (gitracer.cpp:116-125). No reference
   counterpart: it is the independent check of the quirk-free estimator at depth > 1. */
V3 get_li_bsdf_only(const Scene& s, V3 o, V3 d, float tmin, float tmax, Draws& D, int maxdepth, V3 bg)
{
    Hit hit;
    if (!world_closest(s, o, d, tmin, tmax, hit, nullptr)) return bg;
    V3 L = V3{0, 0, 0}, T = V3{1.f, 1.f, 1.f}, rd = d;
    for (int b = 0;; ++b)
    {
        const Mat& mat = s.mats[hit.m];
        V3 wo = neg(rd);
        if (mat.type == 2)
        {
            L = add(L, mulv(T, le_emissive(mat.c, hit.n, wo)));
            break;
        }
        if (b >= maxdepth) break;
        float ux, uy;
        D.sample2d(KS_BRDF, ux, uy);
        V3 wi;
        float pdf = 0.f;
        int type = 0;
        (void)mat_sample(mat, hit, ux, uy, wo, wi, pdf, type);
        wi = nrm(wi);
        V3 f = mat_eval(mat, hit, wo, wi);
        float p = mat_pdf(mat, hit, wo, wi);
        if (!(p > 0.f) || sqn(f) == 0.f) break;
        T = mulv(T, muls(f, std::fabs(dot(hit.n, wi)) / p));
        if (b > 3)
        {
            float rnd = (float)D.rnd() / 2147483647;
            float lum = 0.2126f * T.x + 0.7152f * T.y + 0.0722f * T.z;
            float q = std::min(0.5f, lum);
            if (rnd > q) break;
            T = muls(T, 1.f / q);
        }
        V3 no = hit.p;
        if (!world_closest(s, no, wi, 0.01f, 1000000.f, hit, nullptr)) break;
        rd = wi;
    }
    return L;
}

struct Cam
{
    V3 eye, fwd, right, up;
    float znear, zfar, dimx, dimy;
};
Cam make_cam(const float* c)                                        // perspective_camera.cpp:5-23
{
    Cam k;
    V3 eye = v3(c[0], c[1], c[2]), atp = v3(c[3], c[4], c[5]), up = v3(c[6], c[7], c[8]);
    k.eye = eye;
    k.fwd = nrm(sub(atp, eye));
    k.right = cross(k.fwd, nrm(up));
    k.up = cross(k.right, k.fwd);
    k.znear = c[9];
    k.zfar = c[10];
    k.dimx = 2 * tanf(c[11] * 0.5f) * c[9];
    k.dimy = k.dimx / c[12];
    return k;
}
V3 cam_dir(const Cam& k, float sx, float sy)                         // perspective_camera.cpp:25-38
{
    float hx = sx - 0.5f, hy = sy - 0.5f;
    float cx = hx * k.dimx, cy = hy * k.dimy;
    return nrm(add(add(muls(k.fwd, k.znear), muls(k.right, cx)), muls(k.up, cy)));
}

// One camera sample of ImageRenderer::RenderTile (imagerenderer.cpp:78-90) with a
// StratifiedSampler(grid) image sampler (stratified_sampler.cpp:7-24).
V3 camera_sample(const Scene& s, const Cam& k, int W, int H, int x, int y, int si, int grid, int spp, int depth,
                 uint32_t seed, uint64_t* cnt)
{
    Draws D;
    D.base = ks_base(seed, (uint32_t)(((uint64_t)y * W + x) * spp + si));
    for (int i = 0; i < 8; ++i) D.ctr[i] = 0;
    float ux, uy;
    D.sample2d(KS_IMG, ux, uy);
    int idx = si % (grid * grid);
    float cell = 1.f / grid;
    float sx = ((idx % grid) + ux) * cell, sy = ((idx / grid) + uy) * cell;
    float imx = (float)x / W + (1.f / W) * sx, imy = (float)y / H + (1.f / H) * sy;
    return get_li(s, k.eye, cam_dir(k, imx, imy), k.znear, k.zfar, D, depth, V3{0, 0, 0}, cnt);
}

/* ---------------- glibc rand() (TYPE_3, degree 31) for the McRng sequence check ---------------- */
struct GlibcRand
{
    int32_t r[344];
    int i = 0;
    std::vector<int32_t> v;
    explicit GlibcRand(unsigned seed)
    {
        // Peter Selinger, "The GLIBC pseudo-random number generator" (2007)
        v.resize(34);
        v[0] = (int32_t)(seed == 0 ? 1 : seed);
        for (int k = 1; k < 31; ++k)
        {
            int64_t hi = v[k - 1] / 127773, lo = v[k - 1] % 127773;
            int64_t w = 16807 * lo - 2836 * hi;
            if (w < 0) w += 2147483647;
            v[k] = (int32_t)w;
        }
        for (int k = 31; k < 34; ++k) v[k] = v[k - 31];
        for (int k = 34; k < 344; ++k) v.push_back((int32_t)((uint32_t)v[k - 31] + (uint32_t)v[k - 3]));
    }
    int next()
    {
        int32_t x = (int32_t)((uint32_t)v[v.size() - 31] + (uint32_t)v[v.size() - 3]);
        v.push_back(x);
        return (int)((uint32_t)x >> 1);
    }
};

struct McRng                                                         // mcrng.h:45-76
{
    unsigned z, w;
    unsigned next()
    {
        z = 36969 * (z & 65535) + (z >> 16);
        w = 18000 * (w & 65535) + (w >> 16);
        return (z << 16) + w;
    }
};

/* ---------------- spherical harmonics ---------------- */
inline int shidx(int l, int m) { return l * l + l + m; }             // sh.h:57-60
void legendre(float x, int lmax, float* out)                        // sh.cpp:11-50
{
    out[shidx(0, 0)] = 1.f;
    if (lmax >= 1) out[shidx(1, 0)] = x;
    for (int l = 2; l <= lmax; ++l)
        out[shidx(l, 0)] = ((2 * l - 1) * x * out[shidx(l - 1, 0)] - (l - 1) * out[shidx(l - 2, 0)]) / l;
    float neg1 = -1.f, dfact = 1.f;
    float xroot = std::sqrt(std::max(0.f, 1.f - x * x));
    float xpow = xroot;
    for (int l = 1; l <= lmax; ++l)
    {
        out[shidx(l, l)] = neg1 * dfact * xpow;
        neg1 *= -1.f;
        dfact *= 2 * l + 1;
        xpow *= xroot;
    }
    for (int l = 2; l <= lmax; ++l) out[shidx(l, l - 1)] = x * (2 * l - 1) * out[shidx(l - 1, l - 1)];
    for (int l = 3; l <= lmax; ++l)
        for (int m = 1; m <= l - 2; ++m)
            out[shidx(l, m)] = ((2 * (l - 1) + 1) * x * out[shidx(l - 1, m)] - (l - 1 + m) * out[shidx(l - 2, m)]) / (l - m);
}
float divfact(int a, int b)                                          // sh.cpp:54-71
{
    if (b == 0) return 1.f;
    float fa = (float)a, fb = std::fabs((float)b), v = 1.f;
    for (float x = fa - fb + 1.f; x <= fa + fb; x += 1.f) v *= x;
    return 1.f / v;
}
float Kf(int l, int m) { return std::sqrt((2.f * l + 1.f) * 1.f / (4 * PI) * divfact(l, m)); }   // sh.cpp:74-77
void sh_eval_one(V3 p, int lmax, float* c)                          // sh.cpp:97-153
{
    legendre(p.z, lmax, c);
    std::vector<float> klm((lmax + 1) * (lmax + 1));
    for (int l = 0; l <= lmax; ++l)
        for (int m = -l; m <= l; ++m) klm[shidx(l, m)] = Kf(l, m);
    std::vector<float> cm(lmax + 1), sm(lmax + 1);
    float xylen = std::sqrt(std::max(0.f, 1.f - p.z * p.z));
    if (xylen == 0.f)
    {
        for (int i = 0; i <= lmax; ++i) { sm[i] = 0.f; cm[i] = 1.f; }
    }
    else
    {
        float s = p.y / xylen, cc = p.x / xylen, si = 0, ci = 1;   // SinCosIndexed, sh.cpp:80-92
        for (int i = 0; i <= lmax; ++i)
        {
            sm[i] = si;
            cm[i] = ci;
            float o = si;
            si = si * cc + ci * s;
            ci = ci * cc - o * s;
        }
    }
    const float sqrt2 = std::sqrt(2.f);
    for (int l = 0; l <= lmax; ++l)
    {
        for (int m = -l; m < 0; ++m) c[shidx(l, m)] = sqrt2 * klm[shidx(l, m)] * c[shidx(l, -m)] * sm[-m];
        c[shidx(l, 0)] *= klm[shidx(l, 0)];
        for (int m = 1; m <= l; ++m) c[shidx(l, m)] *= sqrt2 * klm[shidx(l, m)] * cm[m];
    }
}

struct ParForJob
{
    std::atomic<int> next{0};
};

}  // namespace

extern "C" {

void* cref_scene_create(float const* verts, int nverts, float const* normals, int nnormals, int const* tri_v,
                        int const* tri_n, int const* tri_m, int ntris, int const* bundles, int nbundles,
                        float const* materials, int nmaterials, int const* lights, int nlights, float const* cam)
{
    Scene* s = new Scene();
    s->V.assign(verts, verts + 3 * nverts);
    s->N.assign(normals, normals + 3 * nnormals);
    s->tv.assign(tri_v, tri_v + 3 * ntris);
    s->tn.assign(tri_n, tri_n + 3 * ntris);
    s->tm.assign(tri_m, tri_m + ntris);
    s->tb.assign(ntris, 0);
    for (int b = 0; b < nbundles; ++b)
        for (int k = 0; k < bundles[6 * b + 1]; ++k) s->tb[bundles[6 * b] + k] = b;
    for (int i = 0; i < nmaterials; ++i)
    {
        const float* m = materials + 8 * i;
        s->mats.push_back(Mat{(int)m[0], V3{m[1], m[2], m[3]}, m[4], m[5]});
    }
    s->bundle_light.assign(nbundles, -1);
    for (int l = 0; l < nlights; ++l)
    {
        int b = lights[2 * l], sh = lights[2 * l + 1];
        int tri = bundles[6 * b] + sh;
        s->lights.push_back(Light{b, tri, tri_area(*s, tri)});
        s->bundle_light[b] = l;   // bundle_.arealight_ = this, arealight.cpp:22
    }
    std::memcpy(s->cam, cam, sizeof(s->cam));
    Builder(*s).build();
    return s;
}

void cref_scene_destroy(void* h) { delete (Scene*)h; }

void cref_intersect(void* h, float const* rays, int n, int any, int* prim, float* thit)
{
    const Scene& s = *(Scene*)h;
    for (int i = 0; i < n; ++i)
    {
        const float* q = rays + 8 * i;
        V3 o = v3(q[0], q[1], q[2]), d = v3(q[4], q[5], q[6]);
        int p = -1;
        float t = 0, a = 0, b = 0;
        bool hit = bvh_query(s, o, d, q[3], q[7], any != 0, p, t, a, b);
        if (any)
            prim[i] = hit ? 1 : 0;
        else
        {
            prim[i] = hit ? p : -1;
            thit[3 * i] = hit ? t : 0.f;
            thit[3 * i + 1] = hit ? a : 0.f;
            thit[3 * i + 2] = hit ? b : 0.f;
        }
    }
}

/* Same outputs as the reference harness's oref_render_keyed (tile-local, not y-flipped). */
void cref_render_keyed(void* h, int W, int H, int x0, int y0, int tw, int th, int img_grid, int nee_grid, int depth,
                       unsigned seed, float* sum, float* sumsq, float* cnt, float* per_sample, long long* rays)
{
    (void)nee_grid;
    const Scene& s = *(Scene*)h;
    Cam k = make_cam(s.cam);
    int spp = img_grid * img_grid;
    uint64_t c[2] = {0, 0};
    for (int y = y0; y < y0 + th; ++y)
        for (int x = x0; x < x0 + tw; ++x)
        {
            int i = (y - y0) * tw + (x - x0);
            float S[3] = {0, 0, 0}, Q[3] = {0, 0, 0};
            for (int si = 0; si < spp; ++si)
            {
                V3 L = camera_sample(s, k, W, H, x, y, si, img_grid, spp, depth, seed, c);
                S[0] += L.x; S[1] += L.y; S[2] += L.z;
                Q[0] += L.x * L.x; Q[1] += L.y * L.y; Q[2] += L.z * L.z;
                if (per_sample)
                {
                    float* o = per_sample + 3 * ((size_t)i * spp + si);
                    o[0] = L.x; o[1] = L.y; o[2] = L.z;
                }
            }
            for (int kk = 0; kk < 3; ++kk) { sum[3 * i + kk] = S[kk]; sumsq[3 * i + kk] = Q[kk]; }
            cnt[i] = (float)spp;
        }
    if (rays) { rays[0] = (long long)c[0]; rays[1] = (long long)c[1]; }
}

/* The BSDF-sampling-only estimator (get_li_bsdf_only) over a tile of the image: spp samples per
   pixel on the stratified image grid, drawn from the keyed stream of `seed`; Sigma L, Sigma L^2 and
   the count per pixel (tile-local, not y-flipped). Rows are spread over `threads` workers
   (0 = hardware concurrency). */
void cref_render_bsdf_only(void* h, int W, int H, int x0, int y0, int tw, int th, int img_grid, int depth,
                           unsigned seed, float const* bg, double* sum, double* sumsq, float* cnt, int threads)
{
    const Scene& s = *(Scene*)h;
    Cam k = make_cam(s.cam);
    const int spp = img_grid * img_grid;
    const V3 B = bg ? at(bg, 0) : V3{0, 0, 0};
    int nth = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nth; ++t)
        pool.emplace_back([&]() {
            for (;;)
            {
                int yy = next.fetch_add(1);
                if (yy >= th) break;
                int y = y0 + yy;
                for (int x = x0; x < x0 + tw; ++x)
                {
                    int i = yy * tw + (x - x0);
                    double S[3] = {0, 0, 0}, Q[3] = {0, 0, 0};
                    for (int si = 0; si < spp; ++si)
                    {
                        Draws D;
                        D.base = ks_base(seed, (uint32_t)(((uint64_t)y * W + x) * spp + si));
                        for (int c = 0; c < 8; ++c) D.ctr[c] = 0;
                        float ux, uy;
                        D.sample2d(KS_IMG, ux, uy);
                        float cell = 1.f / img_grid;
                        float sx = ((si % img_grid) + ux) * cell, sy = ((si / img_grid) + uy) * cell;
                        float imx = (float)x / W + (1.f / W) * sx, imy = (float)y / H + (1.f / H) * sy;
                        V3 L = get_li_bsdf_only(s, k.eye, cam_dir(k, imx, imy), k.znear, k.zfar, D, depth, B);
                        S[0] += L.x; S[1] += L.y; S[2] += L.z;
                        Q[0] += (double)L.x * L.x; Q[1] += (double)L.y * L.y; Q[2] += (double)L.z * L.z;
                    }
                    for (int c = 0; c < 3; ++c) { sum[3 * i + c] = S[c]; sumsq[3 * i + c] = Q[c]; }
                    cnt[i] = (float)spp;
                }
            }
        });
    for (auto& p : pool) p.join();
}

/* Multi-threaded wall-clock render of the full W x H frame (rows handed out dynamically to
   `threads` workers, 0 = hardware concurrency); returns seconds, rays[0..1] = closest, any. */
double cref_time_render(void* h, int W, int H, int img_grid, int nee_grid, int depth, unsigned seed, int threads,
                        long long* rays)
{
    (void)nee_grid;
    const Scene& s = *(Scene*)h;
    Cam k = make_cam(s.cam);
    int spp = img_grid * img_grid;
    int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int> next{0};
    std::vector<uint64_t> counts(2 * nt, 0);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t]() {
            uint64_t c[2] = {0, 0};
            volatile float sink = 0.f;
            for (;;)
            {
                int y = next.fetch_add(1);
                if (y >= H) break;
                for (int x = 0; x < W; ++x)
                    for (int si = 0; si < spp; ++si) sink = sink + camera_sample(s, k, W, H, x, y, si, img_grid, spp, depth, seed, c).x;
            }
            counts[2 * t] = c[0];
            counts[2 * t + 1] = c[1];
        });
    for (auto& th : pool) th.join();
    double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    uint64_t a = 0, b = 0;
    for (int t = 0; t < nt; ++t) { a += counts[2 * t]; b += counts[2 * t + 1]; }
    if (rays) { rays[0] = (long long)a; rays[1] = (long long)b; }
    return dt;
}

void cref_bsdf(float const* mat, int n, float const* frames, float const* view, float const* u, float const* ea,
               float const* eb, float* out)
{
    Mat m{(int)mat[0], V3{mat[1], mat[2], mat[3]}, mat[4], mat[5]};
    for (int i = 0; i < n; ++i)
    {
        const float* f = frames + 12 * i;
        Hit h;
        h.p = V3{0, 0, 0};
        h.n = v3(f[0], f[1], f[2]);
        h.ng = v3(f[3], f[4], f[5]);
        h.dpdu = v3(f[6], f[7], f[8]);
        h.dpdv = v3(f[9], f[10], f[11]);
        V3 wi = at(view, i), wo;
        float pdf = 0.f;
        int type = 0;
        V3 fs = mat_sample(m, h, u[2 * i], u[2 * i + 1], wi, wo, pdf, type);
        V3 fe = mat_eval(m, h, at(ea, i), at(eb, i));
        float pe = mat_pdf(m, h, at(ea, i), at(eb, i));
        float* o = out + 12 * i;
        o[0] = fs.x; o[1] = fs.y; o[2] = fs.z; o[3] = pdf;
        o[4] = wo.x; o[5] = wo.y; o[6] = wo.z; o[7] = (float)type;
        o[8] = fe.x; o[9] = fe.y; o[10] = fe.z; o[11] = pe;
    }
}

void cref_light(void* h, int light, int n, float const* p, float const* u, float const* w, float* out)
{
    const Scene& s = *(Scene*)h;
    const Light& L = s.lights[light];
    for (int i = 0; i < n; ++i)
    {
        V3 d;
        float pdf = 0.f;
        V3 le = light_sample(s, L, at(p, i), u[2 * i], u[2 * i + 1], d, pdf);
        float gp = light_pdf(s, L, at(p, i), at(w, i));
        float* o = out + 8 * i;
        o[0] = le.x; o[1] = le.y; o[2] = le.z; o[3] = pdf;
        o[4] = d.x; o[5] = d.y; o[6] = d.z; o[7] = gp;
    }
}

void cref_camera_rays(float const* cam, float const* samples, int n, float* rays)
{
    Cam k = make_cam(cam);
    for (int i = 0; i < n; ++i)
    {
        V3 d = cam_dir(k, samples[2 * i], samples[2 * i + 1]);
        float* o = rays + 8 * i;
        o[0] = k.eye.x; o[1] = k.eye.y; o[2] = k.eye.z; o[3] = k.znear;
        o[4] = d.x; o[5] = d.y; o[6] = d.z; o[7] = k.zfar;
    }
}

void cref_mcrng(unsigned seed, int n, unsigned* out)
{
    GlibcRand g(seed);
    McRng r;
    r.z = (unsigned)g.next();   // McRng() : z_(rand_uint()), w_(rand_uint()), mcrng.h:48-52
    r.w = (unsigned)g.next();
    for (int i = 0; i < n; ++i) out[i] = r.next();
}

void cref_stratified(int grid, unsigned seed, int n, float* out)
{
    GlibcRand g(seed);
    McRng r;
    r.z = (unsigned)g.next();
    r.w = (unsigned)g.next();
    float cell = 1.f / grid;
    for (int i = 0; i < n; ++i)
    {
        float uy = r.next() * 2.328306e-10f;   // first draw -> .y (right-to-left argument evaluation)
        float ux = r.next() * 2.328306e-10f;
        int idx = i % (grid * grid);
        out[2 * i] = ((idx % grid) + ux) * cell;
        out[2 * i + 1] = ((idx / grid) + uy) * cell;
    }
}

void cref_sh_eval(float const* dirs, int n, int lmax, float* out)
{
    int nt = (lmax + 1) * (lmax + 1);
    for (int i = 0; i < n; ++i) sh_eval_one(at(dirs, i), lmax, out + (size_t)nt * i);
}

void cref_sh_convolve(int lmax, float const* cin, float* cout)           // sh.cpp:160-180
{
    static const float tbl[18] = {0.8862268925f, 1.0233267546f, 0.4954159260f, 0.0000000000f, -0.1107783690f,
                                  0.0000000000f, 0.0499271341f, 0.0000000000f, -0.0285469331f, 0.0000000000f,
                                  0.0185080823f, 0.0000000000f, -0.0129818395f, 0.0000000000f, 0.0096125342f,
                                  0.0000000000f, -0.0074057109f, 0.0000000000f};
    for (int l = 0; l <= lmax; ++l)
    {
        float lam = std::sqrt((4.f * (float)M_PI) / (2.f * (float)l + 1.f));
        for (int m = -l; m <= l; ++m)
        {
            int o = shidx(l, m);
            for (int c = 0; c < 3; ++c) cout[3 * o + c] = l < 18 ? lam * cin[3 * o + c] * tbl[l] : 0.f;
        }
    }
}

void cref_sh_project(float const* env, int W, int H, int lmax, float* coeffs)   // shproject.cpp:73-130
{
    int nt = (lmax + 1) * (lmax + 1);
    std::vector<float> ylm(nt), st(H), ct(H), sp(W), cp(W);
    float thetastep = PI / H, phistep = 2.f * PI / W, theta0 = PI / H / 2, phi0 = 2.f * PI / W / 2;
    for (int i = 0; i < W; ++i) { sp[i] = std::sin(phi0 + i * phistep); cp[i] = std::cos(phi0 + i * phistep); }
    for (int i = 0; i < H; ++i) { st[i] = std::sin(theta0 + i * thetastep); ct[i] = std::cos(theta0 + i * thetastep); }
    for (int i = 0; i < 3 * nt; ++i) coeffs[i] = 0.f;
    for (int phi = 0; phi < W; ++phi)
        for (int th = 0; th < H; ++th)
        {
            V3 w = nrm(v3(st[th] * cp[phi], ct[th], st[th] * sp[phi]));
            float ux = (float)phi / W, uy = (float)th / H;
            int iu = (int)std::floor(ux * W), iv = (int)std::floor(uy * H);
            V3 le = at(env, W * iv + iu);
            sh_eval_one(w, lmax, ylm.data());
            for (int i = 0; i < nt; ++i)
            {
                V3 add_ = muls(muls(muls(muls(le, ylm[i]), st[th]), PI / H), 2.f * PI / W);
                coeffs[3 * i] += add_.x;
                coeffs[3 * i + 1] += add_.y;
                coeffs[3 * i + 2] += add_.z;
            }
        }
}

void cref_sh_irradiance(int lmax, float const* coeffs, float const* normals, int n, float* out)  // shtracer.cpp:20-36
{
    int nt = (lmax + 1) * (lmax + 1);
    std::vector<float> cc(3 * nt), ylm(nt);
    cref_sh_convolve(lmax, coeffs, cc.data());
    for (int i = 0; i < n; ++i)
    {
        sh_eval_one(at(normals, i), lmax, ylm.data());
        V3 e = V3{0, 0, 0};
        for (int k = 0; k < nt; ++k) e = add(e, muls(muls(at(cc.data(), k), ylm[k]), 0.7f / (float)M_PI));
        out[3 * i] = e.x; out[3 * i + 1] = e.y; out[3 * i + 2] = e.z;
    }
}

/* Probe baking restated (no reference counterpart): for probe p and ray j the keyed
   sample g = p*N + j draws a stratified uniform-sphere direction (stream 4, draws 0/1)
   and GiTracer::GetLi estimates L; coefficients c_i = (4 pi / N) sum_j L Y_i(w_j) are
   accumulated in double. bg (3 floats, may be NULL = black) is World::bgcolor_, the radiance of a
   probe ray that hits nothing. coeffs: nprobes x (lmax+1)^2 x 3; radiance (optional): nprobes x N x 3. */
void cref_bake_sh(void* h, float const* probes, int nprobes, int N, int lmax, int depth, unsigned seed, float tmin,
                  float const* bg, double* coeffs, float* radiance, int threads)
{
    const V3 B = bg ? at(bg, 0) : V3{0, 0, 0};
    const Scene& s = *(Scene*)h;
    int g = (int)std::lround(std::sqrt((double)N));
    int nt = (lmax + 1) * (lmax + 1);
    int nth = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nth; ++t)
        pool.emplace_back([&]() {
            std::vector<float> y(nt);
            for (;;)
            {
                int p = next.fetch_add(1);
                if (p >= nprobes) break;
                std::vector<double> acc(3 * nt, 0.0);
                V3 o = at(probes, p);
                for (int j = 0; j < N; ++j)
                {
                    Draws D;
                    D.base = ks_base(seed, (uint32_t)p * (uint32_t)N + (uint32_t)j);
                    for (int i = 0; i < 8; ++i) D.ctr[i] = 0;
                    float ux, uy;
                    D.sample2d(KS_PROBE, ux, uy);
                    float cell = 1.f / (float)g;
                    float u = ((float)(j % g) + ux) * cell, v = ((float)(j / g) + uy) * cell;
                    float z = 1.f - 2.f * u;
                    float r = std::sqrt(std::max(0.f, 1.f - z * z));
                    float phi = 2.f * PI * v;
                    V3 d = v3(r * cosf(phi), r * sinf(phi), z);
                    V3 L = get_li(s, o, d, tmin, 1000000.f, D, depth, B, nullptr);
                    if (radiance)
                    {
                        float* q = radiance + 3 * ((size_t)p * N + j);
                        q[0] = L.x; q[1] = L.y; q[2] = L.z;
                    }
                    sh_eval_one(d, lmax, y.data());
                    for (int k = 0; k < nt; ++k)
                    {
                        acc[3 * k] += (double)L.x * y[k];
                        acc[3 * k + 1] += (double)L.y * y[k];
                        acc[3 * k + 2] += (double)L.z * y[k];
                    }
                }
                double sc = 4.0 * 3.14159265358979323846 / N;
                for (int k = 0; k < 3 * nt; ++k) coeffs[(size_t)p * 3 * nt + k] = acc[k] * sc;
            }
        });
    for (auto& th : pool) th.join();
}

}  // extern "C"
