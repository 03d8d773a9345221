// Wavefront GI: the per-pixel recursion of ImageRenderer + GiTracer::GetLi +
// DiTracer::GetDi (imagerenderer.cpp:61-101, gitracer.cpp:13-140, ditracer.cpp:57-239)
// restructured as kernels over all in-flight camera samples of a chunk:
//
//   raygen -> [trace -> shade] x max_depth -> trace -> shade(final) -> accumulate
//
// shade(b) first resolves the next-event (NEE) rays its previous pass issued, then
// shades the closest hit of the extension ray, emits the new shadow (any-hit),
// MIS (closest-hit) and extension rays into two compacted queues (one atomic per
// wave and queue), and finalises terminated paths. trace() is a persistent kernel
// whose waves take 64 rays at a time from the closest queue, then from the any
// queue, so every wave runs a single traversal variant.
//
// Random numbers come from the keyed sample stream; the counter of every draw is a
// fixed function of (bounce, light), which reproduces the reference's draw order
// (image sampler, BSDF pre-draw, per-GetDi light+BSDF draws, rand() light pick and
// Russian roulette) with light and BSDF samplers of one sample each.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "internal.h"
#include "shading.h"

namespace bsh {

constexpr int kBlock = 256;
constexpr int kMaxLights = 16;

enum : uint32_t
{
    F_EXT = 1,      // extension ray in flight
    F_NEE = 2,      // NEE rays in flight
    F_TERM = 4,     // path ended; finalise once NEE results are in
    F_DONE = 8,     // final radiance written
    F_NOCLAMP = 16  // primary miss: GetLi returns bg + sum Le unclamped (gitracer.cpp:22-33)
};

enum : uint32_t { Q_EXT = 0, Q_NEE_CLOSEST = 1, Q_NEE_ANY = 2 };

struct RenderDev
{
    SceneDev s;
    CameraDev cam;
    int W, H, cx, cy, cw, ch;
    int grid, spp, s_begin, s_count;
    int depth;
    uint32_t seed;
    float bg0, bg1, bg2;
    int spc;       // samples per pixel in this chunk
    int s0;        // first sample of this chunk
    uint32_t P;    // paths in this chunk
    int nl, K;
    float4 *ray_o, *ray_d, *thr, *thr_nee, *rad, *out;
    int4* hit;
    uint4* meta;
    uint32_t* nee_mask;
    float4 *nee_a, *nee_b, *nee_wi;
    int4* nee_res;
    float4 *qc_o[2], *qc_d[2], *qa_o[2], *qa_d[2];
    int *qc_own[2], *qa_own[2];
    uint32_t* counters;            // per launch L: closest count, closest work, any count, any work
    unsigned long long* stats;     // closest, any, node visits, triangle tests
};

__device__ __forceinline__ uint32_t rand_pick_counter(int b) { return (uint32_t)((b - 1) + (b > 4 ? b - 4 : 0)); }

__global__ __launch_bounds__(kBlock) void k_zero(uint32_t* p, int n)
{
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ImageRenderer sample loop + PerscpectiveCamera::GenerateRay (imagerenderer.cpp:78-87,
// perspective_camera.cpp:25-38). Every path emits exactly one closest ray, at queue slot i.
__global__ __launch_bounds__(kBlock) void k_raygen(RenderDev R)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i == 0)
    {
        R.counters[0] = R.P;
        atomicAdd(&R.stats[0], (unsigned long long)R.P);
    }
    if (i >= R.P) return;
    uint32_t q = i / (uint32_t)R.spc;
    int s = R.s0 + (int)(i - q * (uint32_t)R.spc);
    int x = R.cx + (int)(q % (uint32_t)R.cw);
    int y = R.cy + (int)(q / (uint32_t)R.cw);
    uint32_t g = ((uint32_t)y * (uint32_t)R.W + (uint32_t)x) * (uint32_t)R.spp + (uint32_t)s;
    uint32_t base = ks_base(R.seed, g);
    float ux, uy;
    ks_sample2d(base, KS_IMG, 0, ux, uy);
    int idx = s % (R.grid * R.grid);
    float cell = 1.f / (float)R.grid;
    float sx = ((float)(idx % R.grid) + ux) * cell;
    float sy = ((float)(idx / R.grid) + uy) * cell;
    float imx = (float)x / (float)R.W + (1.f / (float)R.W) * sx;
    float imy = (float)y / (float)R.H + (1.f / (float)R.H) * sy;
    float hx = imx - 0.5f, hy = imy - 0.5f;
    float csx = hx * R.cam.dimx, csy = hy * R.cam.dimy;
    f3 fwd = mk(R.cam.fwd[0], R.cam.fwd[1], R.cam.fwd[2]);
    f3 right = mk(R.cam.right[0], R.cam.right[1], R.cam.right[2]);
    f3 up = mk(R.cam.up[0], R.cam.up[1], R.cam.up[2]);
    f3 d = normalize(R.cam.znear * fwd + csx * right + csy * up);
    f3 o = mk(R.cam.eye[0], R.cam.eye[1], R.cam.eye[2]);
    R.ray_o[i] = st3(o, 0.f);
    R.ray_d[i] = st3(d, 0.f);
    R.thr[i] = make_float4(1.f, 1.f, 1.f, 0.f);
    R.rad[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    R.meta[i] = make_uint4(base, F_EXT, 0u, 0u);
    R.qc_o[0][i] = st3(o, R.cam.znear);
    R.qc_d[0][i] = st3(d, R.cam.zfar);
    R.qc_own[0][i] = (int)((Q_EXT << 29) | i);
}

template <bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_queue(RenderDev R, int L, int buf)
{
    __shared__ uint32_t stk[kMaxStack * kBlock];
    const int lane = threadIdx.x & 63;
    uint32_t* my = stk + threadIdx.x;
    uint32_t vis_nodes = 0, vis_tris = 0;
    for (int pass = 0; pass < 2; ++pass)
    {
        const uint32_t count = R.counters[4 * L + 2 * pass];
        uint32_t* work = &R.counters[4 * L + 2 * pass + 1];
        const float4* qo = pass == 0 ? R.qc_o[buf] : R.qa_o[buf];
        const float4* qd = pass == 0 ? R.qc_d[buf] : R.qa_d[buf];
        const int* qown = pass == 0 ? R.qc_own[buf] : R.qa_own[buf];
        for (;;)
        {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(work, 64u);
            base = __shfl(base, 0);
            if (base >= count) break;
            uint32_t q = base + lane;
            if (q < count)
            {
                float4 a = qo[q], b = qd[q];
                uint32_t own = (uint32_t)qown[q];
                uint32_t kind = own >> 29, idx = own & ((1u << 29) - 1u);
                TraceResult res;
                TraceCounts cnt{0u, 0u};
                if (pass == 0)
                {
                    bool h = traverse<false, kCount>(R.s, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, b.w, my, kBlock,
                                                     res, cnt);
                    int4 o = make_int4(h ? res.prim : -1, __float_as_int(res.t), __float_as_int(res.b1),
                                       __float_as_int(res.b2));
                    if (kind == Q_EXT) R.hit[idx] = o;
                    else R.nee_res[idx] = o;
                }
                else
                {
                    bool h = traverse<true, kCount>(R.s, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, b.w, my, kBlock,
                                                    res, cnt);
                    R.nee_res[idx] = make_int4(h ? 1 : 0, 0, 0, 0);
                }
                vis_nodes += cnt.nodes;
                vis_tris += cnt.tris;
            }
        }
    }
    if (kCount)
    {
        atomicAdd(&R.stats[2], (unsigned long long)vis_nodes);
        atomicAdd(&R.stats[3], (unsigned long long)vis_tris);
    }
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        uint32_t t = __shfl_up(v, d);
        if (lane >= d) v += t;
    }
    return v;
}

__global__ __launch_bounds__(kBlock) void k_shade(RenderDev R, int b, int Lout)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const int K = R.K;
    bool emit_ext = false;
    uint32_t emit_mask = 0;
    f3 org = mk(0.f, 0.f, 0.f), ext_d = mk(0.f, 0.f, 0.f);
    if (i < R.P)
    {
        uint4 meta = R.meta[i];
        uint32_t fl = meta.y & 0xffu;
        int pb = (int)((meta.y >> 8) & 0xffu);
        int ncnt = (int)((meta.y >> 16) & 0xffu);
        if (!(fl & F_DONE))
        {
            f3 rad = ld3(R.rad[i]);
            // ---- 1. resolve the NEE rays issued at bounce pb (DiTracer::GetDi tail) ----
            if (fl & F_NEE)
            {
                uint32_t mask = R.nee_mask[i];
                f3 T = ld3(R.thr_nee[i]);
                f3 ps = ld3(R.ray_o[i]);
                for (int j = 0; j < ncnt; ++j)
                {
                    f3 local = mk(0.f, 0.f, 0.f);
                    size_t ks = (size_t)i * K + 2 * j;
                    if (mask & (1u << (2 * j)))
                    {
                        if (R.nee_res[ks].x == 0) local = local + ld3(R.nee_a[ks]);
                    }
                    if (mask & (1u << (2 * j + 1)))
                    {
                        int4 r = R.nee_res[ks + 1];
                        float4 nb = R.nee_b[ks + 1];
                        int Lid = (int)nb.z;
                        if (r.x >= 0)
                        {
                            f3 p1, p2, p3, n1, n2, n3;
                            int mat, bl;
                            tri_verts(R.s, r.x, p1, p2, p3, n1, n2, n3, mat, bl);
                            if (bl == Lid)   // shadowhit.bundle->GetAreaLight() == &light, ditracer.cpp:202
                            {
                                Frame hh = fill_hit(p1, p2, p3, n1, n2, n3, __int_as_float(r.z), __int_as_float(r.w));
                                f3 wi = ld3(R.nee_wi[ks + 1]);
                                f3 d = hh.p - ps;
                                f3 le = mk(0.f, 0.f, 0.f);
                                if (dot(hh.n, -wi) > 0.f)
                                {
                                    float d2inv = 1.f / sqnorm(d);
                                    Mat lm = load_mat(R.s, mat);
                                    if (lm.type == 2) le = emissive_le(lm.c, hh.n, -wi) * d2inv;
                                    else le = mk(0.f, 0.f, 0.f) * d2inv;  // Material::GetLe default
                                }
                                if (sqnorm(le) > 0.f)
                                {
                                    float4 na = R.nee_a[ks + 1];
                                    local = local + le * ld3(na) * na.w * nb.x * nb.y;
                                }
                            }
                        }
                    }
                    if (pb == 0) rad = rad + T * local;
                    else rad = rad + (T * local) * (float)R.nl;
                }
                fl &= ~F_NEE;
            }
            // ---- 2. shade the extension hit (GiTracer bounce loop body) ----
            if (fl & F_EXT)
            {
                fl &= ~F_EXT;
                int4 hr = R.hit[i];
                if (hr.x < 0)
                {
                    if (b == 0)
                    {
                        rad = mk(R.bg0, R.bg1, R.bg2);
                        fl |= F_NOCLAMP;
                    }
                    fl |= F_TERM;
                }
                else
                {
                    f3 p1, p2, p3, n1, n2, n3;
                    int mat, bl;
                    tri_verts(R.s, hr.x, p1, p2, p3, n1, n2, n3, mat, bl);
                    Frame h = fill_hit(p1, p2, p3, n1, n2, n3, __int_as_float(hr.z), __int_as_float(hr.w));
                    Mat m = load_mat(R.s, mat);
                    f3 wo = -ld3(R.ray_d[i]);
                    uint32_t base = meta.x;
                    f3 T = ld3(R.thr[i]);
                    float ux, uy;
                    if (b == 0)
                    {
                        ks_sample2d(base, KS_BRDF, 0, ux, uy);   // pre-drawn once per path, gitracer.cpp:36-41
                        meta.z = __float_as_uint(ux);
                        meta.w = __float_as_uint(uy);
                    }
                    else
                    {
                        ux = __uint_as_float(meta.z);
                        uy = __uint_as_float(meta.w);
                    }
                    if (b == 0 && m.type == 2)
                    {
                        rad = rad + emissive_le(m.c, h.n, wo);   // gitracer.cpp:61-68
                        fl |= F_TERM;
                    }
                    else
                    {
                        org = h.p;
                        // ---- next-event estimation (gitracer.cpp:70-84 -> ditracer.cpp:57-239) ----
                        if (R.nl > 0)
                        {
                            int jn = b == 0 ? R.nl : 1;
                            int Lsel = 0;
                            if (b != 0) Lsel = (int)(ks_rand(base, rand_pick_counter(b)) % (uint32_t)R.nl);
                            for (int j = 0; j < jn; ++j)
                            {
                                int Lid = b == 0 ? j : Lsel;
                                const LightDev L = R.s.lights[Lid];
                                uint32_t lc = b == 0 ? (uint32_t)(2 * j) : (uint32_t)(2 * R.nl + 2 * (b - 1));
                                uint32_t bc = 2u + lc;
                                float lx, ly, bx, by;
                                ks_sample2d(base, KS_LIGHT, lc, lx, ly);
                                ks_sample2d(base, KS_BRDF, bc, bx, by);
                                size_t ks = (size_t)i * K + 2 * j;
                                // light sample + shadow ray
                                f3 ld;
                                float lpdf = 0.f;
                                f3 le = light_sample(R.s, L, h.p, lx, ly, ld, lpdf);
                                if (lpdf > kMinPdf && sqnorm(le) > 0.f)
                                {
                                    f3 wi = normalize(ld);
                                    float dist = sqrtf(sqnorm(ld));
                                    f3 f = mat_eval(m, h, wi, wo);
                                    float bp = mat_pdf(m, h, wi, wo);
                                    float w = power_heuristic(lpdf, bp);
                                    f3 c = le * f * fabsf(dot(h.n, wi)) * w * (1.f / lpdf);
                                    R.nee_a[ks] = st3(c, 0.f);
                                    R.nee_wi[ks] = st3(wi, dist - 0.01f);
                                    emit_mask |= 1u << (2 * j);
                                }
                                // BSDF sample + MIS closest ray
                                f3 wis;
                                float bpdf = 0.f;
                                int type = 0;
                                f3 fb = mat_sample(m, h, bx, by, wo, wis, bpdf, type);
                                f3 wi = normalize(wis);
                                if (sqnorm(fb) > 0.f && bpdf > kMinPdf)
                                {
                                    float lp = light_pdf(R.s, L, h.p, wi);
                                    if (!(lp < kMinPdf))
                                    {
                                        float w = power_heuristic(bpdf, lp);
                                        R.nee_a[ks + 1] = st3(fb, fabsf(dot(h.n, wi)));
                                        R.nee_b[ks + 1] = make_float4(w, 1.f / bpdf, (float)Lid, 0.f);
                                        R.nee_wi[ks + 1] = st3(wi, 10000000.f);
                                        emit_mask |= 1u << (2 * j + 1);
                                    }
                                }
                            }
                            R.thr_nee[i] = st3(T, 0.f);
                            R.ray_o[i] = st3(h.p, 0.f);
                            R.nee_mask[i] = emit_mask;
                            if (emit_mask) fl |= F_NEE;
                            ncnt = jn;
                        }
                        // ---- continuation: BSDF sample, throughput, Russian roulette ----
                        if (b + 1 < R.depth)
                        {
                            f3 wi;
                            float pdf = 0.f;
                            int type = 0;
                            f3 f = mat_sample(m, h, ux, uy, wo, wi, pdf, type);
                            if (sqnorm(f) == 0.f || pdf < kMinPdf)
                            {
                                fl |= F_TERM;
                            }
                            else
                            {
                                f3 nd = normalize(wi);
                                T = T * (f * (fabsf(dot(h.n, wi)) / pdf));
                                bool alive = true;
                                if (b > 3)
                                {
                                    float rnd = ks_rand_float(base, rand_pick_counter(b) + 1u);
                                    float lum = 0.2126f * T.x + 0.7152f * T.y + 0.0722f * T.z;
                                    float q = (lum < 0.5f) ? lum : 0.5f;
                                    if (rnd > q) alive = false;
                                    else T = T * (1.f / q);
                                }
                                if (alive)
                                {
                                    emit_ext = true;
                                    ext_d = nd;
                                    R.thr[i] = st3(T, 0.f);
                                    R.ray_o[i] = st3(h.p, 0.f);
                                    R.ray_d[i] = st3(nd, 0.f);
                                    fl |= F_EXT;
                                }
                                else
                                {
                                    fl |= F_TERM;
                                }
                            }
                        }
                        else
                        {
                            fl |= F_TERM;
                        }
                    }
                }
            }
            // ---- 3. finalise: clamp(radiance / numsamples, 0, MAXRADIANCE), gitracer.cpp:139 ----
            if ((fl & F_TERM) && !(fl & F_NEE))
            {
                f3 o = rad;
                if (!(fl & F_NOCLAMP))
                {
                    f3 v = rad * (1.f / 1.f);
                    o = mk(clampf(v.x, 0.f, kMaxRadiance), clampf(v.y, 0.f, kMaxRadiance),
                           clampf(v.z, 0.f, kMaxRadiance));
                }
                R.out[i] = st3(o, 0.f);
                fl |= F_DONE;
            }
            R.rad[i] = st3(rad, 0.f);
            meta.y = fl | ((uint32_t)b << 8) | ((uint32_t)ncnt << 16);
            R.meta[i] = meta;
        }
    }
    // ---- queue append: one atomic per wave and queue ----
    uint32_t ncl = (emit_ext ? 1u : 0u), nany = 0;
    for (int k = 0; k < K; ++k)
        if (emit_mask & (1u << k)) { if (k & 1) ncl++; else nany++; }
    uint32_t packed = ncl | (nany << 16);
    uint32_t incl = wave_incl_scan(packed);
    uint32_t total = __shfl(incl, 63);
    uint32_t excl = incl - packed;
    const int lane = threadIdx.x & 63;
    uint32_t cbase = 0, abase = 0;
    if (lane == 63)
    {
        uint32_t tc = total & 0xffffu, ta = total >> 16;
        if (tc) cbase = atomicAdd(&R.counters[4 * Lout + 0], tc);
        if (ta) abase = atomicAdd(&R.counters[4 * Lout + 2], ta);
        if (tc) atomicAdd(&R.stats[0], (unsigned long long)tc);
        if (ta) atomicAdd(&R.stats[1], (unsigned long long)ta);
    }
    cbase = __shfl(cbase, 63);
    abase = __shfl(abase, 63);
    if (ncl | nany)
    {
        const int buf = Lout & 1;
        uint32_t cp = cbase + (excl & 0xffffu), ap = abase + (excl >> 16);
        if (emit_ext)
        {
            R.qc_o[buf][cp] = st3(org, 0.01f);
            R.qc_d[buf][cp] = st3(ext_d, 1000000.f);
            R.qc_own[buf][cp] = (int)((Q_EXT << 29) | i);
            ++cp;
        }
        for (int k = 0; k < K; ++k)
        {
            if (!(emit_mask & (1u << k))) continue;
            size_t ks = (size_t)i * K + k;
            float4 wd = R.nee_wi[ks];
            if (k & 1)
            {
                R.qc_o[buf][cp] = st3(org, 0.01f);
                R.qc_d[buf][cp] = wd;
                R.qc_own[buf][cp] = (int)((Q_NEE_CLOSEST << 29) | (uint32_t)ks);
                ++cp;
            }
            else
            {
                R.qa_o[buf][ap] = st3(org, 0.01f);
                R.qa_d[buf][ap] = wd;
                R.qa_own[buf][ap] = (int)((Q_NEE_ANY << 29) | (uint32_t)ks);
                ++ap;
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_accumulate(RenderDev R, float* accum, float* samples, int first_chunk,
                                                       int accumulate)
{
    uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    uint32_t npix = (uint32_t)R.cw * (uint32_t)R.ch;
    if (q >= npix) return;
    float4 s1, s2;
    if (first_chunk && !accumulate)
    {
        s1 = make_float4(0.f, 0.f, 0.f, 0.f);
        s2 = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    else
    {
        s1 = reinterpret_cast<float4*>(accum)[2 * q];
        s2 = reinterpret_cast<float4*>(accum)[2 * q + 1];
    }
    for (int k = 0; k < R.spc; ++k)
    {
        float4 L = R.out[(size_t)q * R.spc + k];
        // FileImagePlane::WriteSample: buf += L; buf.w += 1 (fileimageplane.cpp:29-34)
        s1.x += L.x;
        s1.y += L.y;
        s1.z += L.z;
        s1.w += 1.f;
        s2.x += L.x * L.x;
        s2.y += L.y * L.y;
        s2.z += L.z * L.z;
        if (samples)
        {
            size_t o = ((size_t)q * R.s_count + (size_t)(R.s0 - R.s_begin + k)) * 3;
            samples[o] = L.x;
            samples[o + 1] = L.y;
            samples[o + 2] = L.z;
        }
    }
    reinterpret_cast<float4*>(accum)[2 * q] = s1;
    reinterpret_cast<float4*>(accum)[2 * q + 1] = s2;
}

static bsh_status ensure_work(bsh_scene* s, size_t P, int K, int nl)
{
    RenderWork& w = s->work;
    BSH_HIP(w.ray_o.alloc(P));
    BSH_HIP(w.ray_d.alloc(P));
    BSH_HIP(w.thr.alloc(P));
    BSH_HIP(w.thr_nee.alloc(P));
    BSH_HIP(w.rad.alloc(P));
    BSH_HIP(w.out.alloc(P));
    BSH_HIP(w.hit.alloc(P));
    BSH_HIP(w.meta.alloc(P));
    BSH_HIP(w.nee_mask.alloc(P));
    size_t slots = P * (size_t)std::max(K, 1);
    BSH_HIP(w.nee_a.alloc(slots));
    BSH_HIP(w.nee_b.alloc(slots));
    BSH_HIP(w.nee_wi.alloc(slots));
    BSH_HIP(w.nee_res.alloc(slots));
    size_t qc = P * (size_t)(1 + nl), qa = P * (size_t)std::max(nl, 1);
    for (int k = 0; k < 2; ++k)
    {
        BSH_HIP(w.qc_o[k].alloc(qc));
        BSH_HIP(w.qc_d[k].alloc(qc));
        BSH_HIP(w.qc_own[k].alloc(qc));
        BSH_HIP(w.qa_o[k].alloc(qa));
        BSH_HIP(w.qa_d[k].alloc(qa));
        BSH_HIP(w.qa_own[k].alloc(qa));
    }
    BSH_HIP(w.stats.alloc(4));
    w.paths = P;
    w.slots = K;
    return BSH_OK;
}

bsh_status render_frame(bsh_scene* s, const bsh_render_params* p, float* accum, float* samples, hipStream_t st)
{
    if (!s || !p || !accum) return fail(BSH_ERR_INVALID, "bsh_render: NULL argument");
    if (!s->committed) return fail(BSH_ERR_STATE, "bsh_render: scene not committed");
    if (!s->has_cam) return fail(BSH_ERR_STATE, "bsh_render: no camera");
    if (p->flags & BSH_RENDER_NO_QUIRKS) return fail(BSH_ERR_UNSUPPORTED, "bsh_render: quirk-free mode not built");
    if (p->width <= 0 || p->height <= 0 || p->img_grid <= 0 || p->max_depth <= 0 || p->max_depth > 250)
        return fail(BSH_ERR_INVALID, "bsh_render: bad frame parameters");
    int nl = (int)s->lights.size();
    if (nl > kMaxLights) return fail(BSH_ERR_UNSUPPORTED, "bsh_render: more than 16 area lights");
    int cw = p->crop_w > 0 ? p->crop_w : p->width, ch = p->crop_h > 0 ? p->crop_h : p->height;
    int cx = p->crop_w > 0 ? p->crop_x : 0, cy = p->crop_h > 0 ? p->crop_y : 0;
    if (cx < 0 || cy < 0 || cx + cw > p->width || cy + ch > p->height)
        return fail(BSH_ERR_INVALID, "bsh_render: crop outside the image");
    int spp = p->img_grid * p->img_grid;
    int sb = p->spp_begin, sc = p->spp_count > 0 ? p->spp_count : spp - p->spp_begin;
    if (sb < 0 || sc <= 0 || sb + sc > spp) return fail(BSH_ERR_INVALID, "bsh_render: bad sample range");
    if ((uint64_t)p->width * p->height * spp > 0xffffffffull)
        return fail(BSH_ERR_UNSUPPORTED, "bsh_render: more than 2^32 camera samples per frame");
    size_t npix = (size_t)cw * ch;
    int spc = p->chunk_spp > 0 ? p->chunk_spp : (int)std::max<size_t>(1, ((size_t)1 << 21) / npix);
    spc = std::min(spc, sc);
    size_t P = npix * (size_t)spc;
    int K = 2 * std::max(nl, 1);
    if (P * (size_t)K >= ((size_t)1 << 29)) return fail(BSH_ERR_UNSUPPORTED, "bsh_render: chunk too large");
    BSH_HIP(hipSetDevice(s->ctx->device));
    bsh_status e = ensure_work(s, P, K, nl);
    if (e != BSH_OK) return e;
    RenderWork& w = s->work;
    const int launches = p->max_depth + 1;
    BSH_HIP(w.counters.alloc((size_t)4 * launches));

    RenderDev R;
    R.s = s->dev();
    R.cam = make_camera(s->cam);
    R.W = p->width;
    R.H = p->height;
    R.cx = cx;
    R.cy = cy;
    R.cw = cw;
    R.ch = ch;
    R.grid = p->img_grid;
    R.spp = spp;
    R.s_begin = sb;
    R.s_count = sc;
    R.depth = p->max_depth;
    R.seed = p->seed;
    R.bg0 = p->bg[0];
    R.bg1 = p->bg[1];
    R.bg2 = p->bg[2];
    R.nl = nl;
    R.K = K;
    R.ray_o = w.ray_o.p; R.ray_d = w.ray_d.p; R.thr = w.thr.p; R.thr_nee = w.thr_nee.p; R.rad = w.rad.p; R.out = w.out.p;
    R.hit = w.hit.p; R.meta = w.meta.p; R.nee_mask = w.nee_mask.p;
    R.nee_a = w.nee_a.p; R.nee_b = w.nee_b.p; R.nee_wi = w.nee_wi.p; R.nee_res = w.nee_res.p;
    for (int k = 0; k < 2; ++k)
    {
        R.qc_o[k] = w.qc_o[k].p; R.qc_d[k] = w.qc_d[k].p; R.qc_own[k] = w.qc_own[k].p;
        R.qa_o[k] = w.qa_o[k].p; R.qa_d[k] = w.qa_d[k].p; R.qa_own[k] = w.qa_own[k].p;
    }
    R.counters = w.counters.p;
    R.stats = w.stats.p;

    const bool count_visits = (p->flags & BSH_RENDER_COUNT) != 0;
    int trace_blocks = s->ctx->num_cus * 5;
    const int nchunks = (sc + spc - 1) / spc;
    // Event triplets around every trace/shade pair; read once after the frame.
    std::vector<hipEvent_t> evs((size_t)3 * nchunks * launches + 2);
    for (auto& ev : evs) BSH_HIP(hipEventCreate(&ev));
    size_t ei = 0;
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(kBlock), 0, st, (uint32_t*)w.stats.p, 8);
    BSH_HIP(hipEventRecord(evs[ei++], st));
    for (int s0 = sb; s0 < sb + sc; s0 += spc)
    {
        int n = std::min(spc, sb + sc - s0);
        R.spc = n;
        R.s0 = s0;
        R.P = (uint32_t)(npix * (size_t)n);
        dim3 pg((R.P + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_zero, dim3((4 * launches + kBlock - 1) / kBlock), dim3(kBlock), 0, st, w.counters.p,
                           4 * launches);
        hipLaunchKernelGGL(k_raygen, pg, dim3(kBlock), 0, st, R);
        for (int b = 0; b <= p->max_depth; ++b)
        {
            BSH_HIP(hipEventRecord(evs[ei++], st));
            if (count_visits)
                hipLaunchKernelGGL(k_trace_queue<true>, dim3(trace_blocks), dim3(kBlock), 0, st, R, b, b & 1);
            else
                hipLaunchKernelGGL(k_trace_queue<false>, dim3(trace_blocks), dim3(kBlock), 0, st, R, b, b & 1);
            BSH_HIP(hipEventRecord(evs[ei++], st));
            hipLaunchKernelGGL(k_shade, pg, dim3(kBlock), 0, st, R, b, b + 1 <= p->max_depth ? b + 1 : b);
            BSH_HIP(hipEventRecord(evs[ei++], st));
        }
        hipLaunchKernelGGL(k_accumulate, dim3((uint32_t)((npix + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, R,
                           accum, samples, s0 == sb ? 1 : 0, (p->flags & BSH_RENDER_ACCUMULATE) ? 1 : 0);
    }
    BSH_HIP(hipEventRecord(evs[ei++], st));
    BSH_HIP(hipGetLastError());
    BSH_HIP(hipEventSynchronize(evs[ei - 1]));
    double trace_ms = 0.0, shade_ms = 0.0;
    uint32_t ntrace = 0, nshade = 0;
    for (size_t k = 1; k + 2 < ei; k += 3)
    {
        float a = 0.f, b2 = 0.f;
        BSH_HIP(hipEventElapsedTime(&a, evs[k], evs[k + 1]));
        BSH_HIP(hipEventElapsedTime(&b2, evs[k + 1], evs[k + 2]));
        trace_ms += a;
        shade_ms += b2;
        ++ntrace;
        ++nshade;
    }
    float tot = 0.f;
    BSH_HIP(hipEventElapsedTime(&tot, evs[0], evs[ei - 1]));
    for (auto& ev : evs) (void)hipEventDestroy(ev);
    unsigned long long st2[4];
    BSH_HIP(hipMemcpy(st2, w.stats.p, sizeof(st2), hipMemcpyDeviceToHost));
    s->stats.rays_closest = st2[0];
    s->stats.rays_any = st2[1];
    s->stats.node_visits = count_visits ? st2[2] : 0;
    s->stats.tri_tests = count_visits ? st2[3] : 0;
    s->stats.camera_samples = (uint64_t)npix * (uint64_t)sc;
    s->stats.trace_ms = trace_ms;
    s->stats.shade_ms = shade_ms;
    s->stats.total_ms = tot;
    s->stats.trace_launches = ntrace;
    s->stats.shade_launches = nshade;
    return BSH_OK;
}

}  // namespace bsh
