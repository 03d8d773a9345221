// Wavefront GI over a job table.
//
// The per-pixel recursion of ImageRenderer + GiTracer::GetLi + DiTracer::GetDi
// (imagerenderer.cpp:61-101, gitracer.cpp:13-140, ditracer.cpp:57-239) is restructured
// as kernels over every in-flight path of a chunk:
//
//   raygen(per job segment) -> [trace -> shade] x depth -> trace -> shade(final)
//                           -> accumulate / SH-project (per job segment)
//
// A job is either a camera frame (bsh_render) or a probe set (bsh_bake_sh); a batch
// (bsh_batch_run) is a list of jobs over any number of committed scenes. Every path
// carries its job index, so one chunk mixes scenes freely. Chunks are cut at pixel /
// probe boundaries, so the per-pixel sums and per-probe projections are formed in the
// same order whatever the batch composition: a job gives bit-identical results alone or
// batched.
//
// shade(b) first resolves the next-event (NEE) rays its previous pass issued, then
// shades the closest hit of the extension ray, emits the new shadow (any-hit), MIS
// (closest-hit) and extension rays into two compacted queues (one atomic per wave and
// queue; each wave's rays grouped by direction octant and NEE slot), and finalises
// terminated paths. trace() is a persistent kernel: on 4-wide BVHs (k_trace_refill) each
// lane holds its own ray, closest and any-hit alike, and a wave claims new rays for its idle
// lanes in bulk; on binary BVHs (k_trace_queue) waves take 64 rays at a time from the
// closest queue, then the any queue.
//
// Guided jobs add, per bounce, the guide network's inputs of the extension hits and one
// inference launch before the shade, and training records compacted after each chunk.
//
// Random numbers come from the keyed sample stream; the counter of every draw is a
// fixed function of (bounce, light), which reproduces the reference's draw order
// (image sampler, BSDF pre-draw, per-GetDi light+BSDF draws, rand() light pick and
// Russian roulette) with light and BSDF samplers of one sample each.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "internal.h"
#include "shading.h"
#include "sh_dev.h"
#include "guide_dev.h"

namespace bsh {




constexpr int kBlock = 256;
// Lights the bounce-0 NEE loop covers (two ray slots each in a 64-bit mask per path).
constexpr int kMaxNeeLights = 32;
// LDS entries of the wavefront trace kernel's traversal stack (16 x 4 B x 256 threads = 16 KB
// per block); deeper entries go to per-thread global slots. Two lanes' trace grids of 3 blocks
// per CU then fit on a CU together with the shade kernels: measured on the 3M-triangle terrain
// at 1080p, 32 entries 2855 Mrays/s, 24 3021, 20 3042, 16 3064, 12 3057, 8 3035.
constexpr int kTraceLds = 16;
// Training records are kept for the first kRecVerts vertices of a path, 5 float4 each.
constexpr int kRecVerts = 4, kRecStride = 5 * kRecVerts;
static_assert(kGuideSeg == kBlock, "guide row segments are the blocks of k_guide_inputs");
constexpr int kMaxTraceBlocksPerCU = 5;
constexpr int kTraceBlocksPerLane = 3;   // the default when two lanes run
// k_trace_refill's launch bound: registers for 6 workgroups per CU (6 waves per SIMD, at most
// 80 VGPRs), so both lanes' default grids (3 per CU each) are resident at once whatever the
// node layout (the quantised layout's decode would otherwise take 89).
constexpr int kTraceResidentBlocks = 2 * kTraceBlocksPerLane;
// The shade kernel is held to 128 VGPRs: 4 waves per SIMD (at 129 it would run 3). Holding the
// plain variant to 5 or 6 waves (96 or 80 VGPRs, 24 or 50 of them spilled) made the 1080p
// terrain frame 1.1 and 1.5 % slower.
constexpr int kShadeWavesPerSimd = 4;
// Statistics buffer: 4 counters, then the guided-row partials, kRowStride apart (128 B), then per
// bounce launch b: closest rays, any-hit rays, and in counting mode node visits and triangle tests
// of the closest and of the any-hit rays (kLaunchStat counters each).
constexpr int kStatRows = 4, kRowParts = 64, kRowStride = 16, kStatTotal = kStatRows + kRowParts * kRowStride;
constexpr int kMaxLaunches = 256, kLaunchStat = 6, kStatAll = kStatTotal + kMaxLaunches * kLaunchStat;
static_assert(kMaxLaunches > 250, "a job's depth is at most 250, so a chunk has at most 251 launches");
// Idle lanes of a wave that trigger a refill in k_trace_refill (when the tuning leaves it 0).
// Measured on the 3M-triangle terrain at 1080p: 8 lanes 733 Mrays/s, 16 1058, 32 1314,
// 48 1381, 56 1393, 64 1353; whole 64-ray batches per wave (k_trace_queue) 1065.
constexpr int kDefaultRefill = 56;
// Queue entries a trace wave claims per atomic on the work counter when the run spans several
// scenes and the tuning leaves claim_rays 0 (a single scene claims exactly the idle lanes).
// Measured: 64 seeded scenes (1k-2M triangles) 2875 -> 3193 Mrays/s; on one 3M-triangle
// scene 128 cost 3 % (the small late-bounce launches lose balance).
constexpr uint32_t kClaimRays = 128;
// A single-chunk run at least this large is split in two for the two wavefront lanes.
constexpr size_t kSplitPaths = (size_t)1 << 22;

enum : uint32_t
{
    F_EXT = 1,      // extension ray in flight
    F_NEE = 2,      // NEE rays in flight
    F_TERM = 4,     // path ended; finalise once NEE results are in
    F_DONE = 8,     // final radiance written
    F_NOCLAMP = 16, // primary miss: GetLi returns bg + sum Le unclamped (gitracer.cpp:22-33)
    F_NMAP = 32     // the NEE rays in flight come from a normal-mapped vertex (reference estimator)
};

enum : uint32_t { Q_EXT = 0, Q_NEE_CLOSEST = 1, Q_NEE_ANY = 2 };

struct RenderDev
{
    const JobDev* jobs;
    uint32_t P;        // paths in this chunk
    int Kshift;        // NEE slots per path = 1 << Kshift
    float4 *ray_o, *ray_d, *thr, *thr_nee, *rad, *out;
    int4* hit;
    uint4* meta;       // x: key base, y: flags | bounce << 8 | nee lights << 16 | job << 24, z/w: BSDF sample
    unsigned long long* nee_mask;   // bit 2j: shadow ray of NEE slot j in flight, 2j + 1: its MIS ray
    float4 *nee_a, *nee_b, *nee_wi;
    int4* nee_res;
    float4 *qc_o[2], *qc_d[2], *qa_o[2], *qa_d[2];
    int *qc_own[2], *qa_own[2];
    uint32_t* counters;            // per launch L: closest count, any count (one 64-bit word, so a
                                   // shade wave claims both with one atomic), closest or joint work,
                                   // any work
    uint32_t* alive[2];            // live paths: shade b reads list b & 1, writes list (b + 1) & 1
    uint32_t* alive_n;             // per bounce: length of its live-path list (b >= 1)
    unsigned long long* stats;     // closest, any, node visits, triangle tests, then kRowParts guided-row partials
    uint32_t* stk_ovf;             // traversal stack entries beyond kTraceLds, [entry][trace thread]
    float4* gin;                   // guide inputs, 2 per row, rows compacted over the guided vertices
    const _Float16* gout;          // guide outputs, kGUsed per row (FP16)
    int* grow;                     // per path: its row this pass, -1 when the vertex is not guided
    uint32_t* gseg;                // per block of kBlock paths: guided rows packed at the front of its segment
    float4* rec;                   // per-path training records: kRecVerts x 5 float4
    float* rec_out;                // appended records, 12 floats each
    uint32_t rec_cap;
    uint32_t* rec_count;
    float4* pdir;                  // guided probe jobs: per path, probe ray direction and projection weight
    const float4* nodes1;          // when every job of the run traces the same scene: its nodes and
    const float4* tris1;           // triangles (else null: each ray looks its scene up by its job)
};

__device__ __forceinline__ uint32_t rand_pick_counter(int b) { return (uint32_t)((b - 1) + (b > 4 ? b - 4 : 0)); }

__global__ __launch_bounds__(kBlock) void k_zero(uint32_t* p, int n)
{
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Every path of a chunk enqueues exactly one primary ray at its own slot.
__global__ void k_set_count(uint32_t* counters, unsigned long long* stats, uint32_t P)
{
    (void)stats;
    if (threadIdx.x == 0) counters[0] = P;
}

// After a chunk: its closest and any-hit ray totals, from every launch's queue counts (added
// atomically: the chunks of two lanes may finish at the same time).
__global__ void k_chunk_rays(const uint32_t* counters, int launches, unsigned long long* stats)
{
    if (threadIdx.x == 0)
    {
        unsigned long long c = 0, a = 0;
        for (int L = 0; L < launches; ++L)
        {
            c += counters[4 * L + 0];
            a += counters[4 * L + 1];
            atomicAdd(&stats[kStatTotal + kLaunchStat * L + 0], (unsigned long long)counters[4 * L + 0]);
            atomicAdd(&stats[kStatTotal + kLaunchStat * L + 1], (unsigned long long)counters[4 * L + 1]);
        }
        atomicAdd(&stats[0], c);
        atomicAdd(&stats[1], a);
    }
}

__device__ __forceinline__ void init_path(const RenderDev& R, uint32_t i, f3 o, f3 d, float tmin, float tmax,
                                          uint32_t base, uint32_t job)
{
    R.ray_o[i] = st3(o, 0.f);
    R.ray_d[i] = st3(d, 0.f);
    R.thr[i] = make_float4(1.f, 1.f, 1.f, 0.f);
    R.rad[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    R.meta[i] = make_uint4(base, F_EXT | (job << 24), 0u, 0u);
    R.qc_o[0][i] = st3(o, tmin);
    R.qc_d[0][i] = st3(d, tmax);
    R.qc_own[0][i] = (int)((Q_EXT << 29) | i);
    if (R.rec)
    {
        for (int k = 0; k < kRecVerts; ++k) R.rec[(size_t)i * kRecStride + 5 * k + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Camera raygen for pixels [u0, u0+nu) of camera job `job`, written at chunk offset c0:
// ImageRenderer sample loop + PerscpectiveCamera::GenerateRay (imagerenderer.cpp:78-87,
// perspective_camera.cpp:25-38).
__global__ __launch_bounds__(kBlock) void k_raygen(RenderDev R, uint32_t job, uint32_t u0, uint32_t nu, uint32_t c0)
{
    const JobDev& J = R.jobs[job];
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= nu * (uint32_t)J.s_count) return;
    uint32_t i = c0 + t;
    uint32_t q = u0 + t / (uint32_t)J.s_count;                 // pixel within the crop
    int s = J.s_begin + (int)(t % (uint32_t)J.s_count);
    int x = J.cx + (int)(q % (uint32_t)J.cw);
    int y = J.cy + (int)(q / (uint32_t)J.cw);
    uint32_t g = ((uint32_t)y * (uint32_t)J.W + (uint32_t)x) * (uint32_t)J.spp + (uint32_t)s;
    uint32_t base = ks_base(J.seed, g);
    // CMJ and Sobol patterns are drawn once per pixel, keyed by the pixel's sample 0
    uint32_t base0 = J.img_smp == SMP_CMJ || J.img_smp == SMP_SOBOL ? ks_base(J.seed, g - (uint32_t)s) : base;
    float sx, sy;
    image_sample2d(J.img_smp, J.grid, s, base, base0, sx, sy);
    float imx = (float)x / (float)J.W + (1.f / (float)J.W) * sx;
    float imy = (float)y / (float)J.H + (1.f / (float)J.H) * sy;
    float hx = imx - 0.5f, hy = imy - 0.5f;
    float csx = hx * J.cam.dimx, csy = hy * J.cam.dimy;
    f3 fwd = mk(J.cam.fwd[0], J.cam.fwd[1], J.cam.fwd[2]);
    f3 right = mk(J.cam.right[0], J.cam.right[1], J.cam.right[2]);
    f3 up = mk(J.cam.up[0], J.cam.up[1], J.cam.up[2]);
    f3 d = normalize(J.cam.znear * fwd + csx * right + csy * up);
    f3 o = mk(J.cam.eye[0], J.cam.eye[1], J.cam.eye[2]);
    init_path(R, i, o, d, J.cam.znear, J.cam.zfar, base, job);
}

// Stratified uniform-sphere direction of probe ray j (sqrt(N) = g cells per axis),
// drawn from stream KS_PROBE.
__device__ __forceinline__ f3 probe_dir(uint32_t base, int j, int g)
{
    float ux, uy;
    ks_sample2d(base, KS_PROBE, 0, ux, uy);
    float cell = 1.f / (float)g;
    float u = ((float)(j % g) + ux) * cell;
    float v = ((float)(j / g) + uy) * cell;
    float z = 1.f - 2.f * u;
    float r = sqrtf(fmaxf(0.f, 1.f - z * z));
    float phi = 2.f * kPi * v;
    return mk(r * cosf(phi), r * sinf(phi), z);
}

// Network inputs of the probes [u0, u0+nu) of a guided probe job, rows row0.. : the probe
// position in the scene box, roughness 0 and a zero normal (no surface has either), so the
// network tells the free-space vertex of a probe from surface vertices.
__global__ __launch_bounds__(kBlock) void k_probe_guide_inputs(RenderDev R, uint32_t job, uint32_t u0, uint32_t nu,
                                                              uint32_t row0)
{
    const JobDev& J = R.jobs[job];
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= nu) return;
    const uint32_t p = u0 + t;
    R.gin[2 * (row0 + t)] = make_float4((J.probes[3 * p] - J.glo[0]) * J.ginv[0],
                                        (J.probes[3 * p + 1] - J.glo[1]) * J.ginv[1],
                                        (J.probes[3 * p + 2] - J.glo[2]) * J.ginv[2], 0.f);
    R.gin[2 * (row0 + t) + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Keyed-stream counters of the probe vertex's guide draws (the per-bounce guide draws use
// counters 6b .. 6b+5).
constexpr uint32_t kProbeGuideCtr = 0xFFF000u;

// Probe ray j of probe p. Plain: the stratified uniform direction of cell j, weight 1. With
// JOB_PROBE_DIR: with probability alpha (the guide's BSDF fraction, here the uniform
// fraction) that direction, else a sample of the probe's vMF mixture (network row row0 + p - u0);
// the projection weight is 1 / (4 pi p_mix), p_mix = alpha / (4 pi) + (1 - alpha) q. Summed
// over the N rays, the cell-restricted uniform densities add up to the uniform density, so
// (4 pi / N) sum_j w_j L_j Y(w_j) stays an unbiased projection. A probe whose network output is
// not finite keeps the plain directions and weight 1.
__global__ __launch_bounds__(kBlock) void k_raygen_probe(RenderDev R, uint32_t job, uint32_t u0, uint32_t nu,
                                                       uint32_t c0, uint32_t row0)
{
    const JobDev& J = R.jobs[job];
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= nu * (uint32_t)J.N) return;
    uint32_t i = c0 + t;
    uint32_t p = u0 + t / (uint32_t)J.N;
    int j = (int)(t % (uint32_t)J.N);
    uint32_t base = ks_base(J.seed, p * (uint32_t)J.N + (uint32_t)j);
    f3 d = probe_dir(base, j, J.g);
    f3 o = mk(J.probes[3 * p], J.probes[3 * p + 1], J.probes[3 * p + 2]);
    float pmix = 0.25f / kPi;
    float w = 1.f;
    if (J.flags & JOB_PROBE_DIR)
    {
        const _Float16* row = R.gout + (size_t)(row0 + (p - u0)) * kGUsed;
        const Lobes G = guide_lobes(row);
        if (lobes_finite(G))
        {
            const float alpha = J.bsdf_frac;
            const float usel = ks_float(ks_u32(base, KS_GUIDE, kProbeGuideCtr));
            if (!(usel < alpha))
                d = guide_sample(G, row, ks_float(ks_u32(base, KS_GUIDE, kProbeGuideCtr + 1)),
                                 ks_float(ks_u32(base, KS_GUIDE, kProbeGuideCtr + 2)),
                                 ks_float(ks_u32(base, KS_GUIDE, kProbeGuideCtr + 3)));
            pmix = alpha * (0.25f / kPi) + (1.f - alpha) * guide_pdf(G, d);
            w = 1.f / (4.f * kPi * pmix);
        }
        R.pdir[i] = make_float4(d.x, d.y, d.z, w);
    }
    init_path(R, i, o, d, J.tmin, 1000000.f, base, job);
    if ((J.flags & JOB_PROBE_DIR) && (J.flags & JOB_RECORD))
    {
        // the probe vertex's training record (slot 0): radiance before it is 0 and its
        // throughput 1, so the incident radiance is the path's whole radiance; normal 0 and
        // .w = 0 mark a record without a cosine factor
        float4* rc = R.rec + (size_t)i * kRecStride;
        rc[0] = make_float4((o.x - J.glo[0]) * J.ginv[0], (o.y - J.glo[1]) * J.ginv[1], (o.z - J.glo[2]) * J.ginv[2], 0.f);
        rc[1] = make_float4(0.f, 0.f, 0.f, 0.f);
        rc[2] = make_float4(d.x, d.y, d.z, pmix);
        rc[3] = make_float4(1.f, 1.f, 1.f, 0.f);
        rc[4] = make_float4(0.f, 0.f, 0.f, 1.f);
    }
}

__device__ __forceinline__ uint32_t path_of(const RenderDev& R, uint32_t kind, uint32_t idx)
{
    return kind == Q_EXT ? idx : (idx >> R.Kshift);
}

// Network inputs of every path whose extension ray hit something at a guided vertex of pass b:
// position in the scene box, roughness (GGX alpha, 1 otherwise), shading normal facing the
// viewer. Each block packs its rows at the front of its own kBlock-row segment (a block-wide
// prefix count) and stores how many it wrote; each path records its row. The block's count
// is also added to the guided-row statistic (one integer atomic per block). From bounce 1 on
// the threads walk the live-path list the shade of this bounce will walk, not all paths.
__global__ __launch_bounds__(kBlock) void k_guide_inputs(RenderDev R, int b)
{
    __shared__ uint32_t wsum[kBlock / 64];
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    uint32_t i = t;
    bool have = t < R.P;
    if (b > 0)
    {
        have = t < R.alive_n[b];
        i = have ? R.alive[b & 1][t] : 0u;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f), c = make_float4(0.f, 0.f, 0.f, 0.f);
    bool want = false;
    uint4 meta = make_uint4(0u, 0u, 0u, 0u);
    if (have) meta = R.meta[i];
    const JobDev& J = R.jobs[meta.y >> 24];
    if (have && (J.flags & JOB_GUIDED) && (J.guide_bounces == 0 || b < J.guide_bounces) && (meta.y & F_EXT) &&
        !(meta.y & F_DONE))
    {
        int4 hr = R.hit[i];
        if (hr.x >= 0)
        {
            want = true;
            f3 p1, p2, p3, n1, n2, n3;
            Mat m;
            tri_verts(J.s, hr.x, p1, p2, p3, n1, n2, n3, m);
            Frame h = fill_hit(p1, p2, p3, n1, n2, n3, __int_as_float(hr.z), __int_as_float(hr.w));
            f3 wo = -ld3(R.ray_d[i]);
            f3 n = dot(h.n, wo) < 0.f ? -h.n : h.n;
            a = make_float4((h.p.x - J.glo[0]) * J.ginv[0], (h.p.y - J.glo[1]) * J.ginv[1],
                            (h.p.z - J.glo[2]) * J.ginv[2], m.type == 1 ? m.alpha : 1.f);
            c = make_float4(n.x, n.y, n.z, 1.f);
        }
    }
    const unsigned long long m = __ballot(want);
    if (lane == 0) wsum[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = blockIdx.x * kBlock, total = 0;
    for (int w = 0; w < kBlock / 64; ++w)
    {
        if (w < wave) base += wsum[w];
        total += wsum[w];
    }
    if (threadIdx.x == 0)
    {
        R.gseg[blockIdx.x] = total;
        // guided rows: an exact integer sum over kRowParts partial counters on separate cache
        // lines (one shared counter serialises the blocks' atomics)
        if (total) atomicAdd(&R.stats[kStatRows + kRowStride * (blockIdx.x % kRowParts)], (unsigned long long)total);
    }
    if (!have) return;
    int row = -1;
    if (want)
    {
        row = (int)(base + __popcll(m & ((1ull << lane) - 1ull)));
        R.gin[2 * row] = a;
        R.gin[2 * row + 1] = c;
    }
    R.grow[i] = row;
}

template <bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_queue(RenderDev R, int L, int buf)
{
    __shared__ uint32_t stk[kTraceLds * kBlock];
    const int lane = threadIdx.x & 63;
    uint32_t* my = stk + threadIdx.x;
    const int ostride = gridDim.x * kBlock;
    uint32_t* ovf = R.stk_ovf + blockIdx.x * kBlock + threadIdx.x;
    uint32_t vis_nodes = 0, vis_tris = 0;
    for (int pass = 0; pass < 2; ++pass)
    {
        if (kCount && pass == 1)
        {
            // the closest-hit queue's counts, then the any-hit queue's
            atomicAdd(&R.stats[kStatTotal + kLaunchStat * L + 2], (unsigned long long)vis_nodes);
            atomicAdd(&R.stats[kStatTotal + kLaunchStat * L + 3], (unsigned long long)vis_tris);
            atomicAdd(&R.stats[2], (unsigned long long)vis_nodes);
            atomicAdd(&R.stats[3], (unsigned long long)vis_tris);
            vis_nodes = vis_tris = 0;
        }
        const uint32_t count = R.counters[4 * L + pass];
        uint32_t* work = &R.counters[4 * L + 2 + pass];
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
                uint32_t job = R.meta[path_of(R, kind, idx)].y >> 24;
                const SceneDev S = R.jobs[job].s;
                TraceResult res;
                TraceCounts cnt{0u, 0u};
                if (pass == 0)
                {
                    bool h = S.width == 4
                                 ? traverse4<false, kCount, kTraceLds>(S, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w,
                                                                       b.w, my, kBlock, res, cnt, ovf, ostride)
                                 : traverse<false, kCount, kTraceLds>(S, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w,
                                                                      b.w, my, kBlock, res, cnt, ovf, ostride);
                    int4 o = make_int4(h ? res.prim : -1, __float_as_int(res.t), __float_as_int(res.b1),
                                       __float_as_int(res.b2));
                    if (kind == Q_EXT) R.hit[idx] = o;
                    else R.nee_res[idx] = o;
                }
                else
                {
                    bool h = S.width == 4
                                 ? traverse4<true, kCount, kTraceLds>(S, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w,
                                                                      b.w, my, kBlock, res, cnt, ovf, ostride)
                                 : traverse<true, kCount, kTraceLds>(S, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w,
                                                                     b.w, my, kBlock, res, cnt, ovf, ostride);
                    R.nee_res[idx] = make_int4(h ? 1 : 0, 0, 0, 0);
                }
                vis_nodes += cnt.nodes;
                vis_tris += cnt.tris;
            }
        }
    }
    if (kCount)
    {
        atomicAdd(&R.stats[kStatTotal + kLaunchStat * L + 4], (unsigned long long)vis_nodes);
        atomicAdd(&R.stats[kStatTotal + kLaunchStat * L + 5], (unsigned long long)vis_tris);
        atomicAdd(&R.stats[2], (unsigned long long)vis_nodes);
        atomicAdd(&R.stats[3], (unsigned long long)vis_tris);
    }
}

// The trace stack with its top entry in a register: entries 0 .. sp-2 in LDS (past kTraceLds in
// the global overflow), entry sp-1 in `top`. A pop returns the register at once and refills it
// from LDS, so the next node's fetch does not wait for an LDS read.
__device__ __forceinline__ void push_top(lu32* stk, gu32* ovf, int ostride, int& sp, uint32_t& top, uint32_t v)
{
    if (sp >= kMaxStack4) return;   // the builder bounds the depth, so this never drops an entry
    if (sp > 0)
    {
        const int p = sp - 1;
        if (p < kTraceLds) stk[p * kBlock] = top;
        else ovf[(p - kTraceLds) * ostride] = top;
    }
    top = v;
    ++sp;
}

__device__ __forceinline__ uint32_t pop_top(lu32* stk, gu32* ovf, int ostride, int& sp, uint32_t& top)
{
    const uint32_t v = top;
    --sp;
    if (sp > 0)
    {
        const int p = sp - 1;
        top = p < kTraceLds ? stk[p * kBlock] : ovf[(p - kTraceLds) * ostride];
    }
    return v;
}

// 4-wide traversal with per-lane refill. The closest-hit queue and the any-hit queue form one
// index range [0, nc + na) drawn by one work counter. Every lane holds its own ray state and
// advances it by one step per loop iteration (a leaf triangle, the sort of a visit's children
// and the next visit, as far as each applies); a finished lane idles until `refill` or
// more lanes of its wave are idle (or the whole wave is), when the wave claims exactly that
// many new rays with one atomic and hands them to the idle lanes in lane order. Lanes of one
// wave may hold closest and any-hit rays of different scenes at once. Node visits and
// accept rules are those of traverse4, so every ray's result is the same. kQ: every scene of
// the launch has GpuNodeQ nodes (else every one has GpuNode4 nodes). kOne: every job traces
// one scene (R.nodes1), whose arrays then stay in scalar registers.
template <bool kCount, bool kQ, bool kOne>
__global__ __launch_bounds__(kBlock, kTraceResidentBlocks) void k_trace_refill(RenderDev R, int L, int buf, int refill,
                                                                               uint32_t claim_rays)
{
    __shared__ uint32_t stk[kTraceLds * kBlock];
    const int lane = threadIdx.x & 63;
    lu32* my = (lu32*)(stk + threadIdx.x);
    const int ostride = gridDim.x * kBlock;
    gu32* ovf = (gu32*)(R.stk_ovf + blockIdx.x * kBlock + threadIdx.x);
    const uint32_t nc = R.counters[4 * L + 0];
    const uint32_t total = nc + R.counters[4 * L + 1];
    uint32_t* work = &R.counters[4 * L + 2];
    const float kInf = __builtin_inff();
    uint32_t vis_nodes = 0, vis_tris = 0, vis_nodes_any = 0, vis_tris_any = 0;
    bool active = false, any = false, exhausted = false;
    // the wave's reserve: queue entries [rb, re) it has claimed and not yet handed to a lane
    uint32_t rb = 0, re = 0;
    f3 o = mk(0.f, 0.f, 0.f), d = mk(0.f, 0.f, 0.f);
    float ix = 0.f, iy = 0.f, iz = 0.f, ox = 0.f, oy = 0.f, oz = 0.f, tmin = 0.f, tcur = 0.f, hb1 = 0.f, hb2 = 0.f;
    int prim = -1, sp = 0;
    uint32_t node = 0, own = 0, top = 0;
    // the last visit's hit leaf children still to test (bit c: child c, tested in child order,
    // triangle loff of the lowest one next) with that node's refs and counts, and its internal
    // children (key = entry distance, kInf if missed), not yet culled and pushed while !sorted
    uint32_t lmask = 0u;
    int loff = 0, lr0 = 0, lr1 = 0, lr2 = 0, lr3 = 0, lc0 = 0, lc1 = 0, lc2 = 0, lc3 = 0;
    float key[4] = {0.f, 0.f, 0.f, 0.f};
    uint32_t id[4] = {0u, 0u, 0u, 0u};
    bool sorted = true;
    const float4* nodes = kOne ? R.nodes1 : nullptr;
    const float4* tris = kOne ? R.tris1 : nullptr;
    for (;;)
    {
        // ---- refill the idle lanes (wave-uniform decision) ----
        // Idle lanes take entries from the reserve first. A wave whose reserve is short claims
        // claim_rays more with one atomic (0: exactly its idle lanes, so no reserve is kept):
        // with many small scenes, whose rays finish fast, the counter then sees a fraction of
        // the refills, and a wave still takes its rays in queue order.
        const unsigned long long idle = __ballot(!active);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if ((!exhausted || rb < re) && (nidle >= (uint32_t)refill || nidle == 64u))
        {
            const uint32_t from_res = min(re - rb, nidle);
            // >= the lanes it must cover, whatever claim_rays the host let through
            const uint32_t want = claim_rays ? max(claim_rays, nidle) : nidle;
            uint32_t nb = 0;
            const bool claim = from_res < nidle && !exhausted;
            if (claim)
            {
                if (lane == 0) nb = atomicAdd(work, want);
                nb = __builtin_amdgcn_readfirstlane(__shfl(nb, 0));
                if (nb + want >= total) exhausted = true;
            }
            if (!active)
            {
                const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                const uint32_t q = rank < from_res ? rb + rank : claim ? nb + (rank - from_res) : total;
                if (q < total)
                {
                    any = q >= nc;
                    const uint32_t k = any ? q - nc : q;
                    const float4 a = gload(any ? R.qa_o[buf] : R.qc_o[buf], k);
                    const float4 b = gload(any ? R.qa_d[buf] : R.qc_d[buf], k);
                    own = (uint32_t)(any ? R.qa_own[buf][k] : R.qc_own[buf][k]);
                    if (!kOne)
                    {
                        const uint32_t job = R.meta[path_of(R, own >> 29, own & ((1u << 29) - 1u))].y >> 24;
                        nodes = R.jobs[job].s.nodes;
                        tris = R.jobs[job].s.tris;
                    }
                    o = mk(a.x, a.y, a.z);
                    d = mk(b.x, b.y, b.z);
                    const float dx = fabsf(d.x) < 1e-20f ? copysignf(1e-20f, d.x) : d.x;
                    const float dy = fabsf(d.y) < 1e-20f ? copysignf(1e-20f, d.y) : d.y;
                    const float dz = fabsf(d.z) < 1e-20f ? copysignf(1e-20f, d.z) : d.z;
                    ix = 1.f / dx;
                    iy = 1.f / dy;
                    iz = 1.f / dz;
                    ox = o.x * ix;
                    oy = o.y * iy;
                    oz = o.z * iz;
                    tmin = a.w;
                    tcur = b.w;
                    prim = -1;
                    hb1 = hb2 = 0.f;
                    node = 0;
                    sp = 0;
                    lmask = 0u;
                    sorted = true;
                    active = true;
                }
            }
            // the reserve left (rb <= re always: a claim hands out at most the idle lanes it
            // was sized for)
            if (claim)
            {
                rb = nb + (nidle - from_res);
                re = nb + want;
            }
            else
                rb += from_res;
        }
        if (__ballot(active) == 0ull) break;
        if (!active) continue;
        // ---- one step: a triangle of the pending leaves, then, once a lane's leaves are
        // done, its node's internal children and the next node's visit ----
        // A visit's leaf triangles are tested one per iteration, before its internal children
        // are culled against the updated tcur, sorted and pushed (the order traverse4 tests and
        // culls in); a lane whose last leaf triangle, or whose stack pop, settles its next
        // node visits that node in the same iteration. A wave so runs a triangle test, a sort
        // and a visit per iteration over whichever of its lanes need them, instead of looping
        // over its busiest lane's triangles while the other lanes wait.
        bool done = false;
        if (lmask)
        {
            const int c = __builtin_ctz(lmask);
            const int lref = c == 0 ? lr0 : c == 1 ? lr1 : c == 2 ? lr2 : lr3;
            const int lcnt = c == 0 ? lc0 : c == 1 ? lc1 : c == 2 ? lc2 : lc3;
            const int tq = lref + loff;
            const float4 v0 = gload(tris, 3 * (size_t)tq + 0);
            const float4 a1 = gload(tris, 3 * (size_t)tq + 1);
            const float4 a2 = gload(tris, 3 * (size_t)tq + 2);
            if (kCount) (any ? vis_tris_any : vis_tris)++;
            float t, u, v;
            if (intersect_face(ld3(v0), ld3(a1), ld3(a2), o, d, tmin, tcur, t, u, v))
            {
                prim = __builtin_bit_cast(int, v0.w);
                tcur = t;
                hb1 = u;
                hb2 = v;
                if (any) done = true;
            }
            if (++loff == lcnt)
            {
                lmask &= lmask - 1u;
                loff = 0;
            }
        }
        if (!done && !sorted && lmask == 0u)
        {
            // internal children still in front of the (possibly updated) closest hit, nearest
            // first: visit it, push the others
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (!(key[c] <= tcur)) key[c] = kInf;
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
                if (sp == 0) done = true;
                else node = pop_top(my, ovf, ostride, sp, top);
            }
            else
            {
#pragma unroll
                for (int k = 3; k >= 1; --k)
                    if (k < n) push_top(my, ovf, ostride, sp, top, id[k]);
                node = id[0];
            }
            sorted = true;
        }
        if (!done && sorted && lmask == 0u)
        {
            // visit the next node; its internal children are sorted in the next iteration,
            // after its leaf triangles
            if (kCount) (any ? vis_nodes_any : vis_nodes)++;
            int ref[4], cn[4];
            float tn[4];
            bool hitc[4];
            node4_test<kQ>(nodes, node, ix, iy, iz, ox, oy, oz, tmin, tcur, tn, hitc, ref, cn);
#pragma unroll
            for (int c = 0; c < 4; ++c)
            {
                key[c] = (hitc[c] && cn[c] == 0) ? tn[c] : kInf;
                id[c] = (uint32_t)ref[c];
            }
            // the hit leaf children, tested from the next iteration on
            lmask = 0u;
#pragma unroll
            for (int c = 0; c < 4; ++c) lmask |= (hitc[c] && cn[c] > 0) ? 1u << c : 0u;
            loff = 0;
            lr0 = ref[0]; lr1 = ref[1]; lr2 = ref[2]; lr3 = ref[3];
            lc0 = cn[0]; lc1 = cn[1]; lc2 = cn[2]; lc3 = cn[3];
            sorted = false;
        }
        if (done)
        {
            const uint32_t kind = own >> 29, idx = own & ((1u << 29) - 1u);
            if (any) R.nee_res[idx] = make_int4(prim >= 0 ? 1 : 0, 0, 0, 0);
            else
            {
                const int4 res = make_int4(prim, __float_as_int(tcur), __float_as_int(hb1), __float_as_int(hb2));
                if (kind == Q_EXT) R.hit[idx] = res;
                else R.nee_res[idx] = res;
            }
            active = false;
        }
    }
    if (kCount)
    {
        unsigned long long* ls = R.stats + kStatTotal + kLaunchStat * L;
        atomicAdd(&ls[2], (unsigned long long)vis_nodes);
        atomicAdd(&ls[3], (unsigned long long)vis_tris);
        atomicAdd(&ls[4], (unsigned long long)vis_nodes_any);
        atomicAdd(&ls[5], (unsigned long long)vis_tris_any);
        atomicAdd(&R.stats[2], (unsigned long long)(vis_nodes + vis_nodes_any));
        atomicAdd(&R.stats[3], (unsigned long long)(vis_tris + vis_tris_any));
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

// kFull = false compiles only the reference estimator (no quirk-free, guided or record
// code), keeping the plain path's registers low. kOne: the run has a single job, so its
// record and scene pointers are read with scalar loads at the start instead of per lane
// after the path's meta word. kLights (with kFull) adds point, directional and environment
// lights and the materials beyond Lambert, Microfacet and Emissive (material_dev.h), kept out
// of the other variants' registers.
template <bool kFull, bool kOne, bool kLights = false>
__global__ __launch_bounds__(kBlock, kShadeWavesPerSimd) void k_shade(RenderDev R, int b, int Lout)
{
    // Bounce 0 shades every path of the chunk; a later bounce only the paths the previous
    // shade left live, in the order it listed them (its wave order, so neighbouring paths
    // stay together). A wave then holds live paths only, instead of the survivors scattered
    // over all paths.
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    uint32_t i = t;
    bool have = t < R.P;
    if (b > 0)
    {
        have = t < R.alive_n[b];
        i = have ? R.alive[b & 1][t] : 0u;
    }
    const int K = 1 << R.Kshift;
    bool emit_ext = false, keep = false;
    unsigned long long emit_mask = 0;
    f3 org = mk(0.f, 0.f, 0.f), ext_d = mk(0.f, 0.f, 0.f);
    if (have)
    {
        uint4 meta = R.meta[i];
        // the extension hit is read with the meta word rather than after it: one level less
        // in the chain of dependent loads (wasted only by paths with just NEE rays pending)
        const int4 hr = R.hit[i];
        uint32_t fl = meta.y & 0xffu;
        int pb = (int)((meta.y >> 8) & 0xffu);
        int ncnt = (int)((meta.y >> 16) & 0xffu);
        const uint32_t job = kOne ? 0u : meta.y >> 24;
        const JobDev& J = R.jobs[job];
        const SceneDev S = J.s;
        const int nl = S.nlights;
        const int jflags = kFull ? J.flags : 0;
        const bool nq = (jflags & JOB_NO_QUIRKS) != 0;
        if (!(fl & F_DONE))
        {
            const float4 rad4 = R.rad[i];
            f3 rad = ld3(rad4);
            // kLights runs keep the path's rand() draws in a running counter (rad.w): Glass and
            // MixedMaterial draw inside Sample, so the draw count depends on the path
            RandStream rs{meta.x, kLights ? __float_as_uint(rad4.w) : 0u};
            const bool had_nee = (fl & F_NEE) != 0;
            // ---- 1. resolve the NEE rays issued at bounce pb (DiTracer::GetDi tail) ----
            if (fl & F_NEE)
            {
                const unsigned long long mask = R.nee_mask[i];
                f3 T = ld3(R.thr_nee[i]);
                f3 ps = ld3(R.ray_o[i]);
                for (int j = 0; j < ncnt; ++j)
                {
                    f3 local = mk(0.f, 0.f, 0.f);
                    size_t ks = (size_t)i * K + 2 * j;
                    if (mask & (1ull << (2 * j)))
                    {
                        if (R.nee_res[ks].x == 0) local = local + ld3(R.nee_a[ks]);
                    }
                    // a normal-mapped vertex traced two BSDF rays for this light: the one that saw
                    // the normal mapped twice (its light sample was unoccluded, so Evaluate ran
                    // first) in slot 2j + 1, the other in slot 2 ncnt + 2j + 1
                    int bbit = 2 * j + 1;
                    size_t bks = ks + 1;
                    if (kLights && (fl & F_NMAP) && (mask & (1ull << (2 * j))) && R.nee_res[ks].x != 0)
                    {
                        bbit = 2 * ncnt + 2 * j + 1;
                        bks = (size_t)i * K + bbit;
                    }
                    if (mask & (1ull << bbit))
                    {
                        int4 r = R.nee_res[bks];
                        float4 nb = R.nee_b[bks];
                        int Lid = (int)nb.z;
                        if (r.x >= 0)
                        {
                            // shadowhit.bundle->GetAreaLight() == &light (ditracer.cpp:202); quirk-free:
                            // the hit must be this light's own triangle
                            bool mine = nq ? (r.x == S.lights[Lid].tri) : (tri_bundle_light(S, r.x) == Lid);
                            if (mine)
                            {
                                f3 p1, p2, p3, n1, n2, n3;
                                Mat lm;
                                tri_verts(S, r.x, p1, p2, p3, n1, n2, n3, lm);
                                Frame hh = fill_hit(p1, p2, p3, n1, n2, n3, __int_as_float(r.z), __int_as_float(r.w));
                                f3 wi = ld3(R.nee_wi[bks]);
                                f3 d = hh.p - ps;
                                f3 le = mk(0.f, 0.f, 0.f);
                                if (dot(hh.n, -wi) > 0.f)
                                {
                                    float d2inv = 1.f / sqnorm(d);
                                    if (lm.type == 2) le = emissive_le(lm.c, hh.n, -wi) * (nq ? 1.f : d2inv);
                                    else le = mk(0.f, 0.f, 0.f) * d2inv;  // Material::GetLe default
                                }
                                if (sqnorm(le) > 0.f)
                                {
                                    float4 na = R.nee_a[bks];
                                    local = local + le * ld3(na) * na.w * nb.x * nb.y;
                                }
                            }
                        }
                        else if (kLights && Lid >= S.narea)
                        {
                            // the ray left the scene: the light's GetLe, nonzero for environment
                            // lights (ditracer.cpp:218-222)
                            f3 le = other_light_le(S, S.lights[Lid], ld3(R.nee_wi[bks]));
                            if (sqnorm(le) > 0.f)
                            {
                                float4 na = R.nee_a[bks];
                                local = local + le * ld3(na) * na.w * nb.x * nb.y;
                            }
                        }
                    }
                    // the reference scales its one uniformly picked light by nl; a quirk-free pick's
                    // 1 / probability is in the slot's contributions already
                    if (pb == 0 || nq) rad = rad + T * local;
                    else rad = rad + (T * local) * (float)nl;
                }
                fl &= ~(F_NEE | F_NMAP);
            }
            // record slot of path vertex b: b, or b + 1 on guided probe paths, whose slot 0 is
            // the probe itself
            const int roff = (jflags & JOB_PROBE_DIR) ? 1 : 0;
            // training record of the vertex shaded in the previous pass: radiance so far
            if ((jflags & JOB_RECORD) && b >= 1 && b - 1 + roff < kRecVerts)
            {
                float4 v = R.rec[(size_t)i * kRecStride + (b - 1 + roff) * 5 + 2];
                if (v.w > 0.f) R.rec[(size_t)i * kRecStride + (b - 1 + roff) * 5 + 4] = st3(rad, 1.f);
            }
            (void)had_nee;
            // ---- 2. shade the extension hit (GiTracer bounce loop body) ----
            if (fl & F_EXT)
            {
                fl &= ~F_EXT;
                if (hr.x < 0)
                {
                    if (b == 0)
                    {
                        rad = mk(J.bg[0], J.bg[1], J.bg[2]);
                        // a camera ray that leaves the scene also gets every light's GetLe
                        // (gitracer.cpp:20-30), which only environment lights make nonzero
                        if (kLights)
                            for (int l = S.narea; l < nl; ++l) rad = rad + other_light_le(S, S.lights[l], ld3(R.ray_d[i]));
                        fl |= F_NOCLAMP;
                    }
                    fl |= F_TERM;
                }
                else
                {
                    f3 p1, p2, p3, n1, n2, n3;
                    Mat m;
                    tri_verts(S, hr.x, p1, p2, p3, n1, n2, n3, m);
                    Frame h;
                    TexCtx tx{S.tex, S.tex_desc, S.mat_maps, 0.f, 0.f};
                    if (kLights && S.tri_uv)
                    {
                        // texture coordinates: FillHit's UV frame and (u, v) (mesh.cpp:72-94)
                        const float4 ua = S.tri_uv[2 * (size_t)hr.x], ub = S.tri_uv[2 * (size_t)hr.x + 1];
                        h = fill_hit_uv(p1, p2, p3, n1, n2, n3, __int_as_float(hr.z), __int_as_float(hr.w), ua.x, ua.y,
                                        ua.z, ua.w, ub.x, ub.y, tx.u, tx.v);
                    }
                    else
                        h = fill_hit(p1, p2, p3, n1, n2, n3, __int_as_float(hr.z), __int_as_float(hr.w));
                    const int mrow = kLights ? __float_as_int(S.stri[(size_t)kShadeF4 * hr.x].w) : 0;
                    if (kLights && S.mat_maps) apply_maps(tx, mrow, m);
                    // NormalMapping (bsdf/normal_mapping.h:57-80): nA, nB = the hit normal mapped
                    // once and twice. The reference estimator keeps the unmapped normal for light
                    // sampling and uses nA / nB as Sample and Evaluate leave them (see the NEE loop);
                    // the quirk-free one maps it once for every use.
                    const int nmap = kLights && S.nmaps ? S.nmaps[mrow] : -1;
                    f3 nA = h.n, nB = h.n;
                    if (kLights && nmap >= 0)
                    {
                        nA = map_normal(tx, nmap, h, h.n);
                        if (nq) h.n = nA;
                        else nB = map_normal(tx, nmap, h, nA);
                    }
                    const bool nmref = kLights && nmap >= 0 && !nq;
                    f3 wo = -ld3(R.ray_d[i]);
                    uint32_t base = meta.x;
                    f3 T = ld3(R.thr[i]);
                    float ux, uy;
                    if (b == 0)
                    {
                        nee_sample2d(J.brdf_smp, base, KS_BRDF, 0, ux, uy);   // pre-drawn once per path, gitracer.cpp:36-41
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
                    else if (nq && m.type == 2)
                    {
                        fl |= F_TERM;   // quirk-free: emitters do not scatter; NEE counted their light
                    }
                    else
                    {
                        org = h.p;
                        // ---- next-event estimation (gitracer.cpp:70-84 -> ditracer.cpp:57-239) ----
                        if (nl > 0)
                        {
                            // every light at bounce 0, then one picked uniformly (gitracer.cpp:72-84);
                            // quirk-free: every light at bounce 0 while they fit the NEE slots, else one
                            // light picked in proportion to its power, weighted by 1 / its probability
                            const bool all = b == 0 && (!nq || nl <= kMaxNeeLights);
                            const int jn = all ? nl : 1;
                            int Lsel = 0;
                            float pickw = 1.f;
                            if (!all)
                            {
                                if (nq)
                                    Lsel = pick_light(S, kLights ? rs.next_float() : ks_rand_float(base, rand_pick_counter(b)),
                                                      pickw);
                                else
                                    Lsel = (int)((kLights ? rs.next_uint() : ks_rand(base, rand_pick_counter(b))) %
                                                 (uint32_t)nl);
                            }
                            for (int j = 0; j < jn; ++j)
                            {
                                int Lid = all ? j : Lsel;
                                const LightDev* L = S.lights + Lid;
                                uint32_t lc = b == 0 ? (uint32_t)(2 * j) : (uint32_t)(2 * nl + 2 * (b - 1));
                                uint32_t bc = 2u + lc;
                                float lx, ly, bx, by;
                                nee_sample2d(J.light_smp, base, KS_LIGHT, lc, lx, ly);
                                nee_sample2d(J.brdf_smp, base, KS_BRDF, bc, bx, by);
                                size_t ks = (size_t)i * K + 2 * j;
                                // light sample + shadow ray
                                f3 ld;
                                float lpdf = 0.f, cosl = 0.f;
                                // point, directional and environment lights exist only in kLights runs
                                const bool area = !kLights || L->type == LIGHT_AREA;
                                // point and directional lights are singular: light sample only, no
                                // MIS (ditracer.cpp:113-119,149)
                                const bool singular = !area && L->type <= LIGHT_DIRECTIONAL;
                                f3 le = area ? light_sample(L, h.p, lx, ly, ld, lpdf, &cosl)
                                             : other_light_sample(S, *L, h.p, h.n, lx, ly, ld, lpdf);
                                if ((nq ? lpdf > 0.f : lpdf > kMinPdf) && sqnorm(le) > 0.f)
                                {
                                    f3 wi = normalize(ld);
                                    float dist = sqrtf(sqnorm(ld));
                                    // a normal-mapped vertex: Evaluate maps the normal once (nA, also
                                    // for the cosine) and GetPdf maps that copy again (nB)
                                    Frame hE = h;
                                    if (nmref) hE.n = nA;
                                    f3 f = kLights ? mat_eval_x(S.mats, tx, m, hE, wi, wo) : mat_eval(m, h, wi, wo);
                                    float w = 1.f;
                                    if (!singular)
                                    {
                                        // the reference asks GetPdf(light dir, view), the density of the
                                        // view direction; MIS needs the density of sampling the light
                                        // direction from the view, GetPdf(view, light dir)
                                        Frame hP = h;
                                        if (nmref) hP.n = nB;
                                        float bp = kLights ? (nq ? mat_pdf_x(S.mats, tx, m, h, wo, wi) : mat_pdf_x(S.mats, tx, m, hP, wi, wo))
                                                           : (nq ? mat_pdf(m, h, wo, wi) : mat_pdf(m, h, wi, wo));
                                        // quirk-free: both MIS weights in solid angle (the reference
                                        // weighs an area light's area pdf against the BSDF's solid-angle
                                        // pdf; environment pdfs are in solid angle already)
                                        float lw = nq && area ? lpdf * sqnorm(ld) / cosl : lpdf;
                                        w = power_heuristic(lw, bp);
                                    }
                                    // GetLe / d^2 is the reference's whole area-to-solid-angle factor,
                                    // which treats the emitter as uniform radiance e; the emitter's
                                    // radiance is GetLe = e cos_l (what the camera and the BSDF branch
                                    // see), so the quirk-free estimator adds the Jacobian's cos_l
                                    if (nq && area) le = le * cosl;
                                    f3 c = le * f * fabsf(dot(hE.n, wi)) * w * (1.f / lpdf);
                                    if (nq) c = c * pickw;
                                    R.nee_a[ks] = st3(c, 0.f);
                                    R.nee_wi[ks] = st3(wi, dist - 0.01f);
                                    emit_mask |= 1ull << (2 * j);
                                }
                                // BSDF sample + MIS closest ray, sampled at frame hs into NEE slot `bit`
                                auto bsdf_nee = [&](const Frame& hs, int bit) {
                                    f3 wis = mk(0.f, 0.f, 0.f);
                                    float bpdf = 0.f;
                                    int type = 0;
                                    f3 fb = singular ? mk(0.f, 0.f, 0.f)
                                                     : kLights ? mat_sample_x(S.mats, tx, m, hs, bx, by, wo, wis, bpdf, type, rs)
                                                               : mat_sample(m, hs, bx, by, wo, wis, bpdf, type);
                                    f3 wi = normalize(wis);
                                    if (!singular && sqnorm(fb) > 0.f && (nq ? bpdf > 0.f : bpdf > kMinPdf))
                                    {
                                        // a specular BSDF sample takes weight 1 without asking the light's
                                        // pdf (ditracer.cpp:167-181)
                                        const bool spec = kLights && (type & BSDF_SPECULAR) != 0;
                                        float lp = spec ? 0.f : area ? light_pdf(L, hs.p, wi) : other_light_pdf(S, *L, hs.n, wi);
                                        // quirk-free: where an environment light's density is 0 the BSDF
                                        // sample carries the whole weight (an area light's density is 0
                                        // only where the ray cannot reach it)
                                        if (spec || (nq ? lp > 0.f || !area : !(lp < kMinPdf)))
                                        {
                                            float w = spec || (nq && !(lp > 0.f)) ? 1.f : power_heuristic(bpdf, lp);
                                            const size_t kb = (size_t)i * K + bit;
                                            R.nee_a[kb] = st3(fb, fabsf(dot(hs.n, wi)));
                                            R.nee_b[kb] = make_float4(nq ? w * pickw : w, 1.f / bpdf, (float)Lid, 0.f);
                                            R.nee_wi[kb] = st3(wi, 10000000.f);
                                            emit_mask |= 1ull << bit;
                                        }
                                    }
                                };
                                if (nmref)
                                {
                                    // Sample maps the normal as the light branch left it: twice where
                                    // the light sample was unoccluded (Evaluate ran), else once. The
                                    // shadow ray decides later, so both rays are traced.
                                    Frame hs = h;
                                    if ((emit_mask >> (2 * j)) & 1ull)
                                    {
                                        hs.n = nB;
                                        bsdf_nee(hs, 2 * j + 1);
                                        hs.n = nA;
                                        bsdf_nee(hs, 2 * jn + 2 * j + 1);
                                    }
                                    else
                                    {
                                        hs.n = nA;
                                        bsdf_nee(hs, 2 * j + 1);
                                    }
                                }
                                else
                                    bsdf_nee(h, 2 * j + 1);
                            }
                            R.thr_nee[i] = st3(T, 0.f);
                            R.ray_o[i] = st3(h.p, 0.f);
                            R.nee_mask[i] = emit_mask;
                            if (emit_mask) fl |= F_NEE;
                            if (emit_mask && nmref) fl |= F_NMAP;
                            ncnt = jn;
                        }
                        // GiTracer's own Sample maps the hit's normal once (gitracer.cpp:97), and
                        // the throughput's cosine reads it
                        if (nmref) h.n = nA;
                        // ---- continuation: BSDF sample, throughput, Russian roulette ----
                        if (b + 1 < J.depth)
                        {
                            f3 wi;
                            float pdf = 0.f;
                            int type = 0;
                            f3 f;
                            bool stop;
                            if (nq)
                            {
                                // quirk-free continuation: a fresh sample per bounce and, for guided
                                // jobs, the one-sample MIS mixture alpha BSDF + (1 - alpha) guide; the
                                // integrand is always Evaluate(view, dir) with pdf GetPdf(view, dir)
                                const uint32_t c6 = 6u * (uint32_t)b;
                                float gx, gy;
                                ks_sample2d(base, KS_GUIDE, c6, gx, gy);
                                bool guided = (jflags & JOB_GUIDED) != 0 && (J.guide_bounces == 0 || b < J.guide_bounces);
                                Lobes G;
                                const int grow = guided ? R.grow[i] : -1;
                                guided = grow >= 0;
                                if (guided)
                                {
                                    G = guide_lobes(R.gout + (size_t)grow * kGUsed);
                                    // a vertex whose network output is not finite samples the BSDF
                                    // alone; the choice depends on the vertex only, so the estimator
                                    // stays unbiased
                                    guided = lobes_finite(G);
                                }
                                // materials with delta lobes (perfect reflection and refraction, glass,
                                // mixtures) continue with their own sample and its weight, unguided
                                const bool own = kLights && m.type >= MAT_REFLECT;
                                guided = guided && !own;
                                float alpha = guided ? J.bsdf_frac : 1.f;
                                float usel = ks_float(ks_u32(base, KS_GUIDE, c6 + 2));
                                f3 fown = mk(0.f, 0.f, 0.f);
                                if (own)
                                {
                                    fown = mat_sample_x(S.mats, tx, m, h, gx, gy, wo, wi, pdf, type, rs);
                                    wi = normalize(wi);
                                }
                                else if (!guided || usel < alpha)
                                {
                                    if (kLights) (void)mat_sample_x(S.mats, tx, m, h, gx, gy, wo, wi, pdf, type, rs);
                                    else (void)mat_sample(m, h, gx, gy, wo, wi, pdf, type);
                                    wi = normalize(wi);
                                }
                                else
                                {
                                    float u0 = ks_float(ks_u32(base, KS_GUIDE, c6 + 3));
                                    float u1 = ks_float(ks_u32(base, KS_GUIDE, c6 + 4));
                                    float u2 = ks_float(ks_u32(base, KS_GUIDE, c6 + 5));
                                    wi = guide_sample(G, R.gout + (size_t)grow * kGUsed, u0, u1, u2);
                                }
                                if (own)
                                    f = fown;
                                else
                                {
                                    f = kLights ? mat_eval_x(S.mats, tx, m, h, wo, wi) : mat_eval(m, h, wo, wi);
                                    float pb = kLights ? mat_pdf_x(S.mats, tx, m, h, wo, wi) : mat_pdf(m, h, wo, wi);
                                    pdf = alpha * pb + (guided ? (1.f - alpha) * guide_pdf(G, wi) : 0.f);
                                }
                                stop = !(pdf > 0.f) || sqnorm(f) == 0.f;
                            }
                            else
                            {
                                f = kLights ? mat_sample_x(S.mats, tx, m, h, ux, uy, wo, wi, pdf, type, rs)
                                            : mat_sample(m, h, ux, uy, wo, wi, pdf, type);
                                stop = sqnorm(f) == 0.f || pdf < kMinPdf;
                            }
                            if (stop)
                            {
                                fl |= F_TERM;
                            }
                            else
                            {
                                f3 nd = normalize(wi);
                                T = T * (f * (fabsf(dot(h.n, wi)) / pdf));
                                if ((jflags & JOB_RECORD) && b + roff < kRecVerts)
                                {
                                    // training record: network input, sampled direction, mixture pdf,
                                    // throughput after this vertex
                                    f3 nn = dot(h.n, wo) < 0.f ? -h.n : h.n;
                                    float4* rc = R.rec + (size_t)i * kRecStride + (b + roff) * 5;
                                    rc[0] = make_float4((h.p.x - J.glo[0]) * J.ginv[0], (h.p.y - J.glo[1]) * J.ginv[1],
                                                        (h.p.z - J.glo[2]) * J.ginv[2], m.type == 1 ? m.alpha : 1.f);
                                    rc[1] = make_float4(nn.x, nn.y, nn.z, 1.f);
                                    rc[2] = make_float4(nd.x, nd.y, nd.z, pdf);
                                    rc[3] = st3(T, 0.f);
                                    rc[4] = make_float4(0.f, 0.f, 0.f, 0.f);
                                }
                                bool alive = true;
                                if (b > 3)
                                {
                                    float rnd = kLights ? rs.next_float() : ks_rand_float(base, rand_pick_counter(b) + 1u);
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
                if (jflags & JOB_RECORD)
                {
#pragma unroll
                    for (int k = 0; k < kRecVerts; ++k)
                    {
                        const float4* rc = R.rec + (size_t)i * kRecStride + k * 5;
                        float4 d = rc[2], snap = rc[4];
                        if (!(d.w > 0.f) || !(snap.w > 0.f)) continue;
                        float4 T = rc[3];
                        f3 L = mk(T.x > 0.f ? (rad.x - snap.x) / T.x : 0.f, T.y > 0.f ? (rad.y - snap.y) / T.y : 0.f,
                                  T.z > 0.f ? (rad.z - snap.z) / T.z : 0.f);
                        float lum = 0.2126f * L.x + 0.7152f * L.y + 0.0722f * L.z;
                        if (lum > 0.f)
                        {
                            // target density of the guide: incident radiance times the cosine
                            // (the whole integrand of a diffuse vertex; the network sees no view
                            // direction, so it cannot learn a glossy lobe)
                            float4 nn = rc[1];
                            float cs = nn.w > 0.f ? fabsf(nn.x * d.x + nn.y * d.y + nn.z * d.z) : 1.f;
                            // the record's weight goes in the free .w of its throughput; the
                            // chunk's compaction pass emits the records with a weight > 0
                            reinterpret_cast<float*>(R.rec + (size_t)i * kRecStride + k * 5 + 3)[3] = lum * cs / d.w;
                        }
                    }
                }
                if (!(fl & F_NOCLAMP) && !nq)
                {
                    f3 v = rad * (1.f / 1.f);
                    o = mk(clampf(v.x, 0.f, kMaxRadiance), clampf(v.y, 0.f, kMaxRadiance),
                           clampf(v.z, 0.f, kMaxRadiance));
                }
                R.out[i] = st3(o, 0.f);
                fl |= F_DONE;
            }
            R.rad[i] = st3(rad, kLights ? __uint_as_float(rs.c) : 0.f);
            meta.y = fl | ((uint32_t)b << 8) | ((uint32_t)ncnt << 16) | (job << 24);
            R.meta[i] = meta;
            keep = !(fl & F_DONE);
        }
    }
    // ---- queue append and the next bounce's live-path list: one atomic per block each ----
    // The trace kernel claims rays in queue order, so the wave writes them grouped: its
    // extension rays by direction octant, then its NEE rays slot by slot (the same light for
    // every lane of a run), instead of lane by lane. Rays claimed together then share an
    // origin region and a direction class. Hits are written per owner, so the order changes
    // no result. The block's four waves add their counts in LDS and one thread claims the
    // block's slots: every claim is a returning atomic on one of two hot cache lines, and
    // per wave they limited the kernel.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned long long mkeep = __ballot(keep);
    uint32_t ncl = (uint32_t)__popcll(__ballot(emit_ext)), nany = 0;
    for (int k = 0; k < K; ++k)
    {
        const uint32_t c = (uint32_t)__popcll(__ballot((emit_mask >> k) & 1ull));
        if (k & 1) ncl += c;
        else nany += c;
    }
    __shared__ uint32_t s_cnt[kBlock / 64][3];   // per wave: closest rays, any-hit rays, live paths
    __shared__ unsigned long long s_qbase;
    __shared__ uint32_t s_lbase;
    if (lane == 0)
    {
        s_cnt[wave][0] = ncl;
        s_cnt[wave][1] = nany;
        s_cnt[wave][2] = (uint32_t)__popcll(mkeep);
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t tc = 0, ta = 0, tl = 0;
        for (int w = 0; w < kBlock / 64; ++w)
        {
            tc += s_cnt[w][0];
            ta += s_cnt[w][1];
            tl += s_cnt[w][2];
        }
        // both queues with one 64-bit atomic: closest in the low word, any-hit in the high
        // word (neither count reaches 2^32, so no carry crosses); the run's ray totals are
        // summed from these counters once per chunk
        s_qbase = (tc | ta) ? atomicAdd(reinterpret_cast<unsigned long long*>(&R.counters[4 * Lout]),
                                        ((unsigned long long)ta << 32) | tc)
                            : 0ull;
        s_lbase = tl ? atomicAdd(&R.alive_n[b + 1], tl) : 0u;
    }
    __syncthreads();
    uint32_t cbase = (uint32_t)s_qbase, abase = (uint32_t)(s_qbase >> 32), lbase = s_lbase;
    for (int w = 0; w < wave; ++w)
    {
        cbase += s_cnt[w][0];
        abase += s_cnt[w][1];
        lbase += s_cnt[w][2];
    }
    if (keep) R.alive[(b + 1) & 1][lbase + (uint32_t)__popcll(mkeep & below)] = i;
    if (ncl | nany)
    {
        const int buf = Lout & 1;
        uint32_t cofs = 0, aofs = 0;
        const int oct = (ext_d.x < 0.f ? 1 : 0) | (ext_d.y < 0.f ? 2 : 0) | (ext_d.z < 0.f ? 4 : 0);
        for (int o = 0; o < 8; ++o)
        {
            const bool mine = emit_ext && oct == o;
            const unsigned long long m = __ballot(mine);
            if (mine)
            {
                const uint32_t cp = cbase + cofs + (uint32_t)__popcll(m & below);
                R.qc_o[buf][cp] = st3(org, 0.01f);
                R.qc_d[buf][cp] = st3(ext_d, 1000000.f);
                R.qc_own[buf][cp] = (int)((Q_EXT << 29) | i);
            }
            cofs += (uint32_t)__popcll(m);
        }
        for (int k = 0; k < K; ++k)
        {
            const bool mine = ((emit_mask >> k) & 1ull) != 0;
            const unsigned long long m = __ballot(mine);
            if (mine)
            {
                size_t ks = (size_t)i * K + k;
                float4 wd = R.nee_wi[ks];
                if (k & 1)
                {
                    const uint32_t cp = cbase + cofs + (uint32_t)__popcll(m & below);
                    R.qc_o[buf][cp] = st3(org, 0.01f);
                    R.qc_d[buf][cp] = wd;
                    R.qc_own[buf][cp] = (int)((Q_NEE_CLOSEST << 29) | (uint32_t)ks);
                }
                else
                {
                    const uint32_t ap = abase + aofs + (uint32_t)__popcll(m & below);
                    R.qa_o[buf][ap] = st3(org, 0.01f);
                    R.qa_d[buf][ap] = wd;
                    R.qa_own[buf][ap] = (int)((Q_NEE_ANY << 29) | (uint32_t)ks);
                }
            }
            if (k & 1) cofs += (uint32_t)__popcll(m);
            else aofs += (uint32_t)__popcll(m);
        }
    }
}

// ---- training records: each chunk's records are compacted in (path, vertex) order ----
// A record of path i, vertex k is emitted iff its direction was stored (pdf > 0) and its
// weight (throughput .w) is positive. Three passes, no atomics, so the order and the
// contents are the same run after run: per-block counts, one scan over the blocks (plus the
// running total of earlier chunks), and the writes.
__device__ __forceinline__ int rec_emitted(const RenderDev& R, uint32_t i, int k)
{
    const float4* rc = R.rec + (size_t)i * kRecStride + k * 5;
    return rc[2].w > 0.f && rc[3].w > 0.f;
}

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t& total)
{
    __shared__ uint32_t wsum[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = wave_incl_scan(v);
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t base = 0;
    total = 0;
    for (int w = 0; w < kBlock / 64; ++w)
    {
        if (w < wave) base += wsum[w];
        total += wsum[w];
    }
    __syncthreads();
    return base + incl - v;
}

__global__ __launch_bounds__(kBlock) void k_rec_count(RenderDev R, uint32_t* blk)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t n = 0;
    if (i < R.P)
        for (int k = 0; k < kRecVerts; ++k) n += rec_emitted(R, i, k);
    uint32_t total;
    (void)block_excl_scan(n, total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// One block: exclusive scan of the per-block counts, offset by the running total.
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void k_rec_scan(const uint32_t* blk, uint32_t nblk, uint32_t* running,
                                                         uint32_t* off)
{
    __shared__ uint32_t ws[kScanBlock / 64];
    const uint32_t per = (nblk + kScanBlock - 1) / kScanBlock;
    const uint32_t b0 = threadIdx.x * per, b1 = min(b0 + per, nblk);
    uint32_t sum = 0;
    for (uint32_t b = b0; b < b1; ++b) sum += blk[b];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = wave_incl_scan(sum);
    if (lane == 63) ws[wave] = incl;
    const uint32_t start = *running;
    __syncthreads();
    uint32_t base = start, total = 0;
    for (int w = 0; w < kScanBlock / 64; ++w)
    {
        if (w < wave) base += ws[w];
        total += ws[w];
    }
    uint32_t o = base + incl - sum;
    for (uint32_t b = b0; b < b1; ++b)
    {
        off[b] = o;
        o += blk[b];
    }
    __syncthreads();
    if (threadIdx.x == 0) *running = start + total;
}

__global__ __launch_bounds__(kBlock) void k_rec_write(RenderDev R, const uint32_t* off)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t mask = 0;
    if (i < R.P)
        for (int k = 0; k < kRecVerts; ++k) mask |= (uint32_t)rec_emitted(R, i, k) << k;
    uint32_t total;
    uint32_t slot = off[blockIdx.x] + block_excl_scan((uint32_t)__popc(mask), total);
    for (int k = 0; k < kRecVerts; ++k)
    {
        if (!(mask & (1u << k))) continue;
        if (slot < R.rec_cap)
        {
            const float4* rc = R.rec + (size_t)i * kRecStride + k * 5;
            float4* q = reinterpret_cast<float4*>(R.rec_out) + (size_t)slot * 3;
            float4 d = rc[2];
            q[0] = rc[0];
            q[1] = rc[1];
            q[2] = make_float4(d.x, d.y, d.z, rc[3].w);
        }
        ++slot;
    }
}

// One thread per pixel of a camera segment: FileImagePlane::WriteSample's buf += L,
// buf.w += 1 (fileimageplane.cpp:29-34) over the pixel's samples in order, plus sum L^2.
__global__ __launch_bounds__(kBlock) void k_accumulate(RenderDev R, uint32_t job, uint32_t u0, uint32_t nu,
                                                       uint32_t c0)
{
    const JobDev& J = R.jobs[job];
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= nu) return;
    uint32_t q = u0 + t;
    float4* acc = reinterpret_cast<float4*>(J.accum);
    float4 s1, s2;
    if (J.accumulate)
    {
        s1 = acc[2 * q];
        s2 = acc[2 * q + 1];
    }
    else
    {
        s1 = make_float4(0.f, 0.f, 0.f, 0.f);
        s2 = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const uint32_t first = c0 + t * (uint32_t)J.s_count;
    for (int k = 0; k < J.s_count; ++k)
    {
        float4 L = R.out[first + k];
        s1.x += L.x;
        s1.y += L.y;
        s1.z += L.z;
        s1.w += 1.f;
        s2.x += L.x * L.x;
        s2.y += L.y * L.y;
        s2.z += L.z * L.z;
        if (J.samples)
        {
            size_t o = ((size_t)q * J.s_count + k) * 3;
            J.samples[o] = L.x;
            J.samples[o + 1] = L.y;
            J.samples[o + 2] = L.z;
        }
    }
    acc[2 * q] = s1;
    acc[2 * q + 1] = s2;
}

// One wave per probe of a probe segment: lane l accumulates rays l, l+64, ... in order,
// then a fixed butterfly reduction; c_i = (4 pi / N) sum_j L_j Y_i(w_j).
template <int LMAX>
__global__ __launch_bounds__(kBlock) void k_project_sh(RenderDev R, uint32_t job, uint32_t u0, uint32_t nu,
                                                     uint32_t c0)
{
    constexpr int NT = (LMAX + 1) * (LMAX + 1);
    const JobDev& J = R.jobs[job];
    uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= nu) return;
    const uint32_t p = u0 + wave;
    const int N = J.N;
    float acc[3 * NT];
#pragma unroll
    for (int k = 0; k < 3 * NT; ++k) acc[k] = 0.f;
    for (int j = lane; j < N; j += 64)
    {
        const uint32_t i = c0 + wave * (uint32_t)N + (uint32_t)j;
        const float4 L = R.out[i];
        f3 d;
        float w = 1.f;
        if (J.flags & JOB_PROBE_DIR)
        {
            const float4 dw = R.pdir[i];
            d = mk(dw.x, dw.y, dw.z);
            w = dw.w;
        }
        else
        {
            d = probe_dir(ks_base(J.seed, p * (uint32_t)N + (uint32_t)j), j, J.g);
        }
        float y[kShMaxTerms];
        sh_evaluate(d, LMAX, y);
        const float lx = L.x * w, ly = L.y * w, lz = L.z * w;
#pragma unroll
        for (int k = 0; k < NT; ++k)
        {
            acc[3 * k + 0] += lx * y[k];
            acc[3 * k + 1] += ly * y[k];
            acc[3 * k + 2] += lz * y[k];
        }
        if (J.ray_dirs)
            reinterpret_cast<float4*>(J.ray_dirs)[(size_t)p * N + j] = make_float4(d.x, d.y, d.z, w);
        if (J.radiance)
        {
            size_t o = ((size_t)p * N + j) * 3;
            J.radiance[o] = L.x;
            J.radiance[o + 1] = L.y;
            J.radiance[o + 2] = L.z;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
        for (int k = 0; k < 3 * NT; ++k) acc[k] += __shfl_xor(acc[k], off);
    if (lane == 0)
    {
        float scale = 4.f * kPi / (float)N;
#pragma unroll
        for (int k = 0; k < 3 * NT; ++k) J.coeffs[(size_t)p * 3 * NT + k] = acc[k] * scale;
    }
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
static bsh_status ensure_work(RenderWork& w, size_t P, int K, int nlmax)
{
    BSH_HIP(w.ray_o.alloc(P));
    BSH_HIP(w.ray_d.alloc(P));
    BSH_HIP(w.thr.alloc(P));
    BSH_HIP(w.thr_nee.alloc(P));
    BSH_HIP(w.rad.alloc(P));
    BSH_HIP(w.out.alloc(P));
    BSH_HIP(w.hit.alloc(P));
    BSH_HIP(w.meta.alloc(P));
    BSH_HIP(w.nee_mask.alloc(P));
    BSH_HIP(w.alive[0].alloc(P));
    BSH_HIP(w.alive[1].alloc(P));
    size_t slots = P * (size_t)K;
    BSH_HIP(w.nee_a.alloc(slots));
    BSH_HIP(w.nee_b.alloc(slots));
    BSH_HIP(w.nee_wi.alloc(slots));
    BSH_HIP(w.nee_res.alloc(slots));
    size_t qc = P * (size_t)(1 + nlmax), qa = P * (size_t)std::max(nlmax, 1);
    for (int k = 0; k < 2; ++k)
    {
        BSH_HIP(w.qc_o[k].alloc(qc));
        BSH_HIP(w.qc_d[k].alloc(qc));
        BSH_HIP(w.qc_own[k].alloc(qc));
        BSH_HIP(w.qa_o[k].alloc(qa));
        BSH_HIP(w.qa_d[k].alloc(qa));
        BSH_HIP(w.qa_own[k].alloc(qa));
    }
    BSH_HIP(w.stats.alloc(kStatAll));
    return BSH_OK;
}

// HIP events around every trace and shade launch on the launch stream, read once at the end.
struct LaunchTimer
{
    std::vector<hipEvent_t> evs;
    std::vector<int> bounce;   // the bounce (launch index within its chunk) of every pass, in order
    size_t ei = 0;
    bsh_status init(size_t pairs)
    {
        evs.resize(5 * pairs + 2);
        for (auto& e : evs) BSH_HIP(hipEventCreate(&e));
        return BSH_OK;
    }
    hipEvent_t next() { return evs[ei++]; }
    ~LaunchTimer()
    {
        for (auto& e : evs) (void)hipEventDestroy(e);
    }
};

#define BSH_TRY(x)                   \
    do                               \
    {                                \
        bsh_status r_ = (x);         \
        if (r_ != BSH_OK) return r_; \
    } while (0)

// A run of consecutive units (pixels or probes) of one job inside a chunk.
struct Segment
{
    uint32_t job, u0, nu, c0;
};
struct Chunk
{
    uint32_t P = 0;
    int depth = 0;
    std::vector<Segment> segs;
};

static std::vector<Chunk> plan_chunks(const std::vector<JobDev>& jobs, size_t pmax)
{
    std::vector<Chunk> out;
    Chunk cur;
    for (uint32_t j = 0; j < jobs.size(); ++j)
    {
        const JobDev& J = jobs[j];
        uint32_t unit = J.type == 0 ? (uint32_t)J.s_count : (uint32_t)J.N;
        uint32_t nunits = J.type == 0 ? (uint32_t)(J.cw * J.ch) : (uint32_t)J.nprobes;
        uint32_t u = 0;
        while (u < nunits)
        {
            size_t room = pmax - cur.P;
            if (room < unit)
            {
                out.push_back(cur);
                cur = Chunk();
                room = pmax;
            }
            uint32_t take = (uint32_t)std::min<size_t>(nunits - u, room / unit);
            cur.segs.push_back(Segment{j, u, take, cur.P});
            cur.P += take * unit;
            cur.depth = std::max(cur.depth, J.depth);
            u += take;
        }
    }
    if (cur.P) out.push_back(cur);
    return out;
}

static dim3 grid_for(size_t n) { return dim3((uint32_t)std::max<size_t>(1, (n + kBlock - 1) / kBlock)); }

// k_trace_refill for one node layout, with the single-scene variant when one scene is traced
template <bool kCount, bool kQ>
static void launch_refill(bool one, int blocks, hipStream_t st, const RenderDev& R, int b, int refill, uint32_t claim)
{
    if (one)
        hipLaunchKernelGGL((k_trace_refill<kCount, kQ, true>), dim3(blocks), dim3(kBlock), 0, st, R, b, b & 1, refill,
                           claim);
    else
        hipLaunchKernelGGL((k_trace_refill<kCount, kQ, false>), dim3(blocks), dim3(kBlock), 0, st, R, b, b & 1, refill,
                           claim);
}

// Enqueues every kernel of a batch on `st`. With T == nullptr no events are recorded
// (the form captured into a hipGraph).
// Orders further streams after a first one (fork) and the first after all of them (join).
struct ForkJoin
{
    static constexpr int kMax = bsh_context::kLanes - 1;
    hipEvent_t f = nullptr, j[kMax] = {};
    bsh_status fork(hipStream_t a, const hipStream_t* b, int n)
    {
        BSH_HIP(hipEventCreateWithFlags(&f, hipEventDisableTiming));
        BSH_HIP(hipEventRecord(f, a));
        for (int k = 0; k < n; ++k)
        {
            BSH_HIP(hipEventCreateWithFlags(&j[k], hipEventDisableTiming));
            BSH_HIP(hipStreamWaitEvent(b[k], f, 0));
        }
        return BSH_OK;
    }
    bsh_status join(hipStream_t a, const hipStream_t* b, int n)
    {
        for (int k = 0; k < n; ++k)
        {
            BSH_HIP(hipEventRecord(j[k], b[k]));
            BSH_HIP(hipStreamWaitEvent(a, j[k], 0));
        }
        return BSH_OK;
    }
    ~ForkJoin()
    {
        if (f) (void)hipEventDestroy(f);
        for (hipEvent_t e : j)
            if (e) (void)hipEventDestroy(e);
    }
};

// One wavefront lane: its device view, its buffers and the stream its chunks run on.
struct Lane
{
    RenderDev R;
    RenderWork* w;
    hipStream_t st;
};

// Enqueues every chunk of the run; chunk k goes to lane k mod nlanes.
static bsh_status enqueue_batch(bsh_context* c, const Lane* lanes, int nlanes, const std::vector<Chunk>& chunks,
                                const std::vector<JobDev>& hjobs, bool count, const bsh_tuning& tune, LaunchTimer* T,
                                GuideRun* gr)
{
    const int trace_blocks = c->num_cus * std::max(1, tune.trace_blocks_per_cu);
    const int refill = tune.refill_lanes == 0 ? kDefaultRefill : tune.refill_lanes;
    const bool one = hjobs.size() == 1;   // every path of the run belongs to job 0
    // per-atomic trace claims: the tuning's, else exactly the idle lanes on one scene and
    // kClaimRays when the run's jobs trace several (R.nodes1 is set only for one scene)
    const uint32_t claim = tune.claim_rays > 0 ? (uint32_t)tune.claim_rays : lanes[0].R.nodes1 ? 0u : kClaimRays;
    for (size_t ci = 0; ci < chunks.size(); ++ci)
    {
        const Chunk& ch = chunks[ci];
        const Lane& ln = lanes[ci % (size_t)nlanes];
        RenderDev R = ln.R;
        RenderWork& w = *ln.w;
        hipStream_t st = ln.st;
        const int launches = ch.depth + 1;
        R.P = ch.P;
        bool full = false, lights = false, wide = refill > 0, qall = true, qany = false;
        for (const Segment& sg : ch.segs)
        {
            qall = qall && hjobs[sg.job].s.qnodes;
            qany = qany || hjobs[sg.job].s.qnodes;
            full = full || hjobs[sg.job].flags != 0;
            // point, directional and environment lights and the materials of rows 3-7 are compiled
            // into the kLights variants only
            lights = lights || hjobs[sg.job].s.narea != hjobs[sg.job].s.nlights || hjobs[sg.job].s.ext_mats;
            wide = wide && hjobs[sg.job].s.width == 4;
        }
        // the refill kernel is compiled for one 4-wide node layout; a chunk that mixes the float
        // and quantised layouts is traced by the queue kernel, which reads the layout per ray
        wide = wide && qall == qany;
        hipLaunchKernelGGL(k_zero, grid_for(w.counters.n), dim3(kBlock), 0, st, w.counters.p, (int)w.counters.n);
        // guided probe jobs: the network rows of their probes (vertex -1 of every probe path)
        // come first, one row per probe, so the probe rays can be drawn from the mixture
        std::vector<uint32_t> seg_row(ch.segs.size(), 0u);
        uint32_t prow = 0;
        for (size_t k = 0; k < ch.segs.size(); ++k)
        {
            const Segment& sg = ch.segs[k];
            if (hjobs[sg.job].type != 1 || !(hjobs[sg.job].flags & JOB_PROBE_DIR)) continue;
            seg_row[k] = prow;
            hipLaunchKernelGGL(k_probe_guide_inputs, grid_for(sg.nu), dim3(kBlock), 0, st, R, sg.job, sg.u0, sg.nu, prow);
            prow += sg.nu;
        }
        if (prow)
        {
            if (!gr || !gr->guide) return fail(BSH_ERR_STATE, "guided probe job without a guide");
            bsh_status gs = guide_forward_rows(gr->guide, R.gin, nullptr, prow, w.gout.p, st);
            if (gs != BSH_OK) return gs;
        }
        for (size_t k = 0; k < ch.segs.size(); ++k)
        {
            const Segment& sg = ch.segs[k];
            const JobDev& J = hjobs[sg.job];
            size_t n = (size_t)sg.nu * (J.type == 0 ? J.s_count : J.N);
            if (J.type == 0) hipLaunchKernelGGL(k_raygen, grid_for(n), dim3(kBlock), 0, st, R, sg.job, sg.u0, sg.nu, sg.c0);
            else
                hipLaunchKernelGGL(k_raygen_probe, grid_for(n), dim3(kBlock), 0, st, R, sg.job, sg.u0, sg.nu, sg.c0,
                                   seg_row[k]);
        }
        hipLaunchKernelGGL(k_set_count, dim3(1), dim3(64), 0, st, w.counters.p, R.stats, ch.P);
        dim3 pg = grid_for(ch.P);
        for (int b = 0; b <= ch.depth; ++b)
        {
            if (T)
            {
                BSH_HIP(hipEventRecord(T->next(), st));
                T->bounce.push_back(b);
            }
            if (wide)
            {
                const bool one = R.nodes1 != nullptr;
                if (count)
                {
                    if (qall) launch_refill<true, true>(one, trace_blocks, st, R, b, refill, claim);
                    else launch_refill<true, false>(one, trace_blocks, st, R, b, refill, claim);
                }
                else if (qall)
                    launch_refill<false, true>(one, trace_blocks, st, R, b, refill, claim);
                else
                    launch_refill<false, false>(one, trace_blocks, st, R, b, refill, claim);
            }
            else if (count)
                hipLaunchKernelGGL(k_trace_queue<true>, dim3(trace_blocks), dim3(kBlock), 0, st, R, b, b & 1);
            else
                hipLaunchKernelGGL(k_trace_queue<false>, dim3(trace_blocks), dim3(kBlock), 0, st, R, b, b & 1);
            if (T) BSH_HIP(hipEventRecord(T->next(), st));
            // no inference at the chunk's last bounce: no job of it samples a continuation there
            const bool infer = gr && gr->guide && b + 1 < ch.depth && (gr->bounces == 0 || b < gr->bounces);
            if (infer)
            {
                hipLaunchKernelGGL(k_guide_inputs, pg, dim3(kBlock), 0, st, R, b);
            }
            if (T) BSH_HIP(hipEventRecord(T->next(), st));
            if (infer)
            {
                bsh_status gs = guide_forward_rows(gr->guide, R.gin, w.gseg.p, pg.x * kBlock, w.gout.p, st);
                if (gs != BSH_OK) return gs;
            }
            if (T) BSH_HIP(hipEventRecord(T->next(), st));
            const int lo = b + 1 <= ch.depth ? b + 1 : b;
            if (lights && one) hipLaunchKernelGGL((k_shade<true, true, true>), pg, dim3(kBlock), 0, st, R, b, lo);
            else if (lights) hipLaunchKernelGGL((k_shade<true, false, true>), pg, dim3(kBlock), 0, st, R, b, lo);
            else if (full && one) hipLaunchKernelGGL((k_shade<true, true>), pg, dim3(kBlock), 0, st, R, b, lo);
            else if (full) hipLaunchKernelGGL((k_shade<true, false>), pg, dim3(kBlock), 0, st, R, b, lo);
            else if (one) hipLaunchKernelGGL((k_shade<false, true>), pg, dim3(kBlock), 0, st, R, b, lo);
            else hipLaunchKernelGGL((k_shade<false, false>), pg, dim3(kBlock), 0, st, R, b, lo);
            if (T) BSH_HIP(hipEventRecord(T->next(), st));
        }
        hipLaunchKernelGGL(k_chunk_rays, dim3(1), dim3(64), 0, st, (const uint32_t*)w.counters.p, launches, R.stats);
        if (R.rec_out)
        {
            hipLaunchKernelGGL(k_rec_count, pg, dim3(kBlock), 0, st, R, w.rec_blk.p);
            hipLaunchKernelGGL(k_rec_scan, dim3(1), dim3(kScanBlock), 0, st, w.rec_blk.p, pg.x, w.rec_count.p,
                               w.rec_off.p);
            hipLaunchKernelGGL(k_rec_write, pg, dim3(kBlock), 0, st, R, w.rec_off.p);
        }
        for (const Segment& sg : ch.segs)
        {
            const JobDev& J = hjobs[sg.job];
            if (J.type == 0)
            {
                hipLaunchKernelGGL(k_accumulate, grid_for(sg.nu), dim3(kBlock), 0, st, R, sg.job, sg.u0, sg.nu, sg.c0);
            }
            else
            {
                dim3 wg = grid_for((size_t)sg.nu * 64);
                switch (J.lmax)
                {
                case 0: hipLaunchKernelGGL(k_project_sh<0>, wg, dim3(kBlock), 0, st, R, sg.job, sg.u0, sg.nu, sg.c0); break;
                case 1: hipLaunchKernelGGL(k_project_sh<1>, wg, dim3(kBlock), 0, st, R, sg.job, sg.u0, sg.nu, sg.c0); break;
                case 2: hipLaunchKernelGGL(k_project_sh<2>, wg, dim3(kBlock), 0, st, R, sg.job, sg.u0, sg.nu, sg.c0); break;
                case 3: hipLaunchKernelGGL(k_project_sh<3>, wg, dim3(kBlock), 0, st, R, sg.job, sg.u0, sg.nu, sg.c0); break;
                default: hipLaunchKernelGGL(k_project_sh<4>, wg, dim3(kBlock), 0, st, R, sg.job, sg.u0, sg.nu, sg.c0); break;
                }
            }
        }
    }
    BSH_HIP(hipGetLastError());
    return BSH_OK;
}


// The guide's position input is the hit point mapped to [0,1]^3 over the scene's box.
static void set_guide_box(const bsh_scene* s, JobDev& J)
{
    for (int k = 0; k < 3; ++k)
    {
        J.glo[k] = s->box_lo[k];
        float e = s->box_hi[k] - s->box_lo[k];
        J.ginv[k] = e > 0.f ? 1.f / e : 0.f;
    }
}

bsh_status make_render_job(bsh_scene* s, const bsh_render_params* p, float* accum, float* samples, JobDev& J)
{
    if (!s || !p || !accum) return fail(BSH_ERR_INVALID, "bsh_render: NULL argument");
    if (!s->committed) return fail(BSH_ERR_STATE, "bsh_render: scene not committed");
    if (!s->has_cam) return fail(BSH_ERR_STATE, "bsh_render: no camera");
    if (p->width <= 0 || p->height <= 0 || p->img_grid <= 0 || p->max_depth <= 0 || p->max_depth > 250)
        return fail(BSH_ERR_INVALID, "bsh_render: bad frame parameters");
    if (p->chunk_spp != 0)
        return fail(BSH_ERR_INVALID, "bsh_render: chunk_spp is reserved and must be 0 (bound the paths per "
                                     "chunk with bsh_tuning.paths_per_chunk)");
    if (!s->nmaps.empty() && s->lights.size() + s->others.size() > (size_t)kMaxNeeLights / 2 &&
        !(p->flags & BSH_RENDER_NO_QUIRKS))
        return fail(BSH_ERR_UNSUPPORTED, "bsh_render: normal maps with more than " + std::to_string(kMaxNeeLights / 2) +
                                             " lights need BSH_RENDER_NO_QUIRKS (the bounce-0 loop traces two BSDF "
                                             "rays per light)");
    if (s->lights.size() + s->others.size() > (size_t)kMaxNeeLights && !(p->flags & BSH_RENDER_NO_QUIRKS))
        return fail(BSH_ERR_UNSUPPORTED, "bsh_render: GiTracer's bounce-0 loop over every light supports at most 32 "
                                         "lights (BSH_RENDER_NO_QUIRKS picks one light per vertex by power)");
    int cw = p->crop_w > 0 ? p->crop_w : p->width, ch = p->crop_h > 0 ? p->crop_h : p->height;
    int cx = p->crop_w > 0 ? p->crop_x : 0, cy = p->crop_h > 0 ? p->crop_y : 0;
    if (cx < 0 || cy < 0 || cx + cw > p->width || cy + ch > p->height)
        return fail(BSH_ERR_INVALID, "bsh_render: crop outside the image");
    if (p->img_sampler < SMP_STRATIFIED || p->img_sampler > SMP_SOBOL)
        return fail(BSH_ERR_INVALID, "bsh_render: unknown image sampler kind");
    for (int k : {p->light_sampler, p->brdf_sampler})
    {
        if (k < SMP_STRATIFIED || k > SMP_SOBOL) return fail(BSH_ERR_INVALID, "bsh_render: unknown light/BSDF sampler kind");
        // a Sobol sequence runs over every draw a pixel makes, in the order its paths make
        // them (sobol_sampler.cpp:32-39); wavefront paths have no such order
        if (k == SMP_SOBOL)
            return fail(BSH_ERR_UNSUPPORTED, "bsh_render: a Sobol light or BSDF sampler is not supported (its sequence "
                                             "index runs over all of a pixel's draws in path order)");
    }
    if (p->img_grid > ((p->img_sampler == SMP_RANDOM || p->img_sampler == SMP_SOBOL) ? (1 << 24) : 4096))
        return fail(BSH_ERR_INVALID, "bsh_render: img_grid too large");
    int spp = sampler_num_samples(p->img_sampler, p->img_grid);
    int sb = p->spp_begin, sc = p->spp_count > 0 ? p->spp_count : spp - p->spp_begin;
    if (sb < 0 || sc <= 0 || sb + sc > spp) return fail(BSH_ERR_INVALID, "bsh_render: bad sample range");
    if ((uint64_t)p->width * p->height * spp > 0xffffffffull)
        return fail(BSH_ERR_UNSUPPORTED, "bsh_render: more than 2^32 camera samples per frame");
    J = JobDev{};
    J.s = s->dev();
    J.cam = make_camera(s->cam);
    J.type = 0;
    J.W = p->width;
    J.H = p->height;
    J.cx = cx;
    J.cy = cy;
    J.cw = cw;
    J.ch = ch;
    J.grid = p->img_grid;
    J.img_smp = p->img_sampler;
    J.light_smp = p->light_sampler;
    J.brdf_smp = p->brdf_sampler;
    J.spp = spp;
    J.s_begin = sb;
    J.s_count = sc;
    J.depth = p->max_depth;
    J.seed = p->seed;
    for (int k = 0; k < 3; ++k) J.bg[k] = p->bg[k];
    J.accumulate = (p->flags & BSH_RENDER_ACCUMULATE) ? 1 : 0;
    J.accum = accum;
    J.samples = samples;
    J.flags = (p->flags & BSH_RENDER_NO_QUIRKS) ? JOB_NO_QUIRKS : 0;
    set_guide_box(s, J);
    return BSH_OK;
}

bsh_status make_probe_job(bsh_scene* s, const float* probes, size_t nprobes, const bsh_probe_params* p,
                          float* coeffs, float* radiance, JobDev& J)
{
    if (!s || !p || !coeffs || (nprobes && !probes)) return fail(BSH_ERR_INVALID, "bsh_bake_sh: NULL argument");
    if (!s->committed) return fail(BSH_ERR_STATE, "bsh_bake_sh: scene not committed");
    if (p->lmax < 0 || p->lmax > kShMaxL) return fail(BSH_ERR_UNSUPPORTED, "bsh_bake_sh: lmax must be 0..4");
    int N = p->rays_per_probe;
    int g = (int)std::lround(std::sqrt((double)(N > 0 ? N : 0)));
    if (N <= 0 || g * g != N) return fail(BSH_ERR_INVALID, "bsh_bake_sh: rays_per_probe must be a square");
    if (p->max_depth <= 0 || p->max_depth > 250) return fail(BSH_ERR_INVALID, "bsh_bake_sh: bad max_depth");
    if (p->chunk_probes != 0)
        return fail(BSH_ERR_INVALID, "bsh_bake_sh: chunk_probes is reserved and must be 0 (bound the paths per "
                                     "chunk with bsh_tuning.paths_per_chunk)");
    if ((uint64_t)nprobes * (uint64_t)N > 0xffffffffull)
        return fail(BSH_ERR_UNSUPPORTED, "bsh_bake_sh: more than 2^32 probe rays");
    if (!s->nmaps.empty() && s->lights.size() + s->others.size() > (size_t)kMaxNeeLights / 2 &&
        !(p->flags & BSH_RENDER_NO_QUIRKS))
        return fail(BSH_ERR_UNSUPPORTED, "bsh_bake_sh: normal maps with more than " + std::to_string(kMaxNeeLights / 2) +
                                             " lights need BSH_RENDER_NO_QUIRKS");
    if (s->lights.size() + s->others.size() > (size_t)kMaxNeeLights && !(p->flags & BSH_RENDER_NO_QUIRKS))
        return fail(BSH_ERR_UNSUPPORTED, "bsh_bake_sh: GiTracer's bounce-0 loop over every light supports at most 32 "
                                         "lights (BSH_RENDER_NO_QUIRKS picks one light per vertex by power)");
    J = JobDev{};
    J.s = s->dev();
    J.type = 1;
    J.depth = p->max_depth;
    J.seed = p->seed;
    for (int k = 0; k < 3; ++k) J.bg[k] = p->bg[k];
    J.probes = probes;
    J.nprobes = (int)nprobes;
    J.N = N;
    J.g = g;
    J.lmax = p->lmax;
    J.tmin = p->tmin;
    J.coeffs = coeffs;
    J.radiance = radiance;
    J.ray_dirs = nullptr;
    J.flags = (p->flags & BSH_RENDER_NO_QUIRKS) ? JOB_NO_QUIRKS : 0;
    set_guide_box(s, J);
    return BSH_OK;
}

bsh_status run_jobs(bsh_context* c, std::vector<JobDev>& hjobs, bool count, const bsh_tuning* tune_in, int use_graph,
                    hipStream_t st, GuideRun* gr)
{
    if (hjobs.empty()) return BSH_OK;
    if (hjobs.size() > 256) return fail(BSH_ERR_UNSUPPORTED, "at most 256 jobs per batch");
    std::lock_guard<std::mutex> lock(c->mu);
    BSH_HIP(hipSetDevice(c->device));
    bsh_tuning tune = tune_in ? *tune_in : c->tuning;
    BSH_TRY(validate_tuning(tune, "run"));
    if (tune.paths_per_chunk <= 0) tune.paths_per_chunk = 1 << 23;
    int nlmax = 1;
    uint32_t maxunit = 1;
    uint64_t samples = 0;
    bool nmaps = false;
    for (const JobDev& J : hjobs)
    {
        // NEE slots: two per light of the bounce-0 loop, two for a picked light (four with
        // normal maps: the second BSDF ray of a normal-mapped vertex)
        nlmax = std::max(nlmax, (J.flags & JOB_NO_QUIRKS) && J.s.nlights > kMaxNeeLights ? 1 : J.s.nlights);
        nmaps = nmaps || (J.s.nmaps && !(J.flags & JOB_NO_QUIRKS));
        maxunit = std::max<uint32_t>(maxunit, J.type == 0 ? (uint32_t)J.s_count : (uint32_t)J.N);
        samples += J.type == 0 ? (uint64_t)J.cw * J.ch * J.s_count : (uint64_t)J.nprobes * J.N;
    }
    int Kshift = 1;
    while ((1 << Kshift) < (nmaps ? 4 : 2) * nlmax) ++Kshift;
    if (Kshift > 6) return fail(BSH_ERR_UNSUPPORTED, "more NEE slots per path than the 64-bit slot mask holds");
    const int K = 1 << Kshift;
    size_t pmax = std::max<size_t>((size_t)tune.paths_per_chunk, maxunit);
    if (pmax * (size_t)K >= ((size_t)1 << 29)) pmax = (((size_t)1 << 29) - 1) / K;
    // the NEE slot buffers (64 B per slot) hold pmax x K slots: at most 2^27 (8 GiB) per lane, so a
    // scene with many lights runs fewer paths per chunk
    if (pmax * (size_t)K > ((size_t)1 << 27)) pmax = std::max<size_t>(((size_t)1 << 27) / K, maxunit);
    if (pmax < maxunit) return fail(BSH_ERR_UNSUPPORTED, "unit larger than the NEE slot space");
    std::vector<Chunk> chunks = plan_chunks(hjobs, pmax);
    // An unguided run that fits one chunk of at least kSplitPaths paths is planned as two, so
    // the two lanes overlap (4096 probes x 1024 rays, depth 6: 1428 -> 1534 Mrays/s; a
    // 1024^2 4-spp frame unchanged; guided passes lost 2 %, and 3 % when re-measured on the
    // quantised trace kernel, and were unchanged with 3 trace workgroups per lane, so they keep
    // one chunk).
    if (chunks.size() == 1 && chunks[0].P >= kSplitPaths && !(gr && gr->guide))
    {
        // whole units per chunk: a half that splits a unit would leave a third chunk, so try
        // a few sizes up from the half
        size_t h = std::max<size_t>((chunks[0].P + 1) / 2, maxunit);
        for (int k = 0; k < 4 && h < chunks[0].P; ++k, h += maxunit)
        {
            std::vector<Chunk> two = plan_chunks(hjobs, h);
            if (two.size() == 2)
            {
                chunks.swap(two);
                break;
            }
        }
    }
    size_t P = 0;
    int maxdepth = 1;
    size_t pairs = 0;
    for (const Chunk& ch : chunks)
    {
        P = std::max<size_t>(P, ch.P);
        maxdepth = std::max(maxdepth, ch.depth);
        pairs += ch.depth + 1;
    }
    const bool guided = gr && gr->guide;
    const bool record = gr && gr->records && gr->cap > 0;
    // Two lanes when the run has chunks to alternate and nothing ties the chunks to one order
    // (training records take a running offset).
    const int nlanes = record || tune.lanes == 1 ? 1 : (int)std::min<size_t>(bsh_context::kLanes, chunks.size());
    // persistent trace workgroups per CU: all of them on one lane; on two lanes each lane's
    // grid leaves room for the other lane's, which the trace kernel's register bound keeps
    // resident beside it (1080p frame, 16 chunks: 3 per CU 2515 Mrays/s, 4 2483, 5 2468, 2 2415;
    // 4096-probe bake, two chunks side by side: 3 per CU 2302-2337, 4 2244-2250, 5 2232, 2
    // 2084-2091)
    if (tune.trace_blocks_per_cu <= 0) tune.trace_blocks_per_cu = nlanes > 1 ? kTraceBlocksPerLane : kMaxTraceBlocksPerCU;
    for (int k = 0; k + 1 < nlanes; ++k)
        if (!c->lane_stream[k]) BSH_HIP(hipStreamCreateWithFlags(&c->lane_stream[k], hipStreamNonBlocking));
    bool probe_dir = false;
    for (const JobDev& J : hjobs) probe_dir = probe_dir || (J.flags & JOB_PROBE_DIR) != 0;
    // 4 queue counters per launch, then, from the next 256-B boundary (a cache line of their
    // own), the live-path counts of bounces 0 .. maxdepth + 1; every chunk clears the buffer
    const size_t alive_off = ((size_t)4 * (maxdepth + 1) + 63) / 64 * 64;
    const size_t trace_threads = (size_t)c->num_cus * std::max(1, tune.trace_blocks_per_cu) * kBlock;
    BSH_HIP(c->jobs.alloc(hjobs.size()));
    BSH_HIP(hipMemcpyAsync(c->jobs.p, hjobs.data(), hjobs.size() * sizeof(JobDev), hipMemcpyHostToDevice, st));
    BSH_TRY(ensure_work(c->work, P, K, nlmax));   // lane 0's buffers also hold the run's statistics
    RenderWork& w = c->work;

    // every lane's buffers and device view (the statistics are lane 0's, shared)
    auto setup_lane = [&](RenderWork& lw, RenderDev& R) -> bsh_status {
        BSH_TRY(ensure_work(lw, P, K, nlmax));
        BSH_HIP(lw.counters.alloc(alive_off + maxdepth + 2));
        if (guided)
        {
            const size_t Pb = (P + kBlock - 1) / kBlock * kBlock;   // rows live in whole segments
            BSH_HIP(lw.gin.alloc(2 * Pb));
            BSH_HIP(lw.gout.alloc((size_t)kGUsed * Pb));
            BSH_HIP(lw.grow.alloc(P));
            BSH_HIP(lw.gseg.alloc(Pb / kBlock));
        }
        if (probe_dir) BSH_HIP(lw.pdir.alloc(P));
        if (record)
        {
            BSH_HIP(lw.rec.alloc((size_t)kRecStride * P));
            BSH_HIP(lw.rec_count.alloc(1));
            BSH_HIP(lw.rec_blk.alloc((P + kBlock - 1) / kBlock));
            BSH_HIP(lw.rec_off.alloc((P + kBlock - 1) / kBlock));
        }
        BSH_HIP(lw.stk_ovf.alloc(trace_threads * (std::max(kMaxStack, kMaxStack4) - kTraceLds)));
        R.jobs = c->jobs.p;
        R.Kshift = Kshift;
        R.ray_o = lw.ray_o.p; R.ray_d = lw.ray_d.p; R.thr = lw.thr.p; R.thr_nee = lw.thr_nee.p; R.rad = lw.rad.p;
        R.out = lw.out.p;
        R.hit = lw.hit.p; R.meta = lw.meta.p; R.nee_mask = lw.nee_mask.p;
        R.nee_a = lw.nee_a.p; R.nee_b = lw.nee_b.p; R.nee_wi = lw.nee_wi.p; R.nee_res = lw.nee_res.p;
        for (int k = 0; k < 2; ++k)
        {
            R.qc_o[k] = lw.qc_o[k].p; R.qc_d[k] = lw.qc_d[k].p; R.qc_own[k] = lw.qc_own[k].p;
            R.qa_o[k] = lw.qa_o[k].p; R.qa_d[k] = lw.qa_d[k].p; R.qa_own[k] = lw.qa_own[k].p;
        }
        R.counters = lw.counters.p;
        R.alive_n = lw.counters.p + alive_off;
        R.alive[0] = lw.alive[0].p;
        R.alive[1] = lw.alive[1].p;
        R.stats = w.stats.p;
        R.stk_ovf = lw.stk_ovf.p;
        R.P = 0;
        R.gin = guided ? lw.gin.p : nullptr;
        R.gout = guided ? lw.gout.p : nullptr;
        R.grow = guided ? lw.grow.p : nullptr;
        R.gseg = guided ? lw.gseg.p : nullptr;
        R.pdir = probe_dir ? lw.pdir.p : nullptr;
        R.nodes1 = hjobs.empty() ? nullptr : hjobs[0].s.nodes;
        R.tris1 = hjobs.empty() ? nullptr : hjobs[0].s.tris;
        for (const JobDev& J : hjobs)
            if (J.s.nodes != R.nodes1 || J.s.tris != R.tris1) R.nodes1 = R.tris1 = nullptr;
        R.rec = record ? lw.rec.p : nullptr;
        R.rec_out = record ? gr->records : nullptr;
        R.rec_cap = record ? (uint32_t)std::min<size_t>(gr->cap, 0xffffffffu) : 0u;
        R.rec_count = record ? lw.rec_count.p : nullptr;
        return BSH_OK;
    };
    Lane lanes[bsh_context::kLanes];
    std::memset(lanes, 0, sizeof lanes);   // zero padding: the device views are compared bytewise below
    lanes[0].w = &c->work;
    lanes[0].st = st;
    BSH_TRY(setup_lane(c->work, lanes[0].R));
    for (int k = 1; k < nlanes; ++k)
    {
        lanes[k].w = &c->lane_work[k - 1];
        lanes[k].st = c->lane_stream[k - 1];
        BSH_TRY(setup_lane(c->lane_work[k - 1], lanes[k].R));
    }
    if (record) hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, w.rec_count.p, 1);

    hipLaunchKernelGGL(k_zero, grid_for(2 * kStatAll), dim3(kBlock), 0, st, (uint32_t*)w.stats.p, 2 * kStatAll);
    bsh_render_stats& S = c->stats;
    S = bsh_render_stats{};
    std::vector<double> launch_ms(maxdepth + 1, 0.0);
    if (use_graph)
    {
        // Capture the whole batch (every chunk's raygen, bounces, accumulation) as one graph
        // and launch it; counters are cleared by kernels, never by memset nodes. The launches
        // bake in the lanes' device views (buffer pointers), the chunk plan and the kernel
        // variants, which follow from the key below; the job table itself is uploaded before
        // every launch. So an unguided run whose key matches the context's last graph replays
        // it without capture or instantiation.
        std::vector<uint8_t> key;
        auto put = [&key](const void* p, size_t n) {
            const uint8_t* b = static_cast<const uint8_t*>(p);
            key.insert(key.end(), b, b + n);
        };
        const int head[3] = {count ? 1 : 0, nlanes, (int)hjobs.size()};
        put(head, sizeof head);
        put(&tune, sizeof tune);
        for (const JobDev& J : hjobs)
        {
            const int f[13] = {J.type, J.s_count, J.N, J.cw, J.ch, J.nprobes, J.depth, J.flags,
                               J.s.narea, J.s.nlights, J.s.ext_mats, J.s.width, J.s.qnodes};
            put(f, sizeof f);
        }
        for (int k = 0; k < nlanes; ++k) put(&lanes[k].R, sizeof(RenderDev));
        const bool cacheable = !gr;
        bool replay = cacheable && c->graph_exec && c->graph_key == key;
        if (!c->graph_stream) BSH_HIP(hipStreamCreateWithFlags(&c->graph_stream, hipStreamNonBlocking));
        hipStream_t cs = c->graph_stream;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        BSH_HIP(hipEventCreate(&e0));
        BSH_HIP(hipEventCreate(&e1));
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = replay ? c->graph_exec : nullptr;
        bsh_status rc = BSH_OK;
        // order the graph stream after prior work on `st` (job table upload, stats clear)
        (void)hipEventRecord(e0, st);
        (void)hipStreamWaitEvent(cs, e0, 0);
        if (!replay)
        {
            c->drop_graph();
            (void)hipStreamSynchronize(cs);
            hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess)
            {
                // the second lane joins the capture through an event recorded on the capture
                // stream and is joined back before it ends, so the graph holds both lanes' chains
                Lane gl[bsh_context::kLanes];
                for (int k = 0; k < nlanes; ++k) gl[k] = lanes[k];
                gl[0].st = cs;
                ForkJoin fj;
                if (nlanes > 1) rc = fj.fork(cs, c->lane_stream, nlanes - 1);
                if (rc == BSH_OK) rc = enqueue_batch(c, gl, nlanes, chunks, hjobs, count, tune, nullptr, gr);
                if (rc == BSH_OK && nlanes > 1) rc = fj.join(cs, c->lane_stream, nlanes - 1);
                hipError_t e2 = hipStreamEndCapture(cs, &graph);
                if (rc == BSH_OK && e2 != hipSuccess) rc = hip_fail(e2, "hipStreamEndCapture");
            }
            else
            {
                rc = hip_fail(e, "hipStreamBeginCapture");
            }
            if (rc == BSH_OK)
            {
                e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
            }
        }
        if (rc == BSH_OK)
        {
            (void)hipEventRecord(e0, cs);
            hipError_t e = hipGraphLaunch(exec, cs);
            if (e != hipSuccess) rc = hip_fail(e, "hipGraphLaunch");
            (void)hipEventRecord(e1, cs);
            if (rc == BSH_OK)
            {
                e = hipEventSynchronize(e1);
                if (e != hipSuccess) rc = hip_fail(e, "graph execution");
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            S.total_ms = ms;
        }
        if (!replay)
        {
            if (rc == BSH_OK && cacheable)
            {
                c->graph_exec = exec;
                c->graph = graph;
                c->graph_key.swap(key);
            }
            else
            {
                if (exec) (void)hipGraphExecDestroy(exec);
                if (graph) (void)hipGraphDestroy(graph);
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        S.graph_replayed = replay ? 1u : 0u;
        if (rc != BSH_OK) return rc;
        S.trace_launches = S.shade_launches = (uint32_t)pairs;
    }
    else
    {
        LaunchTimer T;
        BSH_TRY(T.init(pairs));
        BSH_HIP(hipEventRecord(T.next(), st));
        // the second lane starts after everything before it on `st` (job table, cleared
        // statistics) and `st` continues after both lanes
        ForkJoin fj;
        if (nlanes > 1) BSH_TRY(fj.fork(st, c->lane_stream, nlanes - 1));
        BSH_TRY(enqueue_batch(c, lanes, nlanes, chunks, hjobs, count, tune, &T, gr));
        if (nlanes > 1) BSH_TRY(fj.join(st, c->lane_stream, nlanes - 1));
        BSH_HIP(hipEventRecord(T.next(), st));
        BSH_HIP(hipEventSynchronize(T.evs[T.ei - 1]));
        // per pass: before trace, after trace, after the guide inputs, after inference, after shade
        for (size_t k = 1; k + 4 < T.ei; k += 5)
        {
            float a = 0.f, g = 0.f, b = 0.f;
            BSH_HIP(hipEventElapsedTime(&a, T.evs[k], T.evs[k + 1]));
            BSH_HIP(hipEventElapsedTime(&g, T.evs[k + 2], T.evs[k + 3]));
            BSH_HIP(hipEventElapsedTime(&b, T.evs[k + 3], T.evs[k + 4]));
            S.trace_ms += a;
            S.shade_ms += b;
            launch_ms[T.bounce[(k - 1) / 5]] += a;
            S.trace_launches++;
            S.shade_launches++;
            if (guided && g > 0.f)
            {
                S.guide_ms += g;
                S.guide_launches++;
            }
        }
        float tot = 0.f;
        BSH_HIP(hipEventElapsedTime(&tot, T.evs[0], T.evs[T.ei - 1]));
        S.total_ms = tot;
    }
    std::vector<unsigned long long> st2(kStatAll);
    BSH_HIP(hipMemcpy(st2.data(), w.stats.p, kStatAll * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    c->launch_stats.assign(maxdepth + 1, bsh_launch_stats{});
    for (int b = 0; b <= maxdepth; ++b)
    {
        const unsigned long long* ls = &st2[kStatTotal + kLaunchStat * b];
        bsh_launch_stats& o = c->launch_stats[b];
        o.rays_closest = ls[0];
        o.rays_any = ls[1];
        o.nodes_closest = count ? ls[2] : 0;
        o.tris_closest = count ? ls[3] : 0;
        o.nodes_any = count ? ls[4] : 0;
        o.tris_any = count ? ls[5] : 0;
        o.trace_ms = launch_ms[b];
    }
    S.guide_rows = 0;
    for (int k = 0; k < kRowParts; ++k) S.guide_rows += st2[kStatRows + kRowStride * k];
    S.rays_closest = st2[0];
    S.rays_any = st2[1];
    S.node_visits = count ? st2[2] : 0;
    S.tri_tests = count ? st2[3] : 0;
    S.camera_samples = samples;
    if (record)
    {
        uint32_t nrec = 0;
        BSH_HIP(hipMemcpy(&nrec, w.rec_count.p, 4, hipMemcpyDeviceToHost));
        gr->written = std::min<size_t>(nrec, gr->cap);
    }
    return BSH_OK;
}

}  // namespace bsh
