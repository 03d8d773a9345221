// Diagnostic entry points: run single device routines (BSDFs, area lights, camera,
// keyed stream) over host arrays so tests can pin them against the reference one
// routine at a time. Not used by the render path.
#include <hip/hip_runtime.h>

#include <vector>

#include "internal.h"
#include "shading.h"

namespace bsh {

__global__ void k_diag_bsdf(float4 m0, float4 m1, uint32_t n, const float* frames, const float* view, const float* u,
                            const float* ea, const float* eb, float* out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Mat m;
    m.type = (int)m0.x;
    m.c = mk(m0.y, m0.z, m0.w);
    m.eta = m1.x;
    m.alpha = m1.y;
    m.x = m1.z;
    m.y = m1.w;
    const float* f = frames + 12 * i;
    Frame h;
    h.p = mk(0.f, 0.f, 0.f);
    h.n = mk(f[0], f[1], f[2]);
    h.ng = mk(f[3], f[4], f[5]);
    h.dpdu = mk(f[6], f[7], f[8]);
    h.dpdv = mk(f[9], f[10], f[11]);
    f3 wi = mk(view[3 * i], view[3 * i + 1], view[3 * i + 2]);
    f3 wo;
    float pdf = 0.f;
    int type = 0;
    f3 fs = mat_sample(m, h, u[2 * i], u[2 * i + 1], wi, wo, pdf, type);
    f3 a = mk(ea[3 * i], ea[3 * i + 1], ea[3 * i + 2]), b = mk(eb[3 * i], eb[3 * i + 1], eb[3 * i + 2]);
    f3 fe = mat_eval(m, h, a, b);
    float pe = mat_pdf(m, h, a, b);
    float* o = out + 12 * i;
    o[0] = fs.x; o[1] = fs.y; o[2] = fs.z; o[3] = pdf;
    o[4] = wo.x; o[5] = wo.y; o[6] = wo.z; o[7] = (float)type;
    o[8] = fe.x; o[9] = fe.y; o[10] = fe.z; o[11] = pe;
}

__global__ void k_diag_light(SceneDev s, int light, uint32_t n, const float* p, const float* u, const float* w,
                             float* out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const LightDev* L = s.lights + light;
    f3 pp = mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    f3 d;
    float pdf = 0.f;
    f3 le = light_sample(L, pp, u[2 * i], u[2 * i + 1], d, pdf);
    float gp = light_pdf(L, pp, mk(w[3 * i], w[3 * i + 1], w[3 * i + 2]));
    float* o = out + 8 * i;
    o[0] = le.x; o[1] = le.y; o[2] = le.z; o[3] = pdf;
    o[4] = d.x; o[5] = d.y; o[6] = d.z; o[7] = gp;
}

__global__ void k_diag_camera(CameraDev c, uint32_t n, const float* smp, float* rays)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float hx = smp[2 * i] - 0.5f, hy = smp[2 * i + 1] - 0.5f;
    float csx = hx * c.dimx, csy = hy * c.dimy;
    f3 d = normalize(c.znear * mk(c.fwd[0], c.fwd[1], c.fwd[2]) + csx * mk(c.right[0], c.right[1], c.right[2]) +
                     csy * mk(c.up[0], c.up[1], c.up[2]));
    float* o = rays + 8 * i;
    o[0] = c.eye[0]; o[1] = c.eye[1]; o[2] = c.eye[2]; o[3] = c.znear;
    o[4] = d.x; o[5] = d.y; o[6] = d.z; o[7] = c.zfar;
}

__global__ void k_diag_keyed(uint32_t seed, uint32_t n, const uint32_t* cs, const uint32_t* stream, const uint32_t* ctr,
                             uint32_t* out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = ks_u32(ks_base(seed, cs[i]), stream[i], ctr[i]);
}

// Image samples of pixel p, sample s (one-row image of npix pixels) or the k-th light
// sampler draw of camera sample 0, exactly as k_raygen and k_shade form them.
__global__ void k_diag_samples(int kind, int arg, int image, uint32_t seed, uint32_t n, int spp, float* out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x, y;
    if (image)
    {
        int s = (int)(i % (uint32_t)spp);
        image_sample2d(kind, arg, s, ks_base(seed, i), ks_base(seed, i - (uint32_t)s), x, y);
    }
    else
        nee_sample2d(kind, ks_base(seed, 0u), KS_LIGHT, 2u * i, x, y);
    out[2 * i] = x;
    out[2 * i + 1] = y;
}

// Material::Sample / Evaluate / GetPdf of material row `midx` of the scene (every type), the
// rand() draws of case i keyed by ks_base(seed, i) from counter 0.
__global__ void k_diag_material(SceneDev s, int midx, uint32_t seed, uint32_t n, const float* frames, const float* view,
                                const float* u, const float* ea, const float* eb, float* out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Mat m = load_mat(s.mats, midx);
    TexCtx tx{s.tex, s.tex_desc, s.mat_maps, 0.f, 0.f};   // the probes' hit.uv is (0, 0)
    apply_maps(tx, midx, m);
    const float* f = frames + 12 * i;
    Frame h;
    h.p = mk(0.f, 0.f, 0.f);
    h.n = mk(f[0], f[1], f[2]);
    h.ng = mk(f[3], f[4], f[5]);
    h.dpdu = mk(f[6], f[7], f[8]);
    h.dpdv = mk(f[9], f[10], f[11]);
    f3 wi = mk(view[3 * i], view[3 * i + 1], view[3 * i + 2]);
    f3 wo = mk(0.f, 0.f, 0.f);
    float pdf = 0.f;
    int type = 0;
    RandStream rs{ks_base(seed, i), 0u};
    f3 fs = mat_sample_x(s.mats, tx, m, h, u[2 * i], u[2 * i + 1], wi, wo, pdf, type, rs);
    f3 a = mk(ea[3 * i], ea[3 * i + 1], ea[3 * i + 2]), b = mk(eb[3 * i], eb[3 * i + 1], eb[3 * i + 2]);
    f3 fe = mat_eval_x(s.mats, tx, m, h, a, b);
    float pe = mat_pdf_x(s.mats, tx, m, h, a, b);
    float* o = out + 12 * i;
    o[0] = fs.x; o[1] = fs.y; o[2] = fs.z; o[3] = pdf;
    o[4] = wo.x; o[5] = wo.y; o[6] = wo.z; o[7] = (float)type;
    o[8] = fe.x; o[9] = fe.y; o[10] = fe.z; o[11] = pe;
}

// Light::GetSample / GetPdf / GetLe of light `light` (any type) at points p with shading
// normals nrm, as the shade kernel calls them.
__global__ void k_diag_light_full(SceneDev s, int light, uint32_t n, const float* p, const float* nrm, const float* u,
                                  const float* w, float* out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const LightDev* L = s.lights + light;
    f3 pp = mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]), nn = mk(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
    f3 ww = mk(w[3 * i], w[3 * i + 1], w[3 * i + 2]);
    f3 d = mk(0.f, 0.f, 0.f), le, ge = mk(0.f, 0.f, 0.f);
    float pdf = 0.f, gp;
    if (L->type == LIGHT_AREA)
    {
        le = light_sample(L, pp, u[2 * i], u[2 * i + 1], d, pdf);
        gp = light_pdf(L, pp, ww);
    }
    else
    {
        le = other_light_sample(s, *L, pp, nn, u[2 * i], u[2 * i + 1], d, pdf);
        gp = other_light_pdf(s, *L, nn, ww);
        ge = other_light_le(s, *L, ww);
    }
    float* o = out + 16 * i;
    o[0] = le.x; o[1] = le.y; o[2] = le.z; o[3] = pdf;
    o[4] = d.x; o[5] = d.y; o[6] = d.z; o[7] = gp;
    o[8] = ge.x; o[9] = ge.y; o[10] = ge.z;
    for (int k = 11; k < 16; ++k) o[k] = 0.f;
}

template <class T>
static bsh_status up(DevBuf<T>& d, const T* h, size_t n)
{
    BSH_HIP(d.alloc(n));
    BSH_HIP(hipMemcpy(d.p, h, n * sizeof(T), hipMemcpyHostToDevice));
    return BSH_OK;
}

#define BSH_TRY(x)                        \
    do                                    \
    {                                     \
        bsh_status r_ = (x);              \
        if (r_ != BSH_OK) return r_;      \
    } while (0)

// Streaming read: the memory ceilings the trace kernel is measured against. Each pass reads the
// buffer's 16-B words in coalesced runs of kStreamBlock words per workgroup, four independent
// loads in flight per lane, and XORs them into a per-thread word written at the end (so no load
// can be dropped). per_xcd = 0: the grid partitions the buffer (every byte read once per pass:
// the HBM ceiling for buffers far beyond the 256 MB Infinity Cache, that cache's for buffers
// inside it). per_xcd = 1: workgroups w, w + 8, w + 16, ... (one XCD under round-robin dispatch)
// partition the buffer among themselves, so each XCD reads all of it: a buffer within one XCD's
// 4 MB L2 is then read from L2 on every pass after the first.
constexpr int kStreamBlock = 256;
constexpr int kStreamUnroll = 8;
typedef float sf4 __attribute__((ext_vector_type(4)));
template <bool kNT>
__global__ __launch_bounds__(kStreamBlock) void k_stream_read(const float4* p, size_t n16, int reps, int per_xcd,
                                                              uint32_t* sink)
{
    const size_t group = per_xcd ? blockIdx.x >> 3 : blockIdx.x;
    const size_t groups = per_xcd ? gridDim.x >> 3 : gridDim.x;
    const size_t stride = groups * kStreamBlock;
    const sf4* q = (const sf4*)p;
    uint32_t acc = 0;
    for (int r = 0; r < reps; ++r)
    {
        size_t i = group * kStreamBlock + threadIdx.x;
        for (; i + (kStreamUnroll - 1) * stride < n16; i += kStreamUnroll * stride)
        {
            sf4 v[kStreamUnroll];
#pragma unroll
            for (int k = 0; k < kStreamUnroll; ++k)
                v[k] = kNT ? __builtin_nontemporal_load(q + i + k * stride) : q[i + k * stride];
#pragma unroll
            for (int k = 0; k < kStreamUnroll; ++k)
                acc ^= __float_as_uint(v[k].x) ^ __float_as_uint(v[k].y) ^ __float_as_uint(v[k].z) ^ __float_as_uint(v[k].w);
        }
        for (; i < n16; i += stride)
        {
            const sf4 a = q[i];
            acc ^= __float_as_uint(a.x) ^ __float_as_uint(a.y) ^ __float_as_uint(a.z) ^ __float_as_uint(a.w);
        }
    }
    sink[(size_t)blockIdx.x * kStreamBlock + threadIdx.x] = acc;
}

}  // namespace bsh

using namespace bsh;

extern "C" {

bsh_status bsh_diag_bsdf(bsh_context* ctx, const bsh_material* mat, size_t n, const float* frames, const float* view,
                         const float* u, const float* ea, const float* eb, float* out)
{
    if (!ctx || !mat) return fail(BSH_ERR_INVALID, "bsh_diag_bsdf: NULL argument");
    if (n == 0) return BSH_OK;
    BSH_HIP(hipSetDevice(ctx->device));
    DevBuf<float> df, dv, du, da, db, dout;
    BSH_TRY(up(df, frames, 12 * n));
    BSH_TRY(up(dv, view, 3 * n));
    BSH_TRY(up(du, u, 2 * n));
    BSH_TRY(up(da, ea, 3 * n));
    BSH_TRY(up(db, eb, 3 * n));
    BSH_HIP(dout.alloc(12 * n));
    float4 m0 = make_float4(mat->type, mat->color[0], mat->color[1], mat->color[2]);
    float4 m1 = make_float4(mat->eta, mat->alpha, 0.f, 0.f);
    hipLaunchKernelGGL(k_diag_bsdf, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0, m0, m1, (uint32_t)n, df.p, dv.p,
                       du.p, da.p, db.p, dout.p);
    BSH_HIP(hipGetLastError());
    BSH_HIP(hipMemcpy(out, dout.p, 12 * n * sizeof(float), hipMemcpyDeviceToHost));
    return BSH_OK;
}

bsh_status bsh_diag_light(bsh_scene* s, int light, size_t n, const float* p, const float* u, const float* w, float* out)
{
    if (!s) return fail(BSH_ERR_INVALID, "bsh_diag_light: NULL scene");
    if (!s->committed) return fail(BSH_ERR_STATE, "bsh_diag_light: scene not committed");
    if (light < 0 || (size_t)light >= s->lights.size()) return fail(BSH_ERR_INVALID, "bsh_diag_light: bad area light");
    if (n == 0) return BSH_OK;
    BSH_HIP(hipSetDevice(s->ctx->device));
    DevBuf<float> dp, du, dw, dout;
    BSH_TRY(up(dp, p, 3 * n));
    BSH_TRY(up(du, u, 2 * n));
    BSH_TRY(up(dw, w, 3 * n));
    BSH_HIP(dout.alloc(8 * n));
    hipLaunchKernelGGL(k_diag_light, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0, s->dev(), light, (uint32_t)n,
                       dp.p, du.p, dw.p, dout.p);
    BSH_HIP(hipGetLastError());
    BSH_HIP(hipMemcpy(out, dout.p, 8 * n * sizeof(float), hipMemcpyDeviceToHost));
    return BSH_OK;
}

bsh_status bsh_diag_camera(bsh_context* ctx, const bsh_camera* cam, size_t n, const float* samples, float* rays)
{
    if (!ctx || !cam) return fail(BSH_ERR_INVALID, "bsh_diag_camera: NULL argument");
    if (n == 0) return BSH_OK;
    BSH_HIP(hipSetDevice(ctx->device));
    DevBuf<float> ds, dr;
    BSH_TRY(up(ds, samples, 2 * n));
    BSH_HIP(dr.alloc(8 * n));
    hipLaunchKernelGGL(k_diag_camera, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0, make_camera(*cam),
                       (uint32_t)n, ds.p, dr.p);
    BSH_HIP(hipGetLastError());
    BSH_HIP(hipMemcpy(rays, dr.p, 8 * n * sizeof(float), hipMemcpyDeviceToHost));
    return BSH_OK;
}

bsh_status bsh_diag_keyed(bsh_context* ctx, uint32_t seed, size_t n, const uint32_t* camsample, const uint32_t* stream,
                          const uint32_t* counter, uint32_t* out)
{
    if (!ctx) return fail(BSH_ERR_INVALID, "bsh_diag_keyed: NULL context");
    if (n == 0) return BSH_OK;
    BSH_HIP(hipSetDevice(ctx->device));
    DevBuf<uint32_t> a, b, c, o;
    BSH_TRY(up(a, camsample, n));
    BSH_TRY(up(b, stream, n));
    BSH_TRY(up(c, counter, n));
    BSH_HIP(o.alloc(n));
    hipLaunchKernelGGL(k_diag_keyed, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0, seed, (uint32_t)n, a.p, b.p,
                       c.p, o.p);
    BSH_HIP(hipGetLastError());
    BSH_HIP(hipMemcpy(out, o.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return BSH_OK;
}

bsh_status bsh_diag_light_full(bsh_scene* s, int light, size_t n, const float* p, const float* nrm, const float* u,
                               const float* w, float* out)
{
    if (!s || (n && (!p || !nrm || !u || !w || !out))) return fail(BSH_ERR_INVALID, "bsh_diag_light_full: NULL argument");
    if (!s->committed) return fail(BSH_ERR_STATE, "bsh_diag_light_full: scene not committed");
    if (light < 0 || (size_t)light >= s->h_lights.size()) return fail(BSH_ERR_INVALID, "bsh_diag_light_full: bad light");
    if (n == 0) return BSH_OK;
    BSH_HIP(hipSetDevice(s->ctx->device));
    DevBuf<float> dp, dn, du, dw, o;
    BSH_TRY(up(dp, p, 3 * n));
    BSH_TRY(up(dn, nrm, 3 * n));
    BSH_TRY(up(du, u, 2 * n));
    BSH_TRY(up(dw, w, 3 * n));
    BSH_HIP(o.alloc(16 * n));
    hipLaunchKernelGGL(k_diag_light_full, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0, s->dev(), light,
                       (uint32_t)n, dp.p, dn.p, du.p, dw.p, o.p);
    BSH_HIP(hipGetLastError());
    BSH_HIP(hipMemcpy(out, o.p, 16 * n * sizeof(float), hipMemcpyDeviceToHost));
    return BSH_OK;
}

bsh_status bsh_diag_material(bsh_scene* s, int material, uint32_t seed, size_t n, const float* frames,
                             const float* view, const float* u, const float* ea, const float* eb, float* out)
{
    if (!s || (n && (!frames || !view || !u || !ea || !eb || !out)))
        return fail(BSH_ERR_INVALID, "bsh_diag_material: NULL argument");
    if (!s->committed) return fail(BSH_ERR_STATE, "bsh_diag_material: scene not committed");
    if (material < 0 || (size_t)material >= s->mats.size()) return fail(BSH_ERR_INVALID, "bsh_diag_material: bad row");
    if (n == 0) return BSH_OK;
    BSH_HIP(hipSetDevice(s->ctx->device));
    DevBuf<float> df, dv, du, da, db, o;
    BSH_TRY(up(df, frames, 12 * n));
    BSH_TRY(up(dv, view, 3 * n));
    BSH_TRY(up(du, u, 2 * n));
    BSH_TRY(up(da, ea, 3 * n));
    BSH_TRY(up(db, eb, 3 * n));
    BSH_HIP(o.alloc(12 * n));
    hipLaunchKernelGGL(k_diag_material, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0, s->dev(), material, seed,
                       (uint32_t)n, df.p, dv.p, du.p, da.p, db.p, o.p);
    BSH_HIP(hipGetLastError());
    BSH_HIP(hipMemcpy(out, o.p, 12 * n * sizeof(float), hipMemcpyDeviceToHost));
    return BSH_OK;
}

bsh_status bsh_diag_samples(bsh_context* ctx, int kind, int arg, int image, uint32_t seed, size_t n, float* out)
{
    if (!ctx || (n && !out)) return fail(BSH_ERR_INVALID, "bsh_diag_samples: NULL argument");
    if (kind < SMP_STRATIFIED || kind > SMP_SOBOL || arg <= 0 || arg > 4096)
        return fail(BSH_ERR_INVALID, "bsh_diag_samples: bad sampler");
    if (n == 0) return BSH_OK;
    if (n > (1u << 30)) return fail(BSH_ERR_INVALID, "bsh_diag_samples: too many samples");
    int spp = image ? sampler_num_samples(kind, arg) : 1;
    if (image && n % (size_t)spp) return fail(BSH_ERR_INVALID, "bsh_diag_samples: n is not whole pixels");
    BSH_HIP(hipSetDevice(ctx->device));
    DevBuf<float> o;
    BSH_HIP(o.alloc(2 * n));
    hipLaunchKernelGGL(k_diag_samples, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0, kind, arg, image, seed,
                       (uint32_t)n, spp, o.p);
    BSH_HIP(hipGetLastError());
    BSH_HIP(hipMemcpy(out, o.p, 2 * n * sizeof(float), hipMemcpyDeviceToHost));
    return BSH_OK;
}

bsh_status bsh_diag_stream_read(bsh_context* ctx, const void* buf, size_t bytes, int reps, int per_xcd, double* ms)
{
    if (!ctx || !buf || !ms) return fail(BSH_ERR_INVALID, "bsh_diag_stream_read: NULL argument");
    if (bytes < 16 * kStreamBlock || bytes % 16 || reps < 1)
        return fail(BSH_ERR_INVALID, "bsh_diag_stream_read: bytes must be a multiple of 16 and >= 4 KiB, reps >= 1");
    BSH_HIP(hipSetDevice(ctx->device));
    // 32 workgroups per CU (full occupancy), or 8 with flag 4; a multiple of the 8 XCDs either way
    const uint32_t blocks = (uint32_t)ctx->num_cus * ((per_xcd & 4) ? 8 : 32);
    const bool nt = (per_xcd & 2) != 0;
    per_xcd &= 1;
    DevBuf<uint32_t> sink;
    BSH_HIP(sink.alloc((size_t)blocks * kStreamBlock));
    const size_t n16 = bytes / 16;
    hipEvent_t e0, e1;
    BSH_HIP(hipEventCreate(&e0));
    BSH_HIP(hipEventCreate(&e1));
    // one untimed pass (page mapping, and for cache-resident sizes the first fill), then the timed one
    auto launch = [&](int r) {
        if (nt)
            hipLaunchKernelGGL(k_stream_read<true>, dim3(blocks), dim3(kStreamBlock), 0, 0, (const float4*)buf, n16, r,
                               per_xcd, sink.p);
        else
            hipLaunchKernelGGL(k_stream_read<false>, dim3(blocks), dim3(kStreamBlock), 0, 0, (const float4*)buf, n16, r,
                               per_xcd, sink.p);
    };
    launch(1);
    (void)hipEventRecord(e0, 0);
    launch(reps);
    (void)hipEventRecord(e1, 0);
    hipError_t e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(e, "bsh_diag_stream_read");
    BSH_HIP(hipGetLastError());
    *ms = t;
    return BSH_OK;
}

}  // extern "C"
