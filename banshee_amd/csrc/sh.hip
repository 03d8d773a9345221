// Spherical-harmonic kernels: environment-map projection (ShProjectEnvironmentMap,
// math/shproject.cpp:73-130), irradiance from SH (ShTracer::GetE, tracer/shtracer.cpp:20-36)
// and the basis diagnostic. Probe baking lives with the wavefront kernels in render.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"
#include "sh_dev.h"

namespace bsh {

constexpr int kShBlock = 256;

// Pass 1: each block sums the contributions of kShBlock consecutive texels, in the
// reference's loop order (phi outer, theta inner), with a fixed LDS tree.
template <int LMAX>
__global__ __launch_bounds__(kShBlock) void k_env_partial(const float* env, int W, int H, const float* sinth,
                                                        const float* costh, const float* sinph, const float* cosph,
                                                        float* partial)
{
    constexpr int NT = (LMAX + 1) * (LMAX + 1);
    __shared__ float red[kShBlock];
    int k = blockIdx.x * kShBlock + threadIdx.x;
    float c[3 * NT];
#pragma unroll
    for (int i = 0; i < 3 * NT; ++i) c[i] = 0.f;
    if (k < W * H)
    {
        int phi = k / H, theta = k % H;
        f3 w = normalize(mk(sinth[theta] * cosph[phi], costh[theta], sinth[theta] * sinph[phi]));
        float ux = (float)phi / W, uy = (float)theta / H;
        int iu = (int)floorf(ux * W), iv = (int)floorf(uy * H);
        const float* e = env + 3 * ((size_t)W * iv + iu);
        float y[kShMaxTerms];
        sh_evaluate(w, LMAX, y);
        float a = kPi / H, b = 2.f * kPi / W;
#pragma unroll
        for (int i = 0; i < NT; ++i)
        {
            c[3 * i + 0] = e[0] * y[i] * sinth[theta] * a * b;
            c[3 * i + 1] = e[1] * y[i] * sinth[theta] * a * b;
            c[3 * i + 2] = e[2] * y[i] * sinth[theta] * a * b;
        }
    }
    for (int i = 0; i < 3 * NT; ++i)
    {
        red[threadIdx.x] = c[i];
        __syncthreads();
        for (int st = kShBlock / 2; st > 0; st >>= 1)
        {
            if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(size_t)blockIdx.x * 3 * NT + i] = red[0];
        __syncthreads();
    }
}

// Pass 2: one thread per coefficient sums the block partials in block order.
__global__ void k_env_final(const float* partial, int nblocks, int n, float* coeffs)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int b = 0; b < nblocks; ++b) s += partial[(size_t)b * n + i];
    coeffs[i] = s;
}

__global__ void k_irradiance(int lmax, const float* coeffs, const float* normals, uint32_t n, float* out)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int nt = (lmax + 1) * (lmax + 1);
    float y[kShMaxTerms];
    sh_evaluate(mk(normals[3 * t], normals[3 * t + 1], normals[3 * t + 2]), lmax, y);
    f3 e = mk(0.f, 0.f, 0.f);
    for (int l = 0; l <= lmax; ++l)
    {
        float lam = sh_lambda(l);
        for (int m = -l; m <= l; ++m)
        {
            int o = sh_index(l, m);
            // ShConvolveCosTheta: lambda(l) * cin * s_costheta[l]
            f3 cc = lam * mk(coeffs[3 * o], coeffs[3 * o + 1], coeffs[3 * o + 2]) * sh_costheta(l);
            e = e + y[o] * cc * (0.7f / 3.14159265358979323846f);
        }
    }
    (void)nt;
    out[3 * t] = e.x;
    out[3 * t + 1] = e.y;
    out[3 * t + 2] = e.z;
}

__global__ void k_diag_sh(int lmax, uint32_t n, const float* dirs, float* out)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int nt = (lmax + 1) * (lmax + 1);
    float y[kShMaxTerms];
    sh_evaluate(mk(dirs[3 * t], dirs[3 * t + 1], dirs[3 * t + 2]), lmax, y);
    for (int i = 0; i < nt; ++i) out[(size_t)t * nt + i] = y[i];
}

// The sin/cos tables exactly as the reference computes them on the host
// (shproject.cpp:85-104), laid out as sin(theta)[H], cos(theta)[H], sin(phi)[W], cos(phi)[W];
// writes entries [first, first + n) of that concatenation.
static void env_tables(int W, int H, size_t first, size_t n, float* out)
{
    float thetastep = kPi / H, phistep = 2.f * kPi / W;
    float theta0 = kPi / H / 2, phi0 = 2.f * kPi / W / 2;
    for (size_t k = 0; k < n; ++k)
    {
        size_t i = first + k;
        if (i < (size_t)H) out[k] = std::sin(theta0 + (int)i * thetastep);
        else if (i < 2 * (size_t)H) out[k] = std::cos(theta0 + (int)(i - H) * thetastep);
        else if (i < 2 * (size_t)H + W) out[k] = std::sin(phi0 + (int)(i - 2 * H) * phistep);
        else out[k] = std::cos(phi0 + (int)(i - 2 * H - W) * phistep);
    }
}

static size_t env_partial_floats(int W, int H, int lmax)
{
    size_t nb = ((size_t)W * H + kShBlock - 1) / kShBlock;
    return nb * 3 * (lmax + 1) * (lmax + 1);
}

static bsh_status check_env_args(bsh_context* ctx, const float* env, int width, int height, int lmax,
                                 const float* coeffs, const char* who)
{
    std::string w(who);
    if (!ctx || !env || !coeffs) return fail(BSH_ERR_INVALID, w + ": NULL argument");
    if (width <= 0 || height <= 0) return fail(BSH_ERR_INVALID, w + ": bad size");
    if (lmax < 0 || lmax > kShMaxL) return fail(BSH_ERR_UNSUPPORTED, w + ": lmax must be 0..4");
    if ((size_t)width * height >= ((size_t)1 << 31)) return fail(BSH_ERR_UNSUPPORTED, w + ": too large");
    return BSH_OK;
}

// Both passes over `tab` (the four tables, 2(W+H) floats) and `partial` (one block's
// 3(lmax+1)^2 sums per kShBlock texels, every entry written by pass 1 before pass 2 reads it).
static hipError_t launch_env_projection(const float* env, int W, int H, int lmax, const float* tab, float* partial,
                                        float* coeffs, hipStream_t st)
{
    const float *sth = tab, *cth = tab + H, *sph = tab + 2 * H, *cph = tab + 2 * H + W;
    int nt3 = 3 * (lmax + 1) * (lmax + 1);
    int nb = (int)(((size_t)W * H + kShBlock - 1) / kShBlock);
    dim3 g(nb), b(kShBlock);
    switch (lmax)
    {
    case 0: hipLaunchKernelGGL(k_env_partial<0>, g, b, 0, st, env, W, H, sth, cth, sph, cph, partial); break;
    case 1: hipLaunchKernelGGL(k_env_partial<1>, g, b, 0, st, env, W, H, sth, cth, sph, cph, partial); break;
    case 2: hipLaunchKernelGGL(k_env_partial<2>, g, b, 0, st, env, W, H, sth, cth, sph, cph, partial); break;
    case 3: hipLaunchKernelGGL(k_env_partial<3>, g, b, 0, st, env, W, H, sth, cth, sph, cph, partial); break;
    default: hipLaunchKernelGGL(k_env_partial<4>, g, b, 0, st, env, W, H, sth, cth, sph, cph, partial); break;
    }
    hipLaunchKernelGGL(k_env_final, dim3(1), dim3(128), 0, st, partial, nb, nt3, coeffs);
    return hipGetLastError();
}

// Scratch of the allocation-free projection: one copy per device, held until the process
// ends. Calls on one device share it, so calls that overlap on one device race.
static __device__ __align__(256) float scratch[kEnvScratchFloats];

static std::mutex g_scratch_mu;
static float* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on `device`, looked up once per device with that device
// current (in host code the array's name is only a stand-in for it).
bsh_status env_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(BSH_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return BSH_OK;
    BSH_HIP(hipSetDevice(device));
    void* p = nullptr;
    BSH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(BSH_ERR_STATE, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_dev[device] = (float*)p;
    return BSH_OK;
}

float* env_scratch_ptr(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_dev[device];
}

}  // namespace bsh

using namespace bsh;

extern "C" {

bsh_status bsh_sh_project_env(bsh_context* ctx, const float* env, int width, int height, int lmax, float* coeffs,
                              void* stream)
{
    bsh_status v = check_env_args(ctx, env, width, height, lmax, coeffs, "bsh_sh_project_env");
    if (v != BSH_OK) return v;
    BSH_HIP(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    std::vector<float> tab(2 * ((size_t)width + height));
    env_tables(width, height, 0, tab.size(), tab.data());
    DevBuf<float> d_tab, d_part;
    BSH_HIP(d_tab.alloc(tab.size()));
    BSH_HIP(hipMemcpyAsync(d_tab.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, st));
    BSH_HIP(d_part.alloc(env_partial_floats(width, height, lmax)));
    BSH_HIP(launch_env_projection(env, width, height, lmax, d_tab.p, d_part.p, coeffs, st));
    BSH_HIP(hipStreamSynchronize(st));   // the temporaries are freed on return
    return BSH_OK;
}

bsh_status bsh_sh_project_env_noalloc(bsh_context* ctx, const float* env, int width, int height, int lmax,
                                      float* coeffs, void* stream)
{
    bsh_status v = check_env_args(ctx, env, width, height, lmax, coeffs, "bsh_sh_project_env_noalloc");
    if (v != BSH_OK) return v;
    size_t ntab = 2 * ((size_t)width + height), npart = env_partial_floats(width, height, lmax);
    size_t part_off = (ntab + 63) & ~(size_t)63;   // partials start on a 256-B boundary
    if (part_off + npart > kEnvScratchFloats)
        return fail(BSH_ERR_INVALID, "bsh_sh_project_env_noalloc: a " + std::to_string(width) + "x" +
                                         std::to_string(height) + " map at lmax " + std::to_string(lmax) + " needs " +
                                         std::to_string((part_off + npart) * 4) + " B of scratch, more than its " +
                                         std::to_string(kEnvScratchFloats * 4) + " B; use bsh_sh_project_env");
    BSH_HIP(hipSetDevice(ctx->device));
    float* base = env_scratch_ptr(ctx->device);
    if (!base) return fail(BSH_ERR_STATE, "bsh_sh_project_env_noalloc: scratch of device " +
                                              std::to_string(ctx->device) + " was not set up by bsh_context_create");
    hipStream_t st = (hipStream_t)stream;
    // the host tables go up through a fixed stack buffer, one piece at a time; each piece is
    // waited for before the buffer is refilled
    float piece[2048];
    for (size_t o = 0; o < ntab; o += 2048)
    {
        size_t n = std::min<size_t>(2048, ntab - o);
        env_tables(width, height, o, n, piece);
        BSH_HIP(hipMemcpyAsync(base + o, piece, n * 4, hipMemcpyHostToDevice, st));
        BSH_HIP(hipStreamSynchronize(st));
    }
    BSH_HIP(launch_env_projection(env, width, height, lmax, base, base + part_off, coeffs, st));
    return BSH_OK;
}

bsh_status bsh_sh_irradiance(bsh_context* ctx, int lmax, const float* coeffs, const float* normals, size_t n,
                             float* out, void* stream)
{
    if (!ctx || !coeffs || (n && (!normals || !out))) return fail(BSH_ERR_INVALID, "bsh_sh_irradiance: NULL argument");
    if (lmax < 0 || lmax > kShMaxL) return fail(BSH_ERR_UNSUPPORTED, "bsh_sh_irradiance: lmax must be 0..4");
    if (n == 0) return BSH_OK;
    BSH_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_irradiance, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, lmax, coeffs,
                       normals, (uint32_t)n, out);
    BSH_HIP(hipGetLastError());
    return BSH_OK;
}

bsh_status bsh_diag_sh_eval(bsh_context* ctx, int lmax, size_t n, const float* dirs, float* out)
{
    if (!ctx || (n && (!dirs || !out))) return fail(BSH_ERR_INVALID, "bsh_diag_sh_eval: NULL argument");
    if (lmax < 0 || lmax > kShMaxL) return fail(BSH_ERR_UNSUPPORTED, "bsh_diag_sh_eval: lmax must be 0..4");
    if (n == 0) return BSH_OK;
    BSH_HIP(hipSetDevice(ctx->device));
    int nt = (lmax + 1) * (lmax + 1);
    DevBuf<float> d, o;
    BSH_HIP(d.alloc(3 * n));
    BSH_HIP(o.alloc(nt * n));
    BSH_HIP(hipMemcpy(d.p, dirs, 3 * n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_diag_sh, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0, lmax, (uint32_t)n, d.p, o.p);
    BSH_HIP(hipGetLastError());
    BSH_HIP(hipMemcpy(out, o.p, nt * n * 4, hipMemcpyDeviceToHost));
    return BSH_OK;
}

}  // extern "C"
