// Spherical-harmonic kernels: environment-map projection (ShProjectEnvironmentMap,
// math/shproject.cpp:73-130), irradiance from SH (ShTracer::GetE, tracer/shtracer.cpp:20-36)
// and the basis diagnostic. Probe baking lives with the wavefront kernels in render.hip.
#include <hip/hip_runtime.h>

#include <cmath>
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

}  // namespace bsh

using namespace bsh;

extern "C" {

bsh_status bsh_sh_project_env(bsh_context* ctx, const float* env, int width, int height, int lmax, float* coeffs,
                              void* stream)
{
    if (!ctx || !env || !coeffs) return fail(BSH_ERR_INVALID, "bsh_sh_project_env: NULL argument");
    if (width <= 0 || height <= 0) return fail(BSH_ERR_INVALID, "bsh_sh_project_env: bad size");
    if (lmax < 0 || lmax > kShMaxL) return fail(BSH_ERR_UNSUPPORTED, "bsh_sh_project_env: lmax must be 0..4");
    if ((size_t)width * height >= ((size_t)1 << 31)) return fail(BSH_ERR_UNSUPPORTED, "bsh_sh_project_env: too large");
    BSH_HIP(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    // sin/cos tables exactly as the reference computes them on the host (shproject.cpp:85-104)
    std::vector<float> sth(height), cth(height), sph(width), cph(width);
    float thetastep = kPi / height, phistep = 2.f * kPi / width;
    float theta0 = kPi / height / 2, phi0 = 2.f * kPi / width / 2;
    for (int i = 0; i < width; ++i)
    {
        sph[i] = std::sin(phi0 + i * phistep);
        cph[i] = std::cos(phi0 + i * phistep);
    }
    for (int i = 0; i < height; ++i)
    {
        sth[i] = std::sin(theta0 + i * thetastep);
        cth[i] = std::cos(theta0 + i * thetastep);
    }
    DevBuf<float> d_sth, d_cth, d_sph, d_cph, d_part;
    BSH_HIP(d_sth.alloc(height));
    BSH_HIP(d_cth.alloc(height));
    BSH_HIP(d_sph.alloc(width));
    BSH_HIP(d_cph.alloc(width));
    BSH_HIP(hipMemcpyAsync(d_sth.p, sth.data(), height * 4, hipMemcpyHostToDevice, st));
    BSH_HIP(hipMemcpyAsync(d_cth.p, cth.data(), height * 4, hipMemcpyHostToDevice, st));
    BSH_HIP(hipMemcpyAsync(d_sph.p, sph.data(), width * 4, hipMemcpyHostToDevice, st));
    BSH_HIP(hipMemcpyAsync(d_cph.p, cph.data(), width * 4, hipMemcpyHostToDevice, st));
    int nt3 = 3 * (lmax + 1) * (lmax + 1);
    int nb = (width * height + kShBlock - 1) / kShBlock;
    BSH_HIP(d_part.alloc((size_t)nb * nt3));
    dim3 g(nb), b(kShBlock);
    switch (lmax)
    {
    case 0: hipLaunchKernelGGL(k_env_partial<0>, g, b, 0, st, env, width, height, d_sth.p, d_cth.p, d_sph.p, d_cph.p, d_part.p); break;
    case 1: hipLaunchKernelGGL(k_env_partial<1>, g, b, 0, st, env, width, height, d_sth.p, d_cth.p, d_sph.p, d_cph.p, d_part.p); break;
    case 2: hipLaunchKernelGGL(k_env_partial<2>, g, b, 0, st, env, width, height, d_sth.p, d_cth.p, d_sph.p, d_cph.p, d_part.p); break;
    case 3: hipLaunchKernelGGL(k_env_partial<3>, g, b, 0, st, env, width, height, d_sth.p, d_cth.p, d_sph.p, d_cph.p, d_part.p); break;
    default: hipLaunchKernelGGL(k_env_partial<4>, g, b, 0, st, env, width, height, d_sth.p, d_cth.p, d_sph.p, d_cph.p, d_part.p); break;
    }
    hipLaunchKernelGGL(k_env_final, dim3(1), dim3(128), 0, st, d_part.p, nb, nt3, coeffs);
    BSH_HIP(hipGetLastError());
    BSH_HIP(hipStreamSynchronize(st));   // the temporaries are freed on return
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
