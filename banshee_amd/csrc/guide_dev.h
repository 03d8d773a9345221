// Learned path-guiding distribution: a mixture of 8 von Mises-Fisher lobes whose
// parameters come from a small MLP of (position, normal, roughness). No reference
// counterpart; the plain BSDF-sampling tracer stays the fallback and error judge.
//
// Network: 7 raw inputs v0..v6 (position normalised to the scene box, GGX alpha or 1 for
// Lambert, shading normal) -> 96 encoded features in 6 blocks of 16, block b = band b:
// position p < 14 is sin (p even) or cos (p odd) of 2^b pi v_{p/2}, positions 14 and 15 the
// raw values v_{2b}, v_{2b+1} (zero past v6: 91 features and 5 zeros)
// -> 4 hidden layers of 64 ReLU units -> 48 outputs of which 40 are used:
//   o[0..7]   lobe logits (softmax -> weights pi_k)
//   o[8..31]  lobe mean directions, unnormalised (mu_k = o / |o|)
//   o[32..39] log concentrations, kappa_k = exp(clamp(o, -2, 6))
// vMF density (stable form): kappa / (2 pi (1 - e^{-2 kappa})) exp(kappa (mu.w - 1)).
// Sampling a lobe: W. Jakob, "Numerically stable sampling of the von Mises-Fisher
// distribution on S^2 (and other tricks)", 2012.
#pragma once

#include "bsh_math.h"

namespace bsh {

constexpr int kGIn = 96;       // padded encoded input width
constexpr int kGHid = 64;      // hidden width
constexpr int kGOut = 48;      // padded output width
constexpr int kGUsed = 40;     // outputs the mixture uses
constexpr int kGLobes = 8;
constexpr int kGRaw = 7;
constexpr int kGBands = 6;
constexpr float kLogKappaLo = -2.f, kLogKappaHi = 6.f;

// Parameter layout of the FP32 master vector (weights input-major W[k][n], then bias).
constexpr int kGW1 = 0, kGB1 = kGW1 + kGIn * kGHid;
constexpr int kGW2 = kGB1 + kGHid, kGB2 = kGW2 + kGHid * kGHid;
constexpr int kGW3 = kGB2 + kGHid, kGB3 = kGW3 + kGHid * kGHid;
constexpr int kGW4 = kGB3 + kGHid, kGB4 = kGW4 + kGHid * kGHid;
constexpr int kGW5 = kGB4 + kGHid, kGB5 = kGW5 + kGHid * kGOut;
constexpr int kGParams = kGB5 + kGOut;

// The five always-zero encoded features (raw slots past v6).
BSH_HD bool guide_pad_feature(int j)
{
    int b = j / 16, p = j % 16;
    return p >= 14 && 2 * b + (p - 14) >= kGRaw;
}

// Device-only: the decode, pdf and sampler use the hardware exp2 / log2 / reciprocal /
// reciprocal square root (relative error about 1e-6, far below the FP16 precision of the
// network outputs they read). The sampler and the pdf work from the same decoded lobes, so
// the density evaluated is the density sampled, to float rounding.
__device__ __forceinline__ float gexp(float x) { return __expf(x); }
__device__ __forceinline__ float glog(float x) { return __logf(x); }
__device__ __forceinline__ float grcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float grsq(float x) { return __builtin_amdgcn_rsqf(x); }

struct Lobes
{
    float w[kGLobes];      // mixture weights
    f3 mu[kGLobes];
    float kappa[kGLobes];
};

// Unit mean direction of lobe k from its raw outputs (0,0,1 when they are all zero).
template <class T>
__device__ __forceinline__ f3 lobe_mu(const T* o, int k)
{
    const f3 r = mk((float)o[8 + 3 * k], (float)o[9 + 3 * k], (float)o[10 + 3 * k]);
    const float n2 = sqnorm(r);
    return n2 > 0.f ? r * grsq(n2) : mk(0.f, 0.f, 1.f);
}

template <class T>
__device__ __forceinline__ Lobes guide_lobes(const T* o)
{
    Lobes L;
    float mx = (float)o[0];
    for (int k = 1; k < kGLobes; ++k) mx = fmaxf(mx, (float)o[k]);
    float s = 0.f;
    for (int k = 0; k < kGLobes; ++k)
    {
        L.w[k] = gexp((float)o[k] - mx);
        s += L.w[k];
    }
    const float is = grcp(s);
    for (int k = 0; k < kGLobes; ++k)
    {
        L.w[k] = L.w[k] * is;
        L.mu[k] = lobe_mu(o, k);
        L.kappa[k] = gexp(clampf((float)o[32 + k], kLogKappaLo, kLogKappaHi));
    }
    return L;
}

__device__ __forceinline__ bool lobes_finite(const Lobes& L)
{
    float s = 0.f;
    for (int k = 0; k < kGLobes; ++k) s += L.w[k] + L.kappa[k] + L.mu[k].x + L.mu[k].y + L.mu[k].z;
    return isfinite(s);
}

__device__ __forceinline__ float vmf_pdf(f3 mu, float kappa, f3 w)
{
    return kappa * grcp(2.f * kPi * (1.f - gexp(-2.f * kappa))) * gexp(kappa * (dot(mu, w) - 1.f));
}

__device__ __forceinline__ float guide_pdf(const Lobes& L, f3 w)
{
    float p = 0.f;
    for (int k = 0; k < kGLobes; ++k) p += L.w[k] * vmf_pdf(L.mu[k], L.kappa[k], w);
    return p;
}

// Lobe chosen by u0 through the weight CDF, then Jakob's stable vMF sample with (u1, u2).
// The chosen lobe's direction and concentration are decoded again from the raw outputs `o`
// (the same operations as guide_lobes, so the same values): indexing the Lobes registers by
// a computed lobe would move them to scratch memory.
template <class T>
__device__ __forceinline__ f3 guide_sample(const Lobes& L, const T* o, float u0, float u1, float u2)
{
    int k = kGLobes - 1;
    float c = 0.f;
    bool found = false;
#pragma unroll
    for (int j = 0; j < kGLobes; ++j)
    {
        c += L.w[j];
        bool take = !found && u0 < c;
        k = take ? j : k;
        found = found || take;
    }
    const f3 mu = lobe_mu(o, k);
    const float kap = gexp(clampf((float)o[32 + k], kLogKappaLo, kLogKappaHi));
    float w = 1.f + glog(u1 + (1.f - u1) * gexp(-2.f * kap)) * grcp(kap);
    w = clampf(w, -1.f, 1.f);
    float s = sqrtf(fmaxf(0.f, 1.f - w * w));
    f3 t1 = orthovector(mu);
    f3 t2 = cross(mu, t1);
    // azimuth 2 pi u2; the hardware sine / cosine take revolutions: sin(2 pi u2) = amdgcn_sin(u2)
    return normalize(mu * w + t1 * (s * __builtin_amdgcn_cosf(u2)) + t2 * (s * __builtin_amdgcn_sinf(u2)));
}

}  // namespace bsh
