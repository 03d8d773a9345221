// Sample patterns of the reference's samplers as pure functions of (sample index, pattern
// draws), so a wavefront thread computes its camera sample without the sampler's sequential
// state. The draws come from the keyed stream (bsh_math.h):
//   image sampler, per camera sample (pixel, s):
//     StratifiedSampler(g)  cell s mod g^2 + jitter from draws (KS_IMG, 0/1) of sample s
//     RandomSampler(n)      draws (KS_IMG, 0/1) of sample s
//     RegularSampler(g)     no draws
//     CmjSampler(g)         one pattern draw per pixel: (KS_IMG, 0) of the pixel's sample 0
//     SobolSampler(n)       two scramble draws per pixel: (KS_IMG, 0) and (KS_IMG, 1) of the
//                           pixel's sample 0 (ImageRenderer::Render resets the sampler after
//                           every pixel, imagerenderer.cpp:44-46, sobol_sampler.cpp:41-48)
//   light / BSDF samplers (one sample per call), the k-th call of a stream at counters
//   2k, 2k+1: stratified(1) and random(1) use both, CmjSampler(1) takes its pattern from 2k,
//   RegularSampler(1) draws nothing.
#pragma once

#include "bsh_math.h"

namespace bsh {

enum : int { SMP_STRATIFIED = 0, SMP_REGULAR = 1, SMP_RANDOM = 2, SMP_CMJ = 3, SMP_SOBOL = 4 };

// Samples per pixel of the image sampler built with argument `arg` (Sampler::num_samples):
// g^2 for the grid samplers, n for RandomSampler(n), the next power of two for
// SobolSampler(n) (upper_power_of_two, math/mathutils.h). 0 for an unknown kind.
BSH_HD int sampler_num_samples(int kind, int arg)
{
    switch (kind)
    {
    case SMP_STRATIFIED:
    case SMP_REGULAR:
    case SMP_CMJ: return arg * arg;
    case SMP_RANDOM: return arg;
    case SMP_SOBOL:
    {
        unsigned v = (unsigned)arg - 1u;   // smallest power of two >= arg
        v |= v >> 1;
        v |= v >> 2;
        v |= v >> 4;
        v |= v >> 8;
        v |= v >> 16;
        return (int)(v + 1u);
    }
    default: return 0;
    }
}

// Correlated multi-jittered sampling, A. Kensler, "Correlated Multi-Jittered Sampling",
// Pixar Technical Memo 13-01 (2013), as sampler/cmj_sampler.cpp:3-62 uses it: a hashed
// permutation of [0, l) by cycle walking over the next power of two, a hashed jitter, and
// the sample of index s in an m x n pattern. The constants are the memo's; sample-for-sample
// parity needs exactly these.
BSH_HD uint32_t cmj_permute(uint32_t i, uint32_t l, uint32_t p)
{
    uint32_t w = l - 1;
    w |= w >> 1;
    w |= w >> 2;
    w |= w >> 4;
    w |= w >> 8;
    w |= w >> 16;
    do
    {
        i ^= p;
        i *= 0xe170893dU;
        i ^= p >> 16;
        i ^= (i & w) >> 4;
        i ^= p >> 8;
        i *= 0x0929eb3fU;
        i ^= p >> 23;
        i ^= (i & w) >> 1;
        i *= 1U | p >> 27;
        i *= 0x6935fa69U;
        i ^= (i & w) >> 11;
        i *= 0x74dcb303U;
        i ^= (i & w) >> 2;
        i *= 0x9e501cc3U;
        i ^= (i & w) >> 2;
        i *= 0xc860a3dfU;
        i &= w;
        i ^= i >> 5;
    } while (i >= l);
    return (i + p) % l;
}

BSH_HD float cmj_jitter(uint32_t i, uint32_t p)
{
    i ^= p;
    i ^= i >> 17;
    i ^= i >> 10;
    i *= 0xb36534e5U;
    i ^= i >> 12;
    i ^= i >> 21;
    i *= 0x93fc4795U;
    i ^= 0xdf6e307fU;
    i ^= i >> 17;
    i *= 1U | p >> 18;
    return (float)i * (1.0f / 4294967808.0f);
}

// Sample s (0 <= s < m n) of pattern p. The reference passes p as an int and s as an int;
// the arithmetic below keeps its int/float mix and its order of operations.
BSH_HD void cmj_sample(int s, int m, int n, uint32_t p, float& x, float& y)
{
    int sx = (int)cmj_permute((uint32_t)(s % m), (uint32_t)m, p * 0xa511e9b3U);
    int sy = (int)cmj_permute((uint32_t)(s / m), (uint32_t)n, p * 0x63d83595U);
    float jx = cmj_jitter((uint32_t)s, p * 0xa399d265U);
    float jy = cmj_jitter((uint32_t)s, p * 0x711ad6a5U);
    x = ((float)(s % m) + ((float)sy + jx) / (float)n) / (float)m;
    y = ((float)(s / m) + ((float)sx + jy) / (float)m) / (float)n;
}

// Scrambled radical inverse in base 2 (van der Corput) and the second component of the
// (0,2)-sequence (Sobol's first two dimensions), T. Kollig and A. Keller, "Efficient
// Multidimensional Sampling", Eurographics 2002, with 24 bits kept
// (sampler/sobol_sampler.cpp:5-30).
BSH_HD float sobol_vdc(uint32_t n, uint32_t scramble)
{
    uint32_t r = __builtin_bitreverse32(n) ^ scramble;
    return (float)((r >> 8) & 0xffffffU) / (float)(1 << 24);
}

BSH_HD float sobol_02(uint32_t n, uint32_t scramble)
{
    for (uint32_t v = 1U << 31; n != 0; n >>= 1, v ^= v >> 1)
        if (n & 1U) scramble ^= v;
    return (float)((scramble >> 8) & 0xffffffU) / (float)(1 << 24);
}

// RegularSampler(g) sample s: the cell centres in the order the constructor pushes them,
// x outer and y inner (sampler/regular_sampler.h:46-58).
BSH_HD void regular_sample(int s, int g, float& x, float& y)
{
    float step = 1.f / (float)g;
    float step2 = 1.f / (2.f * (float)g);
    int i = s % (g * g);
    x = (float)(i / g) * step + step2;
    y = (float)(i % g) * step + step2;
}

// The k-th sample of a light or BSDF sampler built with argument 1, from stream `stream`
// at counter c = 2k.
BSH_HD void nee_sample2d(int kind, uint32_t base, uint32_t stream, uint32_t c, float& x, float& y)
{
    if (kind == SMP_REGULAR)
    {
        x = 0.5f;
        y = 0.5f;
    }
    else if (kind == SMP_CMJ)
        cmj_sample(0, 1, 1, ks_u32(base, stream, c), x, y);
    else
        ks_sample2d(base, stream, c, x, y);
}

// The image sample of camera sample s of a pixel. `base` keys sample s, `base0` the pixel's
// sample 0.
BSH_HD void image_sample2d(int kind, int arg, int s, uint32_t base, uint32_t base0, float& x, float& y)
{
    switch (kind)
    {
    case SMP_REGULAR: regular_sample(s, arg, x, y); break;
    case SMP_RANDOM: ks_sample2d(base, KS_IMG, 0, x, y); break;
    case SMP_CMJ: cmj_sample(s % (arg * arg), arg, arg, ks_u32(base0, KS_IMG, 0), x, y); break;
    case SMP_SOBOL:
        x = sobol_vdc((uint32_t)s, ks_u32(base0, KS_IMG, 0));
        y = sobol_02((uint32_t)s, ks_u32(base0, KS_IMG, 1));
        break;
    default:
    {
        // StratifiedSampler::Sample2D (sampler/stratified_sampler.cpp:7-24)
        float ux, uy;
        ks_sample2d(base, KS_IMG, 0, ux, uy);
        int idx = s % (arg * arg);
        float cell = 1.f / (float)arg;
        x = ((float)(idx % arg) + ux) * cell;
        y = ((float)(idx / arg) + uy) * cell;
    }
    }
}

}  // namespace bsh
