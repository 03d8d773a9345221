// Real spherical harmonics in the reference's convention (z is the pole, Condon-Shortley
// phase, index l*l + l + m, m < 0 uses sin(|m| phi)): math/sh.h:51-60, math/sh.cpp:11-180.
// Same float operations in the same order as the reference, for lmax <= kShMaxL.
#pragma once

#include "bsh_math.h"

namespace bsh {

constexpr int kShMaxL = 4;
constexpr int kShMaxTerms = (kShMaxL + 1) * (kShMaxL + 1);

BSH_HD int sh_index(int l, int m) { return l * l + l + m; }

// EvaluateLegendrePolynomial, sh.cpp:11-50
BSH_HD void sh_legendre(float x, int lmax, float* out)
{
    out[sh_index(0, 0)] = 1.f;
    if (lmax >= 1) out[sh_index(1, 0)] = x;
    for (int l = 2; l <= lmax; ++l)
        out[sh_index(l, 0)] = ((2 * l - 1) * x * out[sh_index(l - 1, 0)] - (l - 1) * out[sh_index(l - 2, 0)]) / l;
    float neg = -1.f, dfact = 1.f;
    float xroot = sqrtf(fmaxf(0.f, 1.f - x * x));
    float xpow = xroot;
    for (int l = 1; l <= lmax; ++l)
    {
        out[sh_index(l, l)] = neg * dfact * xpow;
        neg *= -1.f;
        dfact *= 2 * l + 1;
        xpow *= xroot;
    }
    for (int l = 2; l <= lmax; ++l) out[sh_index(l, l - 1)] = x * (2 * l - 1) * out[sh_index(l - 1, l - 1)];
    for (int l = 3; l <= lmax; ++l)
        for (int m = 1; m <= l - 2; ++m)
            out[sh_index(l, m)] =
                ((2 * (l - 1) + 1) * x * out[sh_index(l - 1, m)] - (l - 1 + m) * out[sh_index(l - 2, m)]) / (l - m);
}

// DivFact and K(l, m), sh.cpp:54-77
BSH_HD float sh_divfact(int a, int b)
{
    if (b == 0) return 1.f;
    float fa = (float)a, fb = fabsf((float)b), v = 1.f;
    for (float x = fa - fb + 1.f; x <= fa + fb; x += 1.f) v *= x;
    return 1.f / v;
}
BSH_HD float sh_K(int l, int m) { return sqrtf((2.f * l + 1.f) * 1.f / (4 * kPi) * sh_divfact(l, m)); }

// ShEvaluate, sh.cpp:97-153
BSH_HD void sh_evaluate(f3 p, int lmax, float* coefs)
{
    sh_legendre(p.z, lmax, coefs);
    float cosm[kShMaxL + 1], sinm[kShMaxL + 1];
    float xylen = sqrtf(fmaxf(0.f, 1.f - p.z * p.z));
    if (xylen == 0.f)
    {
        for (int i = 0; i <= lmax; ++i)
        {
            sinm[i] = 0.f;
            cosm[i] = 1.f;
        }
    }
    else
    {
        // SinCosIndexed(p.y / xylen, p.x / xylen, lmax + 1), sh.cpp:80-92
        float s = p.y / xylen, c = p.x / xylen;
        float si = 0.f, ci = 1.f;
        for (int i = 0; i <= lmax; ++i)
        {
            sinm[i] = si;
            cosm[i] = ci;
            float oldsi = si;
            si = si * c + ci * s;
            ci = ci * c - oldsi * s;
        }
    }
    const float sqrt2 = sqrtf(2.f);
    for (int l = 0; l <= lmax; ++l)
    {
        for (int m = -l; m < 0; ++m)
            coefs[sh_index(l, m)] = sqrt2 * sh_K(l, m) * coefs[sh_index(l, -m)] * sinm[-m];
        coefs[sh_index(l, 0)] *= sh_K(l, 0);
        for (int m = 1; m <= l; ++m) coefs[sh_index(l, m)] *= sqrt2 * sh_K(l, m) * cosm[m];
    }
}

// ShConvolveCosTheta table and lambda, sh.cpp:155-180 (lmax <= 4 entries)
BSH_HD float sh_costheta(int l)
{
    const float t[5] = {0.8862268925f, 1.0233267546f, 0.4954159260f, 0.0000000000f, -0.1107783690f};
    return t[l];
}
BSH_HD float sh_lambda(int l) { return sqrtf((4.f * (float)3.14159265358979323846) / (2.f * (float)l + 1.f)); }

}  // namespace bsh
