// The reference's lights other than AreaLight, on the device: PointLight, DirectionalLight,
// EnvironmentLight (cosine-sampled) and EnvironmentLightIs (sampled by a Distribution2D of the
// map's luminance), with the lat-long environment map read as the reference's texture system
// reads it: bilinear filtering with periodic wrap (TextureSystem::Options defaults,
// texture/texturesystem.h:35; OpenImageIO 1.6 InterpBilinear + WrapPeriodic behind it,
// texture/oiio_texturesystem.cpp), texel centres at half-integer coordinates.
#pragma once

#include "bsh_math.h"

namespace bsh {

enum : int { LIGHT_AREA = 0, LIGHT_POINT = 1, LIGHT_DIRECTIONAL = 2, LIGHT_ENV = 3, LIGHT_ENV_IS = 4 };

// EnvironmentLightIs's distribution grid (environment_light_is.cpp:10-11)
constexpr int kEnvDistW = 512, kEnvDistH = 256;

struct EnvDev
{
    const float* tex;   // h rows of w RGB texels; row 0 is v = 0 (theta = 0, the +y pole)
    int w, h;
    // Distribution2D(kEnvDistW, kEnvDistH) (math/distribution2d.cpp:5-27): marginal over rows
    // and one conditional per row, each as Distribution1D's funcvals_, normalised cdf_, funcint_
    const float* mval;   // kEnvDistH
    const float* mcdf;   // kEnvDistH + 1
    const float* cval;   // kEnvDistH x kEnvDistW
    const float* ccdf;   // kEnvDistH x (kEnvDistW + 1)
    const float* cint;   // kEnvDistH
    float mint;
};

BSH_HD int wrap_index(int i, int n) { return ((i % n) + n) % n; }

// Bilinear lookup with periodic wrap at (u, v): (1-t)((1-s)c00 + s c10) + t((1-s)c01 + s c11)
// over the four texels around (u w - 1/2, v h - 1/2).
BSH_HD f3 env_bilinear(const float* tex, int w, int h, float u, float v)
{
    float s = u * (float)w - 0.5f, t = v * (float)h - 0.5f;
    float sf = floorf(s), tf = floorf(t);
    float fs = s - sf, ft = t - tf;
    int x0 = wrap_index((int)sf, w), x1 = wrap_index((int)sf + 1, w);
    int y0 = wrap_index((int)tf, h), y1 = wrap_index((int)tf + 1, h);
    const float* a = tex + 3 * ((size_t)y0 * w + x0);
    const float* b = tex + 3 * ((size_t)y0 * w + x1);
    const float* c = tex + 3 * ((size_t)y1 * w + x0);
    const float* d = tex + 3 * ((size_t)y1 * w + x1);
    f3 r;
    r.x = (1.f - ft) * ((1.f - fs) * a[0] + fs * b[0]) + ft * ((1.f - fs) * c[0] + fs * d[0]);
    r.y = (1.f - ft) * ((1.f - fs) * a[1] + fs * b[1]) + ft * ((1.f - fs) * c[1] + fs * d[1]);
    r.z = (1.f - ft) * ((1.f - fs) * a[2] + fs * b[2]) + ft * ((1.f - fs) * c[2] + fs * d[2]);
    return r;
}

// cartesian_to_spherical (math/mathutils.h:68-74) and the lat-long coordinates the lights
// fetch with: u = phi / 2 pi, v = theta / pi.
BSH_HD void dir_to_uv(f3 d, float& u, float& v, float* theta_out = nullptr)
{
    float temp = atan2f(d.z, d.x);
    float r = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    float phi = temp >= 0.f ? temp : temp + 2.f * kPi;
    float theta = acosf(d.y / r);
    u = phi / (2.f * kPi);
    v = theta / kPi;
    if (theta_out) *theta_out = theta;
}

// scale * pow(texel, 1 / gamma) per channel (environment_light.cpp:43-47)
BSH_HD f3 env_radiance(const EnvDev& E, float scale, float invgamma, float u, float v)
{
    f3 val = env_bilinear(E.tex, E.w, E.h, u, v);
    val = mk(powf(val.x, invgamma), powf(val.y, invgamma), powf(val.z, invgamma));
    return scale * val;
}

// std::lower_bound over cdf[0, n]: the first index whose entry is not less than u.
BSH_HD int cdf_lower_bound(const float* cdf, int n, float u)
{
    int first = 0, len = n;
    while (len > 0)
    {
        int half = len >> 1;
        if (cdf[first + half] < u)
        {
            first += half + 1;
            len -= half + 1;
        }
        else
            len = half;
    }
    return first;
}

// Distribution1D::Sample1D (math/distribution1d.cpp:33-50) over n segments.
BSH_HD float dist1d_sample(const float* val, const float* cdf, float funcint, int n, float u, float& pdf)
{
    int seg = cdf_lower_bound(cdf, n + 1, u);
    seg = seg < 1 ? 1 : (seg > n + 1 ? n + 1 : seg);
    float du = (u - cdf[seg - 1]) / (cdf[seg] - cdf[seg - 1]);
    pdf = val[seg - 1] / funcint;
    return ((float)(seg - 1) + du) / (float)n;
}

// Distribution2D::Sample2D (math/distribution2d.cpp:29-45): row from the marginal by u.y,
// column from that row's conditional by u.x.
BSH_HD void dist2d_sample(const EnvDev& E, float ux, float uy, float& sx, float& sy, float& pdf)
{
    float rpdf, cpdf;
    sy = dist1d_sample(E.mval, E.mcdf, E.mint, kEnvDistH, uy, rpdf);
    int row = (int)(sy * (float)kEnvDistH);
    row = row < 0 ? 0 : (row > kEnvDistH - 1 ? kEnvDistH - 1 : row);
    sx = dist1d_sample(E.cval + (size_t)row * kEnvDistW, E.ccdf + (size_t)row * (kEnvDistW + 1), E.cint[row], kEnvDistW,
                       ux, cpdf);
    pdf = rpdf * cpdf;
}

// Distribution2D::Pdf (math/distribution2d.cpp:47-53)
BSH_HD float dist2d_pdf(const EnvDev& E, float u, float v)
{
    int row = (int)(v * (float)kEnvDistH), col = (int)(u * (float)kEnvDistW);
    row = row < 0 ? 0 : (row > kEnvDistH - 1 ? kEnvDistH - 1 : row);
    col = col < 0 ? 0 : (col > kEnvDistW - 1 ? kEnvDistW - 1 : col);
    if (E.cint[row] * E.mint == 0.f) return 0.f;
    return E.cval[(size_t)row * kEnvDistW + col] * E.mval[row] / (E.cint[row] * E.mint);
}

}  // namespace bsh
