// Device and host math for the GI hot path.
//
// Every routine evaluates the same IEEE-754 binary32 operations, in the same order,
// as the reference code it cites (paths relative to FireRays/Banshee/), so that the
// GPU reproduces the reference's ray/triangle results bit for bit and its shading to
// within the ulp differences of sin/cos/atan. The library is compiled with
// -ffp-contract=off; fused multiply-adds appear only where written explicitly
// (ray/box slabs, which need not match bit for bit because they are conservative).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define BSH_HD __host__ __device__ __forceinline__

namespace bsh {

constexpr float kPi = 3.14159265358979323846f;   // PI, math/mathutils.h:50
constexpr float kMinPdf = 0.05f;                 // MINPDF, tracer/gitracer.cpp:10, ditracer.cpp:17
constexpr float kMaxRadiance = 4.f;              // MAXRADIANCE, tracer/gitracer.cpp:11

struct f3
{
    float x, y, z;
};

BSH_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
BSH_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
BSH_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
BSH_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
BSH_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
BSH_HD f3 operator*(f3 a, float c) { return f3{a.x * c, a.y * c, a.z * c}; }
BSH_HD f3 operator*(float c, f3 a) { return f3{a.x * c, a.y * c, a.z * c}; }
// float3::sqnorm / dot, math/float3.h:50,96-99: ((x*x + y*y) + z*z)
BSH_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BSH_HD float sqnorm(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
// cross, math/float3.h:108-114
BSH_HD f3 cross(f3 a, f3 b)
{
    return f3{a.y * b.z - b.y * a.z, b.x * a.z - a.x * b.z, a.x * b.y - a.y * b.x};
}
// normalize: if sqnorm > 0, v *= 1/sqrt(sqnorm), math/float3.h:51,58,101-106
BSH_HD f3 normalize(f3 v)
{
    float s = sqnorm(v);
    if (s > 0.f)
    {
        float inv = 1.f / sqrtf(s);
        return f3{v.x * inv, v.y * inv, v.z * inv};
    }
    return v;
}
BSH_HD float clampf(float x, float a, float b) { return x < a ? a : (x > b ? b : x); }
BSH_HD f3 ld3(float4 v) { return f3{v.x, v.y, v.z}; }
BSH_HD float4 st3(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// orthovector, math/mathutils.h:216-228 (the double-precision sqrt there rounds to
// the same float as sqrtf: double rounding of a float square root is exact).
BSH_HD f3 orthovector(f3 n)
{
    f3 p;
    if (fabsf(n.z) > 0.707106781186547524401f)
    {
        float k = sqrtf(n.y * n.y + n.z * n.z);
        p.x = 0.f;
        p.y = -n.z / k;
        p.z = n.y / k;
    }
    else
    {
        float k = sqrtf(n.x * n.x + n.y * n.y);
        p.x = -n.y / k;
        p.y = n.x / k;
        p.z = 0.f;
    }
    return p;
}

// map_to_hemisphere(n, s, e=1), math/mathutils.h:231-244. powf(x, 1/(e+1)) with e = 1
// is powf(x, 0.5f), whose correctly rounded value equals sqrtf(x).
BSH_HD f3 map_to_hemisphere_cos(f3 n, float sx, float sy)
{
    f3 u = orthovector(n);
    f3 v = cross(u, n);
    u = cross(n, v);
    float a = 2.f * kPi * sx;
    float sinpsi = sinf(a);
    float cospsi = cosf(a);
    float costheta = sqrtf(1.f - sy);
    float sintheta = sqrtf(1.f - costheta * costheta);
    return normalize(u * sintheta * cospsi + v * sintheta * sinpsi + n * costheta);
}

// PowerHeuristic(1, f, 1, g), math/mathutils.h:302-307
BSH_HD float power_heuristic(float f, float g) { return (f * f) / (f * f + g * g); }

// ---------------------------------------------------------------------------------
// Keyed sample stream. Every random number of camera sample g is
//   mix(mix(mix(seed + golden) ^ g) ^ (stream << 24 | counter))
// with mix = lowbias32 (C. Wellons, "Prospecting for Hash Functions", 2018), and the
// uint -> float conversion of McRng::NextFloat (rng/mcrng.h:54-57).
// ---------------------------------------------------------------------------------
enum : uint32_t { KS_IMG = 0, KS_LIGHT = 1, KS_BRDF = 2, KS_RAND = 3, KS_PROBE = 4, KS_GUIDE = 5 };

BSH_HD uint32_t ks_mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}
BSH_HD uint32_t ks_base(uint32_t seed, uint32_t camsample) { return ks_mix32(ks_mix32(seed + 0x9e3779b9U) ^ camsample); }
BSH_HD uint32_t ks_u32(uint32_t base, uint32_t stream, uint32_t counter) { return ks_mix32(base ^ ((stream << 24) | counter)); }
BSH_HD float ks_float(uint32_t u) { return (float)u * 2.328306e-10f; }
// Sample2D from draws (c, c+1): g++ evaluates float2(NextFloat(), NextFloat()) right to
// left (sampler/stratified_sampler.cpp:10), so the first draw lands in .y.
BSH_HD void ks_sample2d(uint32_t base, uint32_t stream, uint32_t c, float& x, float& y)
{
    y = ks_float(ks_u32(base, stream, c));
    x = ks_float(ks_u32(base, stream, c + 1));
}
// glibc rand() stand-in (RAND_MAX = 2^31-1): used by GiTracer for the light pick and
// Russian roulette (gitracer.cpp:75,119); rand_float = (float)rand()/RAND_MAX.
BSH_HD uint32_t ks_rand(uint32_t base, uint32_t counter) { return ks_u32(base, KS_RAND, counter) >> 1; }
BSH_HD float ks_rand_float(uint32_t base, uint32_t counter) { return (float)(int)ks_rand(base, counter) / 2147483648.f; }

// ---------------------------------------------------------------------------------
// Surface frame at a hit (the parts of ShapeBundle::Hit the shading code reads).
// ---------------------------------------------------------------------------------
struct Frame
{
    f3 p, n, ng, dpdu, dpdv;
};

// Mesh::FillHit, primitive/mesh.cpp:47-97, for identity transforms and no UVs
// (uvs_ = {(0,0)} so det == 0 and the orthovector frame is used).
BSH_HD Frame fill_hit(f3 p1, f3 p2, f3 p3, f3 n1, f3 n2, f3 n3, float a, float b)
{
    Frame h;
    float w = 1.f - a - b;
    h.p = w * p1 + a * p2 + b * p3;
    h.n = normalize(w * n1 + a * n2 + b * n3);
    h.ng = normalize(cross(p3 - p1, p2 - p1));
    if (dot(h.n, h.ng) < 0.f) h.ng = -h.ng;
    h.dpdu = orthovector(h.n);
    h.dpdv = cross(h.n, h.dpdu);
    return h;
}

// Mesh::FillHit with texture coordinates t1..t3 (mesh.cpp:72-94, identity transform): the
// frame from the UV derivatives when their determinant is nonzero, else fill_hit's; writes the
// interpolated (u, v).
BSH_HD Frame fill_hit_uv(f3 p1, f3 p2, f3 p3, f3 n1, f3 n2, f3 n3, float a, float b, float u1, float v1, float u2,
                         float v2, float u3, float v3, float& u, float& v)
{
    Frame h = fill_hit(p1, p2, p3, n1, n2, n3, a, b);
    float du1 = u1 - u3, du2 = u2 - u3, dv1 = v1 - v3, dv2 = v2 - v3;
    f3 dp1 = p1 - p3, dp2 = p2 - p3;
    float det = du1 * dv2 - dv1 * du2;
    if (det != 0.f)
    {
        float invdet = 1.f / det;
        h.dpdu = normalize((dv2 * dp1 - dv1 * dp2) * invdet);
        h.dpdv = normalize((-du2 * dp1 + du1 * dp2) * invdet);
    }
    float w = 1.f - a - b;
    u = w * u1 + a * u2 + b * u3;
    v = w * v1 + a * v2 + b * v3;
    return h;
}

// ---------------------------------------------------------------------------------
// Materials. type 0 Lambert, 1 Microfacet+GGX+FresnelDielectric, 2 Emissive.
// Argument names follow the reference: Sample(hit, u, wi, out wo) samples wo given wi;
// Evaluate/GetPdf(hit, wi, wo).
// ---------------------------------------------------------------------------------
struct Mat
{
    int type;
    f3 c;
    float eta, alpha;
    float x, y;   // material rows 4-7 (material_dev.h): Fresnel flag, MixedMaterial components
};

enum : int { BSDF_REFLECTION = 1, BSDF_DIFFUSE = 4, BSDF_GLOSSY = 16 };

// GgxDistribution::D, bsdf/microfacet_distribution.h:205-220
BSH_HD float ggx_D(float a, f3 m, f3 n)
{
    float ndotm = dot(m, n);
    if (ndotm <= 0.f) return 0.f;
    float ndotm2 = ndotm * ndotm;
    float sinmn = sqrtf(1.f - clampf(ndotm * ndotm, 0.f, 1.f));
    float tanmn = sinmn / ndotm;
    float a2 = a * a;
    return a2 / (kPi * ndotm2 * ndotm2 * (a2 + tanmn * tanmn) * (a2 + tanmn * tanmn));
}
// GgxDistribution::GetPdf, microfacet_distribution.h:247-257
BSH_HD float ggx_pdf(float a, f3 n, f3 wi, f3 wo)
{
    f3 m = normalize(wi + wo);
    float mpdf = ggx_D(a, m, n) * fabsf(dot(n, m));
    return mpdf / (4.f * dot(wo, m));
}
// GgxDistribution::G1 / G, microfacet_distribution.h:259-283
BSH_HD float ggx_G1(float a, f3 v, f3 m, f3 n)
{
    float ndotv = dot(n, v);
    float mdotv = dot(m, v);
    if (ndotv * mdotv <= 0.f) return 0.f;
    float sinnv = sqrtf(1.f - clampf(ndotv * ndotv, 0.f, 1.f));
    float tannv = sinnv / ndotv;
    float a2 = a * a;
    return 2.f / (1.f + sqrtf(1.f + a2 * tannv * tannv));
}
// FresnelDielectric::Evaluate, bsdf/fresnel.h:59-90
BSH_HD float fresnel_dielectric(float etai, float etat, float ndotwi)
{
    float ei = etai, et = etat;
    if (ndotwi < 0.f)
    {
        ei = etat;
        et = etai;
        ndotwi = -ndotwi;
    }
    float sint = (ei / et) * sqrtf(fmaxf(0.f, 1.f - ndotwi * ndotwi));
    if (sint > 1.f) return 1.f;
    float ndotwt = sqrtf(fmaxf(0.f, 1.f - sint * sint));
    float rparl = ((et * ndotwi) - (ei * ndotwt)) / ((et * ndotwi) + (ei * ndotwt));
    float rperp = ((ei * ndotwi) - (et * ndotwt)) / ((ei * ndotwi) + (et * ndotwt));
    return (rparl * rparl + rperp * rperp) * 0.5f;
}

// Emissive::GetLe, material/emissive.h:84-95
BSH_HD f3 emissive_le(f3 e, f3 n, f3 wo)
{
    float ndotwo = dot(n, wo);
    if (ndotwo > 0.f) return e * ndotwo;
    return f3{0.f, 0.f, 0.f};
}

// Microfacet::Evaluate, bsdf/microfacet.h:100-131
BSH_HD f3 microfacet_eval(const Mat& m, const Frame& h, f3 wi, f3 wo)
{
    float sameside = dot(wi, h.n) * dot(wo, h.n);
    if (sameside < 0.f) return f3{0.f, 0.f, 0.f};
    f3 n = dot(h.n, wi) < 0.f ? -h.n : h.n;
    float cos_theta_o = dot(n, wo);
    float cos_theta_i = dot(n, wi);
    if (cos_theta_i == 0.f || cos_theta_o == 0.f) return f3{0.f, 0.f, 0.f};
    f3 wh = normalize(wi + wo);
    float fr = fresnel_dielectric(1.f, m.eta, dot(wi, wh));
    float G = ggx_G1(m.alpha, wi, wh, n) * ggx_G1(m.alpha, wo, wh, n);
    return m.c * (ggx_D(m.alpha, wh, n) * G * fr / (4.f * cos_theta_i * cos_theta_o));
}

// Material::Evaluate (SimpleMaterial forwards, simplematerial.h:66-69)
BSH_HD f3 mat_eval(const Mat& m, const Frame& h, f3 wi, f3 wo)
{
    if (m.type == 0)
    {
        // Lambert::Evaluate, lambert.h:88-106
        float sameside = dot(wi, h.ng) * dot(wo, h.ng);
        if (sameside > 0.f) return (1.f / kPi) * m.c;
        return f3{0.f, 0.f, 0.f};
    }
    if (m.type == 1) return microfacet_eval(m, h, wi, wo);
    return emissive_le(m.c, h.n, wi);   // Emissive::Evaluate returns GetLe(hit, wi), emissive.h:64-69
}

// Material::GetPdf
BSH_HD float mat_pdf(const Mat& m, const Frame& h, f3 wi, f3 wo)
{
    if (m.type == 0)
    {
        // Lambert::GetPdf, lambert.h:108-126
        float sameside = dot(wi, h.ng) * dot(wo, h.ng);
        if (sameside > 0.f)
        {
            f3 n = dot(wi, h.ng) >= 0.f ? h.n : -h.n;
            return dot(n, wo) * (1.f / kPi);
        }
        return 0.f;
    }
    if (m.type == 1)
    {
        // Microfacet::GetPdf, microfacet.h:134-150 (dpdu/dpdv flips do not affect the pdf)
        float sameside = dot(wi, h.n) * dot(wo, h.n);
        if (sameside < 0.f) return 0.f;
        f3 n = dot(h.n, wi) < 0.f ? -h.n : h.n;
        return ggx_pdf(m.alpha, n, wi, wo);
    }
    return 0.f;   // Emissive::GetPdf, emissive.h:72-75
}

// Material::Sample: returns f, writes wo, pdf and the BSDF type flags.
BSH_HD f3 mat_sample(const Mat& m, const Frame& h, float ux, float uy, f3 wi, f3& wo, float& pdf, int& type)
{
    if (m.type == 0)
    {
        // Lambert::Sample, lambert.h:66-85
        type = BSDF_REFLECTION | BSDF_DIFFUSE;
        f3 n = dot(wi, h.ng) >= 0.f ? h.n : -h.n;
        wo = map_to_hemisphere_cos(n, ux, uy);
        float invpi = 1.f / kPi;
        pdf = dot(n, wo) * invpi;
        return invpi * m.c;
    }
    if (m.type == 1)
    {
        // Microfacet::Sample, microfacet.h:80-97 -> GgxDistribution::Sample,
        // microfacet_distribution.h:223-245
        type = BSDF_REFLECTION | BSDF_GLOSSY;
        f3 n = h.n, du = h.dpdu, dv = h.dpdv;
        if (dot(h.n, wi) < 0.f)
        {
            n = -n;
            du = -du;
            dv = -dv;
        }
        float temp = atanf(m.alpha * sqrtf(ux) / sqrtf(1.f - ux));
        float theta = (temp >= 0.f) ? temp : (temp + 2.f * kPi);
        float costheta = cosf(theta);
        float sintheta = sqrtf(1.f - clampf(costheta * costheta, 0.f, 1.f));
        float phi = 2.f * kPi * uy;
        float cosphi = cosf(phi);
        float sinphi = sinf(phi);
        f3 wh = normalize(du * sintheta * cosphi + dv * sintheta * sinphi + n * costheta);
        wo = -wi + 2.f * dot(wi, wh) * wh;
        pdf = ggx_pdf(m.alpha, n, wi, wo);
        return microfacet_eval(m, h, wi, wo);
    }
    // Emissive::Sample, emissive.h:55-61: wo is left as the caller's zero float3
    type = 0;
    wo = f3{0.f, 0.f, 0.f};
    pdf = 0.f;
    return emissive_le(m.c, h.n, wi);
}

// Möller-Trumbore exactly as Mesh::IntersectFace, primitive/mesh.cpp:7-45, with e1 = p2-p1
// and e2 = p3-p1 precomputed (the same float subtractions).
BSH_HD bool intersect_face(f3 p1, f3 e1, f3 e2, f3 o, f3 d, float tmin, float tmax, float& t, float& b1o,
                           float& b2o)
{
    f3 s1 = cross(d, e2);
    float invd = 1.f / dot(s1, e1);
    f3 dd = o - p1;
    float b1 = dot(dd, s1) * invd;
    if (b1 < 0.f || b1 > 1.f) return false;
    f3 s2 = cross(dd, e1);
    float b2 = dot(d, s2) * invd;
    if (b2 < 0.f || b1 + b2 > 1.f) return false;
    float tmp = dot(e2, s2) * invd;
    if (tmp > tmin && tmp <= tmax)
    {
        t = tmp;
        b1o = b1;
        b2o = b2;
        return true;
    }
    return false;
}

}  // namespace bsh
