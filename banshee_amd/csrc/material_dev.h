// The reference's materials beyond SimpleMaterial(Lambert | Microfacet) and Emissive:
// SimpleMaterial over OrenNayar, PerfectReflect or PerfectRefract, Glass and MixedMaterial
// (bsdf/orennayar.h, perfect_reflect.h, perfect_refract.h, material/glass.h,
// mixedmaterial.h), in the reference's float operations. Glass and MixedMaterial draw from
// the global rand() inside Sample (glass.h:62, mixedmaterial.h:60-65); here those draws come
// from the path's keyed rand() stream, advancing the path's running counter.
#pragma once

#include "bsh_math.h"
#include "light_dev.h"

namespace bsh {

enum : int
{
    MAT_LAMBERT = 0,
    MAT_MICROFACET = 1,
    MAT_EMISSIVE = 2,
    MAT_ORENNAYAR = 3,   // kd = c, kr = alpha
    MAT_REFLECT = 4,     // PerfectReflect(eta, ks = c, FresnelDielectric when x != 0)
    MAT_REFRACT = 5,     // PerfectRefract(eta, ks = c, FresnelDielectric when x != 0)
    MAT_GLASS = 6,       // Glass(eta, ks = c)
    MAT_MIXED = 7        // MixedMaterial(eta) over the component rows alpha, x, y (-1: none)
};
enum : int { BSDF_TRANSMISSION = 2, BSDF_SPECULAR = 8 };

// Texture maps of material rows (GET_VALUE, bsdf/bsdf.h:42): a row's colour map replaces its kd
// (Lambert, OrenNayar) or ks (Microfacet, PerfectReflect, PerfectRefract, Glass), its kr map
// OrenNayar's kr (the map's first channel); each read at the hit's (u, v) with the texture
// system's default filter, bilinear with periodic wrap (texture/texturesystem.h:35).
struct TexCtx
{
    const float* tex;   // all maps' RGB texels
    const int4* desc;   // per map: width, height, first float (low and high 32 bits)
    const int2* maps;   // per material row: colour map, kr map (-1: none); NULL when no row has one
    float u, v;
};

BSH_HD f3 tex_sample(const TexCtx& t, int id)
{
    const int4 d = t.desc[id];
    const size_t off = (size_t)(uint32_t)d.z | ((size_t)(uint32_t)d.w << 32);
    return env_bilinear(t.tex + off, d.x, d.y, t.u, t.v);
}

// The closest texel with periodic wrap (the texture system's kPoint filter): texel
// (floor(u w), floor(v h)).
BSH_HD f3 tex_point(const TexCtx& t, int id)
{
    const int4 d = t.desc[id];
    const size_t off = (size_t)(uint32_t)d.z | ((size_t)(uint32_t)d.w << 32);
    const int x = wrap_index((int)floorf(t.u * (float)d.x), d.x), y = wrap_index((int)floorf(t.v * (float)d.y), d.y);
    const float* p = t.tex + off + 3 * ((size_t)y * d.x + x);
    return mk(p[0], p[1], p[2]);
}

// Bsdf::MapNormal (bsdf/bsdf.h:97-113) applied to normal n: the map's closest texel as the
// tangent-space normal 2 texel - 1 in the frame (du, dv, n), du and dv Gram-Schmidt
// orthogonalised from the hit's dpdu and dpdv against n, in the reference's operation order.
BSH_HD f3 map_normal(const TexCtx& t, int id, const Frame& h, f3 n)
{
    const float ndotdu = dot(n, h.dpdu);
    const f3 du = normalize(h.dpdu - ndotdu * n);
    const float dudotdv = dot(du, h.dpdv);
    const f3 dv = normalize(h.dpdv - ndotdu * n - dudotdv * du);
    const f3 tn = normalize(2.f * tex_point(t, id) - mk(1.f, 1.f, 1.f));
    return normalize(n * tn.z + du * tn.x - dv * tn.y);
}

BSH_HD void apply_maps(const TexCtx& t, int row, Mat& m)
{
    if (!t.maps) return;
    const int2 mp = t.maps[row];
    if (mp.x >= 0) m.c = tex_sample(t, mp.x);
    if (mp.y >= 0) m.alpha = tex_sample(t, mp.y).x;
}

// The path's keyed rand() stream (glibc rand() and rand_float, math/mathutils.h:59-65).
struct RandStream
{
    uint32_t base, c;
    BSH_HD uint32_t next_uint() { return ks_rand(base, c++); }
    BSH_HD float next_float() { return ks_rand_float(base, c++); }
};

BSH_HD int bsdf_flags(int type)
{
    switch (type)
    {
    case MAT_LAMBERT:
    case MAT_ORENNAYAR: return BSDF_REFLECTION | BSDF_DIFFUSE;
    case MAT_MICROFACET: return BSDF_REFLECTION | BSDF_GLOSSY;
    case MAT_REFLECT: return BSDF_REFLECTION | BSDF_SPECULAR;
    case MAT_REFRACT: return BSDF_TRANSMISSION | BSDF_SPECULAR;
    default: return 0;
    }
}

// OrenNayar::Evaluate (bsdf/orennayar.h:90-156), kr = alpha
BSH_HD f3 orennayar_eval(const Mat& m, const Frame& h, f3 wi, f3 wo)
{
    float sameside = dot(wi, h.n) * dot(wo, h.n);
    if (sameside < 0.f) return mk(0.f, 0.f, 0.f);
    f3 n = h.n, s = h.dpdu;
    if (dot(wi, h.n) < 0.f)
    {
        n = -n;
        s = -s;
    }
    float kr = m.alpha;
    float invpi = 1.f / kPi;
    float r2 = kr * kr;
    float a = 1.f - 0.5f * r2 / (r2 + 0.33f);
    float b = 0.45f * r2 / (r2 + 0.09f);
    float cos_theta_i = dot(n, wi);
    float cos_theta_o = dot(n, wo);
    float sin_theta_i = sqrtf(1.f - cos_theta_i * cos_theta_i);
    float sin_theta_o = sqrtf(1.f - cos_theta_o * cos_theta_o);
    float cos_phi_i = dot(wi, s);
    float cos_phi_o = dot(wo, s);
    float sin_phi_i = sqrtf(1.f - cos_phi_i * cos_phi_i);
    float sin_phi_o = sqrtf(1.f - cos_phi_o * cos_phi_o);
    float maxcos = 0.f;
    if (sin_theta_i > 0.f && sin_theta_o > 0.f)
    {
        float c = cos_phi_i * cos_phi_o + sin_phi_i * sin_phi_o;
        maxcos = 0.f < c ? c : 0.f;   // std::max(0.f, c)
    }
    float sin_alpha, tan_beta;
    if (cos_theta_i > cos_theta_o)
    {
        sin_alpha = sin_theta_o;
        tan_beta = sin_theta_i / cos_theta_i;
    }
    else
    {
        sin_alpha = sin_theta_i;
        tan_beta = sin_theta_o / cos_theta_o;
    }
    return m.c * mk(invpi, invpi, invpi) * (a + b * maxcos * sin_alpha * tan_beta);
}

// OrenNayar::GetPdf (orennayar.h:158-173)
BSH_HD float orennayar_pdf(const Frame& h, f3 wi, f3 wo)
{
    float sameside = dot(wi, h.n) * dot(wo, h.n);
    if (sameside > 0.f)
    {
        f3 n = dot(wi, h.n) >= 0.f ? h.n : -h.n;
        return dot(n, wo) * (1.f / kPi);
    }
    return 0.f;
}

// Bsdf::Sample of the single BSDFs (component rows of a MixedMaterial and SimpleMaterials).
BSH_HD f3 bsdf_sample(const Mat& m, const Frame& h, float ux, float uy, f3 wi, f3& wo, float& pdf)
{
    if (m.type == MAT_ORENNAYAR)
    {
        // OrenNayar::Sample (orennayar.h:72-87)
        f3 n = dot(wi, h.n) >= 0.f ? h.n : -h.n;
        wo = map_to_hemisphere_cos(n, ux, uy);
        pdf = dot(n, wo) * (1.f / kPi);
        return orennayar_eval(m, h, wi, wo);
    }
    if (m.type == MAT_REFLECT)
    {
        // PerfectReflect::Sample (perfect_reflect.h:71-85)
        f3 n = dot(wi, h.n) >= 0.f ? h.n : -h.n;
        wo = normalize(2.f * dot(n, wi) * n - wi);
        pdf = 1.f;
        float reflectance = m.x != 0.f ? fresnel_dielectric(1.f, m.eta, dot(wi, h.n)) : 1.f;
        float ndotwi = dot(n, wi);
        return ndotwi > 1.1920929e-7f ? (reflectance * m.c * (1.f / ndotwi)) : mk(0.f, 0.f, 0.f);
    }
    if (m.type == MAT_REFRACT)
    {
        // PerfectRefract::Sample (perfect_refract.h:69-106)
        f3 n;
        float eta;
        float ndotwi = dot(wi, h.n);
        if (ndotwi >= 0.f)
        {
            n = h.n;
            eta = m.eta;
        }
        else
        {
            n = -h.n;
            eta = 1.f / m.eta;
            ndotwi = -ndotwi;
        }
        float reflectance = m.x != 0.f ? fresnel_dielectric(1.f, m.eta, dot(wi, h.n)) : 0.f;
        if (reflectance < 1.f)
        {
            float q0 = 1.f - (1.f - ndotwi * ndotwi) / (eta * eta);
            float q = 0.f < q0 ? q0 : 0.f;   // std::max(0.f, ...)
            wo = normalize(-(1.f / eta) * wi - (sqrtf(q) - ndotwi / eta) * n);
            pdf = 1.f;
            return ndotwi > 1.1920929e-7f ? ((1.f / (eta * eta)) * (1.f - reflectance) * m.c * (1.f / ndotwi))
                                           : mk(0.f, 0.f, 0.f);
        }
        pdf = 0.f;
        return mk(0.f, 0.f, 0.f);
    }
    int t = 0;
    return mat_sample(m, h, ux, uy, wi, wo, pdf, t);   // Lambert, Microfacet
}

BSH_HD f3 bsdf_eval(const Mat& m, const Frame& h, f3 wi, f3 wo)
{
    if (m.type == MAT_ORENNAYAR) return orennayar_eval(m, h, wi, wo);
    if (m.type == MAT_REFLECT || m.type == MAT_REFRACT) return mk(0.f, 0.f, 0.f);   // delta BSDFs
    return mat_eval(m, h, wi, wo);
}

BSH_HD float bsdf_pdf(const Mat& m, const Frame& h, f3 wi, f3 wo)
{
    if (m.type == MAT_ORENNAYAR) return orennayar_pdf(h, wi, wo);
    if (m.type == MAT_REFLECT || m.type == MAT_REFRACT) return 0.f;
    return mat_pdf(m, h, wi, wo);
}

// A material row of the scene's table (two float4: (type, c), (eta, alpha, x, y)).
BSH_HD Mat load_mat(const float4* mats, int idx)
{
    float4 a = mats[2 * idx], b = mats[2 * idx + 1];
    Mat m;
    m.type = (int)a.x;
    m.c = mk(a.y, a.z, a.w);
    m.eta = b.x;
    m.alpha = b.y;
    m.x = b.z;
    m.y = b.w;
    return m;
}

// The components of a MixedMaterial as AddBsdf files them (mixedmaterial.h:151-161): reflecting
// BSDFs in brdfs_, the others in btdfs_, each in the order added (rows alpha, x, y; -1: none).
BSH_HD void mixed_counts(const float4* mats, const Mat& m, int& nr, int& nt)
{
    nr = nt = 0;
    const float ids[3] = {m.alpha, m.x, m.y};
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        int id = (int)ids[k];
        if (id < 0) continue;
        if (bsdf_flags((int)mats[2 * id].x) & BSDF_REFLECTION) ++nr;
        else ++nt;
    }
}

// The idx-th reflecting (or transmitting) component.
BSH_HD Mat mixed_pick(const float4* mats, const TexCtx& tx, const Mat& m, bool reflecting, int idx)
{
    const float ids[3] = {m.alpha, m.x, m.y};
    int sel = -1, seen = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        int id = (int)ids[k];
        if (id < 0) continue;
        bool r = (bsdf_flags((int)mats[2 * id].x) & BSDF_REFLECTION) != 0;
        if (r == reflecting)
        {
            if (seen == idx) sel = id;
            ++seen;
        }
    }
    Mat c = load_mat(mats, sel < 0 ? 0 : sel);
    apply_maps(tx, sel < 0 ? 0 : sel, c);
    return c;
}

// Material::Sample of every material type; rs advances for Glass and MixedMaterial.
static __device__ __noinline__ f3 mat_sample_x(const float4* mats, const TexCtx& tx, const Mat& m, const Frame& h, float ux, float uy, f3 wi, f3& wo, float& pdf,
                       int& type, RandStream& rs)
{
    if (m.type <= MAT_EMISSIVE) return mat_sample(m, h, ux, uy, wi, wo, pdf, type);
    if (m.type == MAT_GLASS)
    {
        // Glass::Sample (glass.h:60-80) over PerfectReflect / PerfectRefract with FresnelDielectric
        float rnd = rs.next_float();
        float r = fresnel_dielectric(1.f, m.eta, dot(h.n, wi));
        Mat b = m;
        b.x = 1.f;
        if (rnd < r)
        {
            b.type = MAT_REFLECT;
            type = bsdf_flags(MAT_REFLECT);
            f3 f = bsdf_sample(b, h, ux, uy, wi, wo, pdf);
            pdf *= r;
            return f;
        }
        b.type = MAT_REFRACT;
        type = bsdf_flags(MAT_REFRACT);
        f3 f = bsdf_sample(b, h, ux, uy, wi, wo, pdf);
        pdf *= (1.f - r);
        return f;
    }
    if (m.type == MAT_MIXED)
    {
        // MixedMaterial::Sample (mixedmaterial.h:57-87)
        int nr, nt;
        mixed_counts(mats, m, nr, nt);
        float reflectance = fresnel_dielectric(1.f, m.eta, dot(h.n, wi));
        float rnd = rs.next_float();
        const bool refl = nt == 0 || rnd < reflectance;
        const int cnt = refl ? nr : nt;
        const int idx = (int)(rs.next_uint() % (uint32_t)cnt);
        const Mat c = mixed_pick(mats, tx, m, refl, idx);
        f3 f = bsdf_sample(c, h, ux, uy, wi, wo, pdf);
        type = bsdf_flags(c.type);
        pdf *= (1.f / (float)cnt);
        return f;
    }
    type = bsdf_flags(m.type);
    return bsdf_sample(m, h, ux, uy, wi, wo, pdf);
}

// Material::Evaluate of every material type (Glass: 0, glass.h:87-90).
static __device__ __noinline__ f3 mat_eval_x(const float4* mats, const TexCtx& tx, const Mat& m, const Frame& h, f3 wi, f3 wo)
{
    if (m.type <= MAT_EMISSIVE) return mat_eval(m, h, wi, wo);
    if (m.type == MAT_GLASS) return mk(0.f, 0.f, 0.f);
    if (m.type == MAT_MIXED)
    {
        // MixedMaterial::Evaluate (mixedmaterial.h:126-148): the sum over the side's components
        const bool refl = dot(h.n, wi) > 0.f;
        const float ids[3] = {m.alpha, m.x, m.y};
        f3 f = mk(0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            int id = (int)ids[k];
            if (id < 0) continue;
            Mat c = load_mat(mats, id);
            apply_maps(tx, id, c);
            if (((bsdf_flags(c.type) & BSDF_REFLECTION) != 0) == refl) f = f + bsdf_eval(c, h, wi, wo);
        }
        return f;
    }
    return bsdf_eval(m, h, wi, wo);
}

// Material::GetPdf of every material type (Glass: 0, glass.h:82-85).
static __device__ __noinline__ float mat_pdf_x(const float4* mats, const TexCtx& tx, const Mat& m, const Frame& h, f3 wi, f3 wo)
{
    if (m.type <= MAT_EMISSIVE) return mat_pdf(m, h, wi, wo);
    if (m.type == MAT_GLASS) return 0.f;
    if (m.type == MAT_MIXED)
    {
        // MixedMaterial::GetPdf (mixedmaterial.h:97-124): the mean over the side's components
        const bool refl = dot(h.n, wi) > 0.f;
        const float ids[3] = {m.alpha, m.x, m.y};
        float pdf = 0.f;
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            int id = (int)ids[k];
            if (id < 0) continue;
            Mat c = load_mat(mats, id);
            apply_maps(tx, id, c);
            if (((bsdf_flags(c.type) & BSDF_REFLECTION) != 0) == refl)
            {
                pdf += bsdf_pdf(c, h, wi, wo);
                ++cnt;
            }
        }
        pdf /= (float)cnt;
        return pdf;
    }
    return bsdf_pdf(m, h, wi, wo);
}

}  // namespace bsh
