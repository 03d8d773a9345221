/*
 * TEST INFRASTRUCTURE ONLY. This file is compiled together with the
 * reference's own sources (FireRays/Banshee, read in place) into
 * oracle/_ref/liboracle_ref.so by oracle/Makefile. It never ships in the
 * product and the product never links or loads it; only tests/, the
 * fixture generator under tests/golden/ and bench.py's CPU-baseline leg do.
 *
 * It adds, without editing any reference source:
 *   - NullTextureSystem: the reference's TextureSystem is abstract
 *     (Banshee/texture/texturesystem.h:56-59); constant-colour BSDFs never
 *     call it.
 *   - RecordingMesh: a Mesh subclass that records the winning global
 *     primitive index and barycentrics of a closest hit (ShapeBundle::Hit has
 *     no primitive index, Banshee/primitive/shapebundle.h:156-176).
 *   - RecordingAccel: an Intersectable decorator installed over World::accel_
 *     (Banshee/world/world.h:80) that counts and optionally records queries.
 *   - KeyedRng: an Rng (Banshee/rng/rng.h:39-53) drawing from the keyed
 *     sample stream, plus link-time wrapping of rand()/srand() so GiTracer's
 *     light pick and Russian roulette (gitracer.cpp:75,119) draw from it too.
 *   - MapTextureSystem: in-memory float maps served by name. The reference's
 *     only TextureSystem reads textures through OpenImageIO 1.6
 *     (texture/oiio_texturesystem.cpp), which is absent here; this restates
 *     the two filters and wrap modes it maps to: InterpClosest (kPoint) reads
 *     texel (floor(u w), floor(v h)); InterpBilinear (kBilinear) blends the
 *     four texels around (u w - 1/2, v h - 1/2), texel centres at half-integer
 *     coordinates; WrapPeriodic (kRepeat) wraps texel indices, WrapMirror
 *     (kMirror) reflects them.
 *   - KeyedCamera: a PerscpectiveCamera whose GenerateRay (a virtual of
 *     Banshee/camera/camera.h:53) first moves the keyed stream to the next
 *     camera sample, so the light, BSDF and rand() draws of a camera sample are
 *     keyed by it whatever the image sampler draws.
 */
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "world/world.h"
#include "primitive/mesh.h"
#include "accelerator/bvh.h"
#include "material/simplematerial.h"
#include "material/emissive.h"
#include "bsdf/lambert.h"
#include "bsdf/microfacet.h"
#include "bsdf/fresnel.h"
#include "bsdf/orennayar.h"
#include "bsdf/perfect_reflect.h"
#include "bsdf/perfect_refract.h"
#include "bsdf/normal_mapping.h"
#include "material/glass.h"
#include "material/mixedmaterial.h"
#include "light/arealight.h"
#include "light/pointlight.h"
#include "light/directional_light.h"
#include "light/environment_light.h"
#include "light/environment_light_is.h"
#include "math/distribution1d.h"
#include "math/distribution2d.h"
#include "camera/perspective_camera.h"
#include "tracer/gitracer.h"
#include "sampler/stratified_sampler.h"
#include "sampler/regular_sampler.h"
#include "sampler/random_sampler.h"
#include "sampler/cmj_sampler.h"
#include "sampler/sobol_sampler.h"
#include "rng/mcrng.h"
#include "renderer/imagerenderer.h"
#include "renderer/mt_imagerenderer.h"
#include "util/progressreporter.h"
#include "imageplane/imageplane.h"
#include "math/sh.h"
#include "math/shproject.h"
#include "math/mathutils.h"

#include "../keyed_stream.h"

/* Defined with external linkage in Banshee/math/shproject.cpp:73 but not declared in its header. */
void ShProjectEnvironmentMap(float3 const* envmap, int width, int height, int lmax, float3* coeffs);

extern "C" int __real_rand(void);
extern "C" void __real_srand(unsigned);

namespace {

class NullTextureSystem : public TextureSystem
{
public:
    float3 Sample(std::string const&, float2 const&, float2 const&, Options const&) const { return float3(0, 0, 0); }
    void GetTextureInfo(std::string const&, TextureDesc& d) const { d.width = d.height = 0; }
};

NullTextureSystem g_texsys;

class MapTextureSystem : public TextureSystem
{
public:
    struct Map
    {
        int w = 0, h = 0;
        std::vector<float> t;   // h rows of w RGB texels
    };
    std::map<std::string, Map> maps;

    static int wrap(int i, int n, bool mirror)
    {
        if (!mirror) return ((i % n) + n) % n;
        int k = ((i % (2 * n)) + 2 * n) % (2 * n);
        return k < n ? k : 2 * n - 1 - k;
    }
    static float3 texel(Map const& m, int x, int y)
    {
        float const* p = &m.t[3 * ((size_t)y * m.w + x)];
        return float3(p[0], p[1], p[2]);
    }
    float3 Sample(std::string const& name, float2 const& uv, float2 const&, Options const& o) const
    {
        auto it = maps.find(name);
        if (it == maps.end()) return float3(0, 0, 0);
        Map const& m = it->second;
        bool mirror = o.wrapmode == Options::kMirror;
        if (o.filter == Options::kPoint)
            return texel(m, wrap((int)std::floor(uv.x * (float)m.w), m.w, mirror),
                         wrap((int)std::floor(uv.y * (float)m.h), m.h, mirror));
        float s = uv.x * (float)m.w - 0.5f, t = uv.y * (float)m.h - 0.5f;
        float sf = std::floor(s), tf = std::floor(t);
        float fs = s - sf, ft = t - tf;
        int x0 = wrap((int)sf, m.w, mirror), x1 = wrap((int)sf + 1, m.w, mirror);
        int y0 = wrap((int)tf, m.h, mirror), y1 = wrap((int)tf + 1, m.h, mirror);
        float3 a = texel(m, x0, y0), b = texel(m, x1, y0), c = texel(m, x0, y1), d = texel(m, x1, y1);
        return (1.f - ft) * ((1.f - fs) * a + fs * b) + ft * ((1.f - fs) * c + fs * d);
    }
    void GetTextureInfo(std::string const& name, TextureDesc& d) const
    {
        auto it = maps.find(name);
        d.width = it == maps.end() ? 0 : it->second.w;
        d.height = it == maps.end() ? 0 : it->second.h;
    }
};

/* ---------------- keyed stream context (single-threaded use) ---------------- */
struct KeyCtx
{
    bool     on = false;
    uint32_t seed = 0;
    uint64_t img_calls = 0;
    uint32_t base = 0;
    uint32_t camsample = 0;
    uint32_t ctr[8] = {0};
    /* tile mapping: sequential camera sample q -> global index */
    int x0 = 0, y0 = 0, tw = 1, W = 1, spp = 1;
    /* image draws: img_draws == 0 -> two per camera sample (stratified, random), keyed by
       that sample; img_draws = d > 0 -> d per pixel (CMJ 1, Sobol 2), keyed by the pixel's
       sample 0. cam_keyed: KeyedCamera, not the image draws, moves the other streams to
       the next camera sample. stride: counter step per draw of each stream (2 for a
       CmjSampler(1), whose one draw per sample takes the first of the sample's two counters). */
    int img_draws = 0;
    bool cam_keyed = false;
    uint64_t cam_calls = 0;
    uint32_t stride[8] = {1, 1, 1, 1, 1, 1, 1, 1};
};
KeyCtx g_key;

/* Global camera-sample index of the q-th camera sample in the tile's loop order
   (y outer, x inner, s innermost). */
uint32_t key_camsample(uint64_t q)
{
    uint64_t pix = q / (uint64_t)g_key.spp;
    uint64_t s = q % (uint64_t)g_key.spp;
    uint64_t x = g_key.x0 + pix % (uint64_t)g_key.tw;
    uint64_t y = g_key.y0 + pix / (uint64_t)g_key.tw;
    return (uint32_t)((y * (uint64_t)g_key.W + x) * (uint64_t)g_key.spp + s);
}

void key_enter_sample(uint32_t camsample)
{
    g_key.camsample = camsample;
    g_key.base = ks_base(g_key.seed, camsample);
    for (int i = 0; i < 8; ++i) g_key.ctr[i] = 0;
}

class KeyedRng : public Rng
{
public:
    explicit KeyedRng(uint32_t stream) : stream_(stream) {}
    unsigned NextUint() const
    {
        if (stream_ == KS_IMG)
        {
            uint64_t k = g_key.img_calls++;
            if (g_key.img_draws > 0)
            {
                uint64_t d = (uint64_t)g_key.img_draws;
                uint32_t cs0 = key_camsample((k / d) * (uint64_t)g_key.spp);
                return ks_u32(ks_base(g_key.seed, cs0), KS_IMG, (uint32_t)(k % d));
            }
            uint32_t cs = key_camsample(k >> 1);
            if ((k & 1) == 0 && !g_key.cam_keyed) key_enter_sample(cs);
            return ks_u32(ks_base(g_key.seed, cs), KS_IMG, (uint32_t)(k & 1));
        }
        uint32_t c = g_key.ctr[stream_];
        g_key.ctr[stream_] += g_key.stride[stream_];
        return ks_u32(g_key.base, stream_, c);
    }
    float NextFloat() const { return ks_float(NextUint()); }
    Rng* Clone() const { return new KeyedRng(stream_); }
private:
    uint32_t stream_;
};

/* The reference's perspective camera, keying the stream per camera sample when asked. */
class KeyedCamera : public PerscpectiveCamera
{
public:
    using PerscpectiveCamera::PerscpectiveCamera;
    void GenerateRay(float2 const& sample, ray& r) const
    {
        if (g_key.on && g_key.cam_keyed) key_enter_sample(key_camsample(g_key.cam_calls++));
        PerscpectiveCamera::GenerateRay(sample, r);
    }
};

/* Probe RNG used to detect which Sample2D component takes the first draw. */
class SeqRng : public Rng
{
public:
    unsigned NextUint() const { return n_++; }
    float NextFloat() const { return (n_++ == 0) ? 0.25f : 0.75f; }
    Rng* Clone() const { return new SeqRng(); }
private:
    mutable unsigned n_ = 0;
};

/* ---------------- recording of closest hits ---------------- */
struct HitRecord { int prim; float t, a, b; };
thread_local HitRecord tl_rec;

class RecordingMesh : public Mesh
{
public:
    RecordingMesh(float const* v, int nv, float const* n, int nn, int const* vi, int const* ni,
                  int const* mi, int nf, int global_start, float const* uv = nullptr, int nuv = 0,
                  int const* ui = nullptr)
        : Mesh(v, nv, 0, n, nn, 0, uv, nuv, 0, vi, 0, ni, 0, ui, 0, mi, (int)sizeof(int), nf)
        , start_(global_start)
    {
    }
    bool IntersectShape(std::size_t idx, ray const& r, Hit& hit) const
    {
        if (!Mesh::IntersectShape(idx, r, hit)) return false;
        matrix m, minv;
        GetTransform(m, minv);
        ray ro = transform_ray(r, minv);
        float t, a, b;
        if (IntersectFace(GetFaces()[idx], ro, hit.t, t, a, b))
        {
            tl_rec.prim = start_ + (int)idx;
            tl_rec.t = t;
            tl_rec.a = a;
            tl_rec.b = b;
        }
        return true;
    }
    bool IntersectShape(std::size_t idx, ray const& r) const { return Mesh::IntersectShape(idx, r); }
    int start_;
};

/* ---------------- counting / recording accelerator ---------------- */
struct alignas(64) Slot { uint64_t closest = 0, any = 0; };
Slot g_slots[256];
std::atomic<int> g_slot_next(0);
thread_local int tl_slot = -1;
inline Slot& my_slot()
{
    if (tl_slot < 0) tl_slot = g_slot_next.fetch_add(1) & 255;
    return g_slots[tl_slot];
}

struct RayRec { float r[8]; int kind; int prim; float t, a, b; };

class RecordingAccel : public Intersectable
{
public:
    explicit RecordingAccel(std::unique_ptr<Intersectable> inner) : inner_(std::move(inner)) {}
    bool Intersect(ray const& r, ShapeBundle::Hit& hit) const
    {
        my_slot().closest++;
        if (!recording_) return inner_->Intersect(r, hit);
        tl_rec.prim = -1;
        bool h = inner_->Intersect(r, hit);
        Record(r, 0, h ? tl_rec.prim : -1, h ? tl_rec.t : 0.f, h ? tl_rec.a : 0.f, h ? tl_rec.b : 0.f);
        return h;
    }
    bool Intersect(ray const& r) const
    {
        my_slot().any++;
        bool h = inner_->Intersect(r);
        if (recording_) Record(r, 1, h ? 1 : 0, 0.f, 0.f, 0.f);
        return h;
    }
    void Record(ray const& r, int kind, int prim, float t, float a, float b) const
    {
        uint64_t k = seen_++;
        if (stride_ <= 0 || (k % (uint64_t)stride_) != 0) return;
        if ((int)recs_.size() >= max_) return;
        RayRec rr;
        rr.r[0] = r.o.x; rr.r[1] = r.o.y; rr.r[2] = r.o.z; rr.r[3] = r.t.x;
        rr.r[4] = r.d.x; rr.r[5] = r.d.y; rr.r[6] = r.d.z; rr.r[7] = r.t.y;
        rr.kind = kind; rr.prim = prim; rr.t = t; rr.a = a; rr.b = b;
        recs_.push_back(rr);
    }
    std::unique_ptr<Intersectable> inner_;
    mutable bool recording_ = false;
    mutable int stride_ = 1;
    mutable int max_ = 0;
    mutable uint64_t seen_ = 0;
    mutable std::vector<RayRec> recs_;
};

void reset_counts()
{
    for (auto& s : g_slots) { s.closest = 0; s.any = 0; }
}
void read_counts(long long* out)
{
    uint64_t c = 0, a = 0;
    for (auto& s : g_slots) { c += s.closest; a += s.any; }
    if (out) { out[0] = (long long)c; out[1] = (long long)a; }
}

/* ---------------- image plane keeping sum, sum of squares, count ---------------- */
class StatsPlane : public ImagePlane
{
public:
    StatsPlane(int W, int H, int x0, int y0, int tw, int th, int spp, bool per_sample)
        : ImagePlane(int2(W, H), nullptr), x0_(x0), y0_(y0), tw_(tw), th_(th), spp_(spp)
        , sum_(3 * tw * th, 0.f), sq_(3 * tw * th, 0.f), n_(tw * th, 0.f)
    {
        if (per_sample) samples_.reserve((size_t)3 * tw * th * spp);
        keep_ = per_sample;
    }
    void WriteSample(int2 const& pos, float3 const& v)
    {
        int lx = pos.x - x0_, ly = pos.y - y0_;
        if (lx < 0 || ly < 0 || lx >= tw_ || ly >= th_) return;
        int i = ly * tw_ + lx;
        sum_[3 * i] += v.x; sum_[3 * i + 1] += v.y; sum_[3 * i + 2] += v.z;
        sq_[3 * i] += v.x * v.x; sq_[3 * i + 1] += v.y * v.y; sq_[3 * i + 2] += v.z * v.z;
        n_[i] += 1.f;
        if (keep_) { samples_.push_back(v.x); samples_.push_back(v.y); samples_.push_back(v.z); }
    }
    int x0_, y0_, tw_, th_, spp_;
    bool keep_;
    std::vector<float> sum_, sq_, n_, samples_;
};

struct Scene
{
    MapTextureSystem texsys;   // the environment lights' map, named "env"
    World world;
    std::vector<RecordingMesh*> meshes;
    RecordingAccel* accel = nullptr;
};

/* The texture map names of material row i: maps[2 i] (colour) and maps[2 i + 1] (kr) name
   "map<k>" of the scene's texture system, or nothing. */
struct RowMaps
{
    std::string color, kr;
};
RowMaps row_maps(int const* maps, int i)
{
    RowMaps r;
    if (maps && maps[2 * i] >= 0) r.color = "map" + std::to_string(maps[2 * i]);
    if (maps && maps[2 * i + 1] >= 0) r.kr = "map" + std::to_string(maps[2 * i + 1]);
    return r;
}

/* The single BSDF of material rows 0, 1, 3, 4 and 5 (also a MixedMaterial's components), with
   its colour (kd / ks) and kr maps. */
Bsdf* make_bsdf(TextureSystem const& ts, float const* rows, int i, int const* maps)
{
    float const* m = rows + 8 * i;
    int type = (int)m[0];
    float3 c(m[1], m[2], m[3]);
    RowMaps r = row_maps(maps, i);
    if (type == 0) return new Lambert(ts, c, r.color);
    if (type == 1) return new Microfacet(ts, m[4], c, r.color, "", new FresnelDielectric(), new GgxDistribution(m[5]));
    if (type == 3) return new OrenNayar(ts, c, m[5], r.color, r.kr);
    if (type == 4) return new PerfectReflect(ts, m[4], c, r.color, "", m[6] != 0.f ? new FresnelDielectric() : nullptr);
    return new PerfectRefract(ts, m[4], c, r.color, "", m[6] != 0.f ? new FresnelDielectric() : nullptr);
}

/* Material row (bsh_material): 0 Lambert, 1 Microfacet, 2 Emissive, 3 OrenNayar, 4 PerfectReflect,
   5 PerfectRefract, 6 Glass, 7 MixedMaterial over the rows m[5], m[6], m[7] (-1: none). */
/* Per material row a normal map ("map<k>") that wraps the row's BSDF in NormalMapping
   (bsdf/normal_mapping.h:40-90), or -1; set only while oref_scene_create_normal_mapped builds
   the materials. */
int const* g_row_nmaps = nullptr;

Material* make_material(float const* rows, int i, TextureSystem const& ts = g_texsys, int const* maps = nullptr)
{
    float const* m = rows + 8 * i;
    int type = (int)m[0];
    float3 c(m[1], m[2], m[3]);
    if (type == 2) return new Emissive(c);
    if (g_row_nmaps && g_row_nmaps[i] >= 0 && type != 6 && type != 7)
        return new SimpleMaterial(new NormalMapping(make_bsdf(ts, rows, i, maps), "map" + std::to_string(g_row_nmaps[i])));
    if (type == 6) return new Glass(ts, m[4], c, row_maps(maps, i).color);
    if (type == 7)
    {
        MixedMaterial* mm = new MixedMaterial(m[4]);
        for (int k = 5; k < 8; ++k)
            if ((int)m[k] >= 0) mm->AddBsdf(make_bsdf(ts, rows, (int)m[k], maps));
        return mm;
    }
    return new SimpleMaterial(make_bsdf(ts, rows, i, maps));
}

/* ---------------- the reference's other samplers on the keyed stream ---------------- */
/* kind: 0 StratifiedSampler, 1 RegularSampler, 2 RandomSampler, 3 CmjSampler, 4 SobolSampler */
Sampler* make_sampler(int kind, int arg, uint32_t stream)
{
    switch (kind)
    {
    case 1: return new RegularSampler(arg);
    case 2: return new RandomSampler(arg, new KeyedRng(stream));
    case 3: return new CmjSampler(arg, new KeyedRng(stream));
    case 4: return new SobolSampler(arg, new KeyedRng(stream));
    default: return new StratifiedSampler(arg, new KeyedRng(stream));
    }
}

int sampler_spp(int kind, int arg)
{
    if (kind == 2) return arg;
    if (kind == 4) return (int)upper_power_of_two((unsigned long)arg);
    return arg * arg;
}

void key_setup(unsigned seed, int x0, int y0, int tw, int W, int spp, int img_kind, int light_kind, int brdf_kind)
{
    g_key = KeyCtx();
    g_key.on = true;
    g_key.seed = seed;
    g_key.x0 = x0; g_key.y0 = y0; g_key.tw = tw; g_key.W = W; g_key.spp = spp;
    g_key.cam_keyed = true;
    g_key.img_draws = img_kind == 3 ? 1 : img_kind == 4 ? 2 : 0;
    g_key.stride[KS_LIGHT] = light_kind == 3 ? 2 : 1;
    g_key.stride[KS_BRDF] = brdf_kind == 3 ? 2 : 1;
}

} // namespace

extern "C" {

int __wrap_rand(void)
{
    if (g_key.on) return ks_rand(g_key.base, g_key.ctr[KS_RAND]++);
    return __real_rand();
}

void __wrap_srand(unsigned s) { __real_srand(s); }

/* 1 if the first draw of a Sample2D lands in .x, 0 if it lands in .y. */
int oref_sample2d_first_is_x(void)
{
    StratifiedSampler s(1, new SeqRng());
    float2 v = s.Sample2D();
    return v.x == 0.25f ? 1 : 0;
}

/* oref_scene_create with texture coordinates (nuvs (u, v) pairs, three indices per triangle),
   texture maps (ntex maps: tex_desc = width, height, first float of texels, each) named "map<k>"
   in the scene's texture system, and per material row its colour and kr map (-1: none). */
void* oref_scene_create_textured(float const* verts, int nverts, float const* normals, int nnormals,
                                 int const* tri_v, int const* tri_n, int const* tri_m, int ntris,
                                 int const* bundles, int nbundles, float const* materials, int nmaterials,
                                 int const* lights, int nlights, float const* cam, float const* uvs, int nuvs,
                                 int const* tri_uv, float const* texels, int const* tex_desc, int ntex,
                                 int const* mat_maps)
{
    (void)nverts; (void)nnormals; (void)ntris;
    Scene* s = new Scene();
    for (int k = 0; k < ntex; ++k)
    {
        MapTextureSystem::Map mp;
        mp.w = tex_desc[3 * k];
        mp.h = tex_desc[3 * k + 1];
        mp.t.assign(texels + tex_desc[3 * k + 2], texels + tex_desc[3 * k + 2] + 3 * (size_t)mp.w * mp.h);
        s->texsys.maps["map" + std::to_string(k)] = mp;
    }
    for (int i = 0; i < nmaterials; ++i)
        s->world.materials_.push_back(std::unique_ptr<Material>(make_material(materials, i, s->texsys, mat_maps)));
    for (int b = 0; b < nbundles; ++b)
    {
        int tb = bundles[6 * b], tc = bundles[6 * b + 1], vb = bundles[6 * b + 2], vc = bundles[6 * b + 3];
        int nb = bundles[6 * b + 4], nc = bundles[6 * b + 5];
        std::vector<int> vi(3 * tc), ni(3 * tc);
        for (int k = 0; k < 3 * tc; ++k)
        {
            vi[k] = tri_v[3 * tb + k] - vb;
            ni[k] = tri_n[3 * tb + k] - nb;
        }
        RecordingMesh* m = new RecordingMesh(verts + 3 * vb, vc, normals + 3 * nb, nc, vi.data(), ni.data(),
                                             tri_m + tb, tc, tb, uvs, uvs ? nuvs : 0, tri_uv ? tri_uv + 3 * tb : nullptr);
        s->meshes.push_back(m);
        s->world.shapebundles_.push_back(std::unique_ptr<ShapeBundle>(m));
    }
    for (int l = 0; l < nlights; ++l)
    {
        int b = lights[2 * l], shape = lights[2 * l + 1];
        RecordingMesh* m = s->meshes[b];
        int mat = m->GetFaces()[shape].m;
        s->world.lights_.push_back(std::unique_ptr<Light>(new AreaLight(shape, *m, *s->world.materials_[mat])));
    }
    s->world.camera_.reset(new KeyedCamera(float3(cam[0], cam[1], cam[2]), float3(cam[3], cam[4], cam[5]),
                                                  float3(cam[6], cam[7], cam[8]), float2(cam[9], cam[10]),
                                                  cam[11], cam[12]));
    s->world.Commit();
    s->accel = new RecordingAccel(std::move(s->world.accel_));
    s->world.accel_.reset(s->accel);
    return s;
}

/* oref_scene_create_textured with a normal map per material row (-1: none) around the rows'
   single BSDFs. */
void* oref_scene_create_normal_mapped(float const* verts, int nverts, float const* normals, int nnormals,
                                      int const* tri_v, int const* tri_n, int const* tri_m, int ntris,
                                      int const* bundles, int nbundles, float const* materials, int nmaterials,
                                      int const* lights, int nlights, float const* cam, float const* uvs, int nuvs,
                                      int const* tri_uv, float const* texels, int const* tex_desc, int ntex,
                                      int const* mat_maps, int const* nmaps)
{
    g_row_nmaps = nmaps;
    void* s = oref_scene_create_textured(verts, nverts, normals, nnormals, tri_v, tri_n, tri_m, ntris, bundles,
                                         nbundles, materials, nmaterials, lights, nlights, cam, uvs, nuvs, tri_uv,
                                         texels, tex_desc, ntex, mat_maps);
    g_row_nmaps = nullptr;
    return s;
}

void* oref_scene_create(float const* verts, int nverts, float const* normals, int nnormals,
                        int const* tri_v, int const* tri_n, int const* tri_m, int ntris,
                        int const* bundles, int nbundles, float const* materials, int nmaterials,
                        int const* lights, int nlights, float const* cam)
{
    return oref_scene_create_textured(verts, nverts, normals, nnormals, tri_v, tri_n, tri_m, ntris, bundles, nbundles,
                                      materials, nmaterials, lights, nlights, cam, nullptr, 0, nullptr, nullptr, nullptr,
                                      0, nullptr);
}

void oref_scene_destroy(void* h) { delete (Scene*)h; }

/* rays: n x 8 (ox,oy,oz,tmin,dx,dy,dz,tmax). closest: prim (-1 miss), thit n x 3 (t,b1,b2).
   any: prim = occluded (1/0). */
void oref_intersect(void* h, float const* rays, int n, int any, int* prim, float* thit)
{
    Scene* s = (Scene*)h;
    for (int i = 0; i < n; ++i)
    {
        float const* q = rays + 8 * i;
        ray r(float3(q[0], q[1], q[2]), float3(q[4], q[5], q[6]), float2(q[3], q[7]));
        if (any)
        {
            prim[i] = s->world.Intersect(r) ? 1 : 0;
        }
        else
        {
            ShapeBundle::Hit hit;
            tl_rec.prim = -1;
            bool h2 = s->world.Intersect(r, hit);
            prim[i] = h2 ? tl_rec.prim : -1;
            thit[3 * i] = h2 ? tl_rec.t : 0.f;
            thit[3 * i + 1] = h2 ? tl_rec.a : 0.f;
            thit[3 * i + 2] = h2 ? tl_rec.b : 0.f;
        }
    }
}

/* Full reference Hit for closest queries: n x 20 floats
   (t, p3, n3, ng3, dpdu3, dpdv3, uv2, m, prim). */
void oref_intersect_fill(void* h, float const* rays, int n, float* out)
{
    Scene* s = (Scene*)h;
    for (int i = 0; i < n; ++i)
    {
        float const* q = rays + 8 * i;
        ray r(float3(q[0], q[1], q[2]), float3(q[4], q[5], q[6]), float2(q[3], q[7]));
        ShapeBundle::Hit hit;
        tl_rec.prim = -1;
        float* o = out + 20 * i;
        if (!s->world.Intersect(r, hit)) { for (int k = 0; k < 20; ++k) o[k] = 0.f; o[19] = -1.f; continue; }
        o[0] = hit.t;
        o[1] = hit.p.x; o[2] = hit.p.y; o[3] = hit.p.z;
        o[4] = hit.n.x; o[5] = hit.n.y; o[6] = hit.n.z;
        o[7] = hit.ng.x; o[8] = hit.ng.y; o[9] = hit.ng.z;
        o[10] = hit.dpdu.x; o[11] = hit.dpdu.y; o[12] = hit.dpdu.z;
        o[13] = hit.dpdv.x; o[14] = hit.dpdv.y; o[15] = hit.dpdv.z;
        o[16] = hit.uv.x; o[17] = hit.uv.y; o[18] = (float)hit.m; o[19] = (float)tl_rec.prim;
    }
}

void oref_record_begin(void* h, int max_rays, int stride)
{
    Scene* s = (Scene*)h;
    s->accel->recs_.clear();
    s->accel->recording_ = true;
    s->accel->max_ = max_rays;
    s->accel->stride_ = stride;
    s->accel->seen_ = 0;
}

/* rays n x 8, kind n, prim n, thit n x 3. Returns the number of records and stops recording. */
int oref_record_fetch(void* h, float* rays, int* kind, int* prim, float* thit)
{
    Scene* s = (Scene*)h;
    s->accel->recording_ = false;
    int n = (int)s->accel->recs_.size();
    for (int i = 0; i < n; ++i)
    {
        RayRec const& r = s->accel->recs_[i];
        if (rays) std::memcpy(rays + 8 * i, r.r, 8 * sizeof(float));
        if (kind) kind[i] = r.kind;
        if (prim) prim[i] = r.prim;
        if (thit) { thit[3 * i] = r.t; thit[3 * i + 1] = r.a; thit[3 * i + 2] = r.b; }
    }
    return n;
}

/* GI render of a tile with the keyed sample stream, single-threaded reference
   ImageRenderer::RenderTile (imagerenderer.cpp:61-101). Image sampler =
   StratifiedSampler(img_grid), light = brdf = StratifiedSampler(nee_grid),
   tracer = GiTracer(depth). Outputs are tile-local, row-major, NOT y-flipped. */
void oref_render_keyed(void* h, int W, int H, int x0, int y0, int tw, int th, int img_grid, int nee_grid,
                       int depth, unsigned seed, float* sum, float* sumsq, float* cnt, float* per_sample,
                       long long* rays)
{
    Scene* s = (Scene*)h;
    int spp = img_grid * img_grid;
    StatsPlane plane(W, H, x0, y0, tw, th, spp, per_sample != nullptr);
    g_key = KeyCtx();
    g_key.on = true;
    g_key.seed = seed;
    g_key.x0 = x0; g_key.y0 = y0; g_key.tw = tw; g_key.W = W; g_key.spp = spp;
    reset_counts();
    {
        ImageRenderer r(plane, new GiTracer(depth), new StratifiedSampler(img_grid, new KeyedRng(KS_IMG)),
                        new StratifiedSampler(nee_grid, new KeyedRng(KS_LIGHT)),
                        new StratifiedSampler(nee_grid, new KeyedRng(KS_BRDF)));
        r.RenderTile(s->world, int2(x0, y0), int2(tw, th));
    }
    g_key.on = false;
    read_counts(rays);
    std::memcpy(sum, plane.sum_.data(), plane.sum_.size() * sizeof(float));
    std::memcpy(sumsq, plane.sq_.data(), plane.sq_.size() * sizeof(float));
    std::memcpy(cnt, plane.n_.data(), plane.n_.size() * sizeof(float));
    if (per_sample) std::memcpy(per_sample, plane.samples_.data(), plane.samples_.size() * sizeof(float));
}

/* Full-frame GI through ImageRenderer::Render (imagerenderer.cpp:9-59), which resets the
   three samplers after every pixel: image sampler <img_kind>(img_arg), light and BSDF
   samplers <light_kind>(1), <brdf_kind>(1), GiTracer(depth), all on the keyed stream.
   Outputs row-major over the frame, NOT y-flipped; per_sample in (pixel, s) order. */
void oref_render_keyed_frame(void* h, int W, int H, int img_kind, int img_arg, int light_kind, int brdf_kind,
                             int depth, unsigned seed, float* sum, float* sumsq, float* cnt, float* per_sample,
                             long long* rays)
{
    Scene* s = (Scene*)h;
    int spp = sampler_spp(img_kind, img_arg);
    StatsPlane plane(W, H, 0, 0, W, H, spp, per_sample != nullptr);
    key_setup(seed, 0, 0, W, W, spp, img_kind, light_kind, brdf_kind);
    reset_counts();
    {
        Sampler* img = make_sampler(img_kind, img_arg, KS_IMG);
        Sampler* light = make_sampler(light_kind, 1, KS_LIGHT);
        Sampler* brdf = make_sampler(brdf_kind, 1, KS_BRDF);
        ImageRenderer r(plane, new GiTracer(depth), img, light, brdf);
        r.Render(s->world);
    }
    g_key.on = false;
    read_counts(rays);
    std::memcpy(sum, plane.sum_.data(), plane.sum_.size() * sizeof(float));
    std::memcpy(sumsq, plane.sq_.data(), plane.sq_.size() * sizeof(float));
    std::memcpy(cnt, plane.n_.data(), plane.n_.size() * sizeof(float));
    if (per_sample) std::memcpy(per_sample, plane.samples_.data(), plane.samples_.size() * sizeof(float));
}

/* The image samples of <kind>(arg) for pixels 0..npix-1 of a one-row image keyed by seed,
   with Reset() after every pixel as ImageRenderer::Render does: out[(p*spp + s)*2 + {0,1}]. */
void oref_image_samples(int kind, int arg, unsigned seed, int npix, float* out)
{
    int spp = sampler_spp(kind, arg);
    key_setup(seed, 0, 0, npix, npix, spp, kind, 0, 0);
    {
        std::unique_ptr<Sampler> smp(make_sampler(kind, arg, KS_IMG));
        for (int p = 0; p < npix; ++p)
        {
            for (int k = 0; k < spp; ++k)
            {
                float2 v = smp->Sample2D();
                out[2 * ((size_t)p * spp + k)] = v.x;
                out[2 * ((size_t)p * spp + k) + 1] = v.y;
            }
            smp->Reset();
        }
    }
    g_key.on = false;
}

/* n successive samples of a light sampler <kind>(1) within camera sample 0 keyed by seed. */
void oref_nee_samples(int kind, unsigned seed, int n, float* out)
{
    key_setup(seed, 0, 0, 1, 1, 1, 0, kind, kind);
    key_enter_sample(0);
    {
        std::unique_ptr<Sampler> smp(make_sampler(kind, 1, KS_LIGHT));
        for (int i = 0; i < n; ++i)
        {
            float2 v = smp->Sample2D();
            out[2 * i] = v.x;
            out[2 * i + 1] = v.y;
        }
    }
    g_key.on = false;
}

/* Wall-clock timing of the reference renderer as shipped: McRng samplers seeded by
   srand(seed), GiTracer(depth), MtImageRenderer with its thread_pool (mt != 0) or
   the single-threaded ImageRenderer. rays[0..1] = closest, any-hit query counts. */
double oref_time_render(void* h, int W, int H, int img_grid, int nee_grid, int depth, unsigned seed, int mt,
                        long long* rays)
{
    Scene* s = (Scene*)h;
    StatsPlane plane(W, H, 0, 0, W, H, img_grid * img_grid, false);
    g_key.on = false;
    std::srand(seed);
    reset_counts();
    auto t0 = std::chrono::steady_clock::now();
    if (mt)
    {
        MtImageRenderer r(plane, new GiTracer(depth), new StratifiedSampler(img_grid, new McRng()),
                          new StratifiedSampler(nee_grid, new McRng()), new StratifiedSampler(nee_grid, new McRng()));
        r.Render(s->world);
    }
    else
    {
        ImageRenderer r(plane, new GiTracer(depth), new StratifiedSampler(img_grid, new McRng()),
                        new StratifiedSampler(nee_grid, new McRng()), new StratifiedSampler(nee_grid, new McRng()));
        r.Render(s->world);
    }
    auto t1 = std::chrono::steady_clock::now();
    read_counts(rays);
    return std::chrono::duration<double>(t1 - t0).count();
}

/* BSDF probes. mat: 8 floats. frames n x 12 (n, ng, dpdu, dpdv); view n x 3 (the
   Sample() "wi" argument); u n x 2; ea/eb n x 3: arguments of Evaluate/GetPdf.
   out n x 12: sample f3, pdf, wo3, type, eval f3, getpdf. */
/* Material::Sample / Evaluate / GetPdf of world.materials_[midx] (any row type, MixedMaterial
   with its components), rand() of case i drawn from the keyed stream ks_base(seed, i) from
   counter 0: out n x 12 as oref_bsdf. */
void oref_material(void* h, int midx, unsigned seed, int n, float const* frames, float const* view, float const* u,
                   float const* ea, float const* eb, float* out)
{
    Scene* s = (Scene*)h;
    Material const& m = *s->world.materials_[midx];
    for (int i = 0; i < n; ++i)
    {
        float const* f = frames + 12 * i;
        ShapeBundle::Hit hit;
        hit.t = 1.f;
        hit.p = float3(0, 0, 0);
        hit.n = float3(f[0], f[1], f[2]);
        hit.ng = float3(f[3], f[4], f[5]);
        hit.dpdu = float3(f[6], f[7], f[8]);
        hit.dpdv = float3(f[9], f[10], f[11]);
        hit.uv = float2(0, 0);
        hit.m = midx;
        hit.bundle = nullptr;
        g_key = KeyCtx();
        g_key.on = true;
        g_key.seed = seed;
        g_key.base = ks_base(seed, (uint32_t)i);
        float3 wi(view[3 * i], view[3 * i + 1], view[3 * i + 2]);
        float3 wo;
        float pdf = 0.f;
        int type = 0;
        ShapeBundle::Hit h1 = hit;
        float3 fs = m.Sample(h1, float2(u[2 * i], u[2 * i + 1]), wi, wo, pdf, type);
        g_key.on = false;
        ShapeBundle::Hit h2 = hit;
        float3 a(ea[3 * i], ea[3 * i + 1], ea[3 * i + 2]), b(eb[3 * i], eb[3 * i + 1], eb[3 * i + 2]);
        float3 fe = m.Evaluate(h2, a, b);
        ShapeBundle::Hit h3 = hit;
        float pe = m.GetPdf(h3, a, b);
        float* o = out + 12 * i;
        o[0] = fs.x; o[1] = fs.y; o[2] = fs.z; o[3] = pdf;
        o[4] = wo.x; o[5] = wo.y; o[6] = wo.z; o[7] = (float)type;
        o[8] = fe.x; o[9] = fe.y; o[10] = fe.z; o[11] = pe;
    }
}

void oref_bsdf(float const* mat, int n, float const* frames, float const* view, float const* u,
               float const* ea, float const* eb, float* out)
{
    std::unique_ptr<Material> m(make_material(mat, 0));
    for (int i = 0; i < n; ++i)
    {
        float const* f = frames + 12 * i;
        ShapeBundle::Hit hit;
        hit.t = 1.f;
        hit.p = float3(0, 0, 0);
        hit.n = float3(f[0], f[1], f[2]);
        hit.ng = float3(f[3], f[4], f[5]);
        hit.dpdu = float3(f[6], f[7], f[8]);
        hit.dpdv = float3(f[9], f[10], f[11]);
        hit.uv = float2(0, 0);
        hit.m = 0;
        hit.bundle = nullptr;
        float3 wi(view[3 * i], view[3 * i + 1], view[3 * i + 2]);
        float3 wo;
        float pdf = 0.f;
        int type = 0;
        ShapeBundle::Hit h1 = hit;
        float3 fs = m->Sample(h1, float2(u[2 * i], u[2 * i + 1]), wi, wo, pdf, type);
        ShapeBundle::Hit h2 = hit;
        float3 a(ea[3 * i], ea[3 * i + 1], ea[3 * i + 2]), b(eb[3 * i], eb[3 * i + 1], eb[3 * i + 2]);
        float3 fe = m->Evaluate(h2, a, b);
        ShapeBundle::Hit h3 = hit;
        float pe = m->GetPdf(h3, a, b);
        float* o = out + 12 * i;
        o[0] = fs.x; o[1] = fs.y; o[2] = fs.z; o[3] = pdf;
        o[4] = wo.x; o[5] = wo.y; o[6] = wo.z; o[7] = (float)type;
        o[8] = fe.x; o[9] = fe.y; o[10] = fe.z; o[11] = pe;
    }
}

/* Area light probes on scene light `light`. p n x 3 shading points, u n x 2,
   w n x 3 directions for GetPdf. out n x 8: le3, pdf, d3, getpdf. */
/* ShapeBundle::SetTransform(m, minv) of every bundle (n x 32 floats: row-major m, then row-major
   minv), then World::Commit again (the BVH over the transformed bounds) under the recording
   accelerator. */
void oref_scene_set_transforms(void* h, float const* xf, int n)
{
    Scene* s = (Scene*)h;
    for (int b = 0; b < n && b < (int)s->meshes.size(); ++b)
    {
        float const* m = xf + 32 * b;
        float const* mi = m + 16;
        s->meshes[b]->SetTransform(matrix(m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9], m[10], m[11],
                                          m[12], m[13], m[14], m[15]),
                                   matrix(mi[0], mi[1], mi[2], mi[3], mi[4], mi[5], mi[6], mi[7], mi[8], mi[9], mi[10],
                                          mi[11], mi[12], mi[13], mi[14], mi[15]));
    }
    s->world.accel_.reset();
    s->world.Commit();
    s->accel = new RecordingAccel(std::move(s->world.accel_));
    s->world.accel_.reset(s->accel);
}

/* The map the scene's environment lights read ("env"): h rows of w RGB floats. */
void oref_scene_set_env(void* h, float const* tex, int w, int hgt)
{
    Scene* s = (Scene*)h;
    MapTextureSystem::Map m;
    m.w = w;
    m.h = hgt;
    m.t.assign(tex, tex + 3 * (size_t)w * hgt);
    s->texsys.maps["env"] = m;
}

/* Appends a light to World::lights_ after the area lights: type 1 PointLight(v, e), 2
   DirectionalLight(v, e), 3 EnvironmentLight(env, scale, gamma), 4 EnvironmentLightIs(env,
   scale, gamma) (its ctor reads the map, so set the map first). */
void oref_scene_add_light(void* h, int type, float const* v, float const* e, float scale, float gamma)
{
    Scene* s = (Scene*)h;
    Light* L = nullptr;
    if (type == 1) L = new PointLight(float3(v[0], v[1], v[2]), float3(e[0], e[1], e[2]));
    else if (type == 2) L = new DirectionalLight(float3(v[0], v[1], v[2]), float3(e[0], e[1], e[2]));
    else if (type == 3) L = new EnvironmentLight(s->texsys, "env", scale, gamma);
    else L = new EnvironmentLightIs(s->texsys, "env", scale, gamma);
    s->world.lights_.push_back(std::unique_ptr<Light>(L));
}

/* Light::GetSample / GetPdf / GetLe of light `light` at points p with shading normals nrm:
   out[16 i ..] = le(3), pdf, d(3), GetPdf(w), GetLe(ray along w)(3), 0 x 5. */
void oref_light_full(void* h, int light, int n, float const* p, float const* nrm, float const* u, float const* w,
                     float* out)
{
    Scene* s = (Scene*)h;
    Light const& L = *s->world.lights_[light];
    for (int i = 0; i < n; ++i)
    {
        ShapeBundle::Hit hit;
        hit.p = float3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        hit.n = float3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
        float3 d;
        float pdf = 0.f;
        float3 le = L.GetSample(hit, float2(u[2 * i], u[2 * i + 1]), d, pdf);
        float3 ww(w[3 * i], w[3 * i + 1], w[3 * i + 2]);
        float gp = L.GetPdf(hit, ww);
        ray r;
        r.o = hit.p;
        r.d = ww;
        float3 ge = L.GetLe(r);
        float* o = out + 16 * i;
        o[0] = le.x; o[1] = le.y; o[2] = le.z; o[3] = pdf;
        o[4] = d.x; o[5] = d.y; o[6] = d.z; o[7] = gp;
        o[8] = ge.x; o[9] = ge.y; o[10] = ge.z;
        for (int k = 11; k < 16; ++k) o[k] = 0.f;
    }
}

/* The texture system's lookup of the scene's map: out[3 i ..] = Sample("env", uv_i) with the
   given filter (0 point, 1 bilinear) and wrap (0 repeat, 1 mirror). */
void oref_texture_sample(void* h, int n, float const* uv, int filter, int wrapmode, float* out)
{
    Scene* s = (Scene*)h;
    TextureSystem::Options o(filter ? TextureSystem::Options::kBilinear : TextureSystem::Options::kPoint,
                             wrapmode ? TextureSystem::Options::kMirror : TextureSystem::Options::kRepeat);
    for (int i = 0; i < n; ++i)
    {
        float3 v = s->texsys.Sample("env", float2(uv[2 * i], uv[2 * i + 1]), float2(0, 0), o);
        out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
    }
}

/* Distribution2D(n, m, values) (math/distribution2d.cpp): Sample2D at k sample pairs
   (out_s: sample uv, pdf) and Pdf at k query points (out_p). */
void oref_dist2d(int n, int m, float const* values, int k, float const* u, float const* q, float* out_s, float* out_p)
{
    Distribution2D d(n, m, values);
    for (int i = 0; i < k; ++i)
    {
        float pdf = 0.f;
        float2 sm = d.Sample2D(float2(u[2 * i], u[2 * i + 1]), pdf);
        out_s[3 * i] = sm.x; out_s[3 * i + 1] = sm.y; out_s[3 * i + 2] = pdf;
        out_p[i] = d.Pdf(float2(q[2 * i], q[2 * i + 1]));
    }
}

void oref_light(void* h, int light, int n, float const* p, float const* u, float const* w, float* out)
{
    Scene* s = (Scene*)h;
    Light const& L = *s->world.lights_[light];
    for (int i = 0; i < n; ++i)
    {
        ShapeBundle::Hit hit;
        hit.p = float3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        float3 d;
        float pdf = 0.f;
        float3 le = L.GetSample(hit, float2(u[2 * i], u[2 * i + 1]), d, pdf);
        float gp = L.GetPdf(hit, float3(w[3 * i], w[3 * i + 1], w[3 * i + 2]));
        float* o = out + 8 * i;
        o[0] = le.x; o[1] = le.y; o[2] = le.z; o[3] = pdf;
        o[4] = d.x; o[5] = d.y; o[6] = d.z; o[7] = gp;
    }
}

/* camera: 13 floats (eye3, at3, up3, znear, zfar, fovy, aspect); samples n x 2; rays n x 8 */
void oref_camera_rays(float const* cam, float const* samples, int n, float* rays)
{
    PerscpectiveCamera c(float3(cam[0], cam[1], cam[2]), float3(cam[3], cam[4], cam[5]),
                         float3(cam[6], cam[7], cam[8]), float2(cam[9], cam[10]), cam[11], cam[12]);
    for (int i = 0; i < n; ++i)
    {
        ray r;
        c.GenerateRay(float2(samples[2 * i], samples[2 * i + 1]), r);
        float* o = rays + 8 * i;
        o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z; o[3] = r.t.x;
        o[4] = r.d.x; o[5] = r.d.y; o[6] = r.d.z; o[7] = r.t.y;
    }
}

void oref_mcrng(unsigned seed, int n, unsigned* out)
{
    std::srand(seed);
    McRng r;
    for (int i = 0; i < n; ++i) out[i] = r.NextUint();
}

void oref_stratified(int grid, unsigned seed, int n, float* out)
{
    std::srand(seed);
    StratifiedSampler s(grid, new McRng());
    for (int i = 0; i < n; ++i)
    {
        float2 v = s.Sample2D();
        out[2 * i] = v.x;
        out[2 * i + 1] = v.y;
    }
}

void oref_sh_eval(float const* dirs, int n, int lmax, float* out)
{
    int nt = NumShTerms(lmax);
    for (int i = 0; i < n; ++i)
        ShEvaluate(float3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), lmax, out + (size_t)nt * i);
}

void oref_sh_convolve(int lmax, float const* cin, float* cout)
{
    int nt = NumShTerms(lmax);
    std::vector<float3> a(nt), b(nt);
    for (int i = 0; i < nt; ++i) a[i] = float3(cin[3 * i], cin[3 * i + 1], cin[3 * i + 2]);
    ShConvolveCosTheta(lmax, a.data(), b.data());
    for (int i = 0; i < nt; ++i) { cout[3 * i] = b[i].x; cout[3 * i + 1] = b[i].y; cout[3 * i + 2] = b[i].z; }
}

void oref_sh_project(float const* env, int W, int H, int lmax, float* coeffs)
{
    int nt = NumShTerms(lmax);
    std::vector<float3> e((size_t)W * H), c(nt);
    for (size_t i = 0; i < (size_t)W * H; ++i) e[i] = float3(env[3 * i], env[3 * i + 1], env[3 * i + 2]);
    ShProjectEnvironmentMap(e.data(), W, H, lmax, c.data());
    for (int i = 0; i < nt; ++i) { coeffs[3 * i] = c[i].x; coeffs[3 * i + 1] = c[i].y; coeffs[3 * i + 2] = c[i].z; }
}

/* ShTracer::GetE (shtracer.cpp:20-36) is private; this composes the same two public
   reference functions the same way: convolve, evaluate, sum ylm * c * (0.7 / pi). */
void oref_sh_irradiance(int lmax, float const* coeffs, float const* normals, int n, float* out)
{
    int nt = NumShTerms(lmax);
    std::vector<float3> c(nt), cc(nt);
    std::vector<float> ylm(nt);
    for (int i = 0; i < nt; ++i) c[i] = float3(coeffs[3 * i], coeffs[3 * i + 1], coeffs[3 * i + 2]);
    ShConvolveCosTheta(lmax, c.data(), cc.data());
    for (int i = 0; i < n; ++i)
    {
        ShEvaluate(float3(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]), lmax, ylm.data());
        float3 e;
        for (int k = 0; k < nt; ++k) e += ylm[k] * cc[k] * (0.7f / (float)M_PI);
        out[3 * i] = e.x; out[3 * i + 1] = e.y; out[3 * i + 2] = e.z;
    }
}

} // extern "C"
