/* The reference's own host code driving the GPU path through include/banshee_world.h.

   A Cornell-style World is built with the reference's API (Mesh, SimpleMaterial(Lambert),
   Emissive, AreaLight, PerscpectiveCamera, World::Commit), then:
   1. rendered by the reference's ImageRenderer + GiTracer on the CPU and by HipImageRenderer on
      the GPU: per-pixel means agree within 3 standard errors;
   2. queried through World::Intersect with World::accel_ swapped from the reference's Bvh to
      HipAccelerator: closest hits (t, point) and occlusion agree ray for ray;
   3. rendered again by the reference's CPU GiTracer with HipAccelerator installed (every ray a
      GPU query): the image equals the one over the reference's Bvh.
   Prints one JSON line; exit status 0 when every check holds. Built by oracle/Makefile
   (target adapter) from the reference's sources where they are present. */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "banshee_world.h"

#include "bsdf/lambert.h"
#include "camera/perspective_camera.h"
#include "light/arealight.h"
#include "material/emissive.h"
#include "material/simplematerial.h"
#include "renderer/imagerenderer.h"
#include "rng/mcrng.h"
#include "sampler/stratified_sampler.h"
#include "texture/texturesystem.h"
#include "tracer/gitracer.h"
#include "util/progressreporter.h"

namespace {

class NoTextures : public TextureSystem
{
public:
    float3 Sample(std::string const&, float2 const&, float2 const&, Options const&) const { return float3(0, 0, 0); }
    void GetTextureInfo(std::string const&, TextureDesc& d) const { d.width = d.height = 0; }
};

// Per-pixel sum, sum of squares and count of the samples ImageRenderer adds.
class MomentsPlane : public ImagePlane
{
public:
    MomentsPlane(int w, int h) : ImagePlane(int2(w, h), nullptr), s(3 * w * h), q(3 * w * h), n(w * h) {}
    void WriteSample(int2 const& p, float3 const& v)
    {
        int i = p.y * resolution().x + p.x;
        s[3 * i] += v.x; s[3 * i + 1] += v.y; s[3 * i + 2] += v.z;
        q[3 * i] += v.x * v.x; q[3 * i + 1] += v.y * v.y; q[3 * i + 2] += v.z * v.z;
        n[i] += 1;
    }
    std::vector<double> s, q, n;
};

struct Quad
{
    float p[4][3];
    float nrm[3];
    int mat;
};

}  // namespace

int main()
{
    NoTextures tex;
    World world;

    bsh_world::SceneDesc desc;
    const float kd[5][3] = {{0.7f, 0.7f, 0.7f}, {0.7f, 0.2f, 0.2f}, {0.2f, 0.6f, 0.2f}, {0.6f, 0.6f, 0.5f}, {0, 0, 0}};
    for (int m = 0; m < 4; ++m)
    {
        world.materials_.push_back(std::unique_ptr<Material>(
            new SimpleMaterial(new Lambert(tex, float3(kd[m][0], kd[m][1], kd[m][2])))));
        bsh_material row = {BSH_LAMBERT, {kd[m][0], kd[m][1], kd[m][2]}, 0.f, 0.f, {0.f, 0.f}};
        desc.materials.push_back(row);
    }
    world.materials_.push_back(std::unique_ptr<Material>(new Emissive(float3(15.f, 15.f, 15.f))));
    desc.materials.push_back(bsh_material{BSH_EMISSIVE, {15.f, 15.f, 15.f}, 0.f, 0.f, {0.f, 0.f}});
    const Quad quads[] = {
        {{{0, 0, 0}, {2, 0, 0}, {2, 0, 2}, {0, 0, 2}}, {0, 1, 0}, 0},           // floor
        {{{0, 2, 0}, {0, 2, 2}, {2, 2, 2}, {2, 2, 0}}, {0, -1, 0}, 0},          // ceiling
        {{{0, 0, 0}, {0, 2, 0}, {2, 2, 0}, {2, 0, 0}}, {0, 0, 1}, 3},           // back
        {{{0, 0, 0}, {0, 0, 2}, {0, 2, 2}, {0, 2, 0}}, {1, 0, 0}, 1},           // left
        {{{2, 0, 0}, {2, 2, 0}, {2, 2, 2}, {2, 0, 2}}, {-1, 0, 0}, 2},          // right
        {{{0.5f, 0.6f, 0.5f}, {1.1f, 0.6f, 0.5f}, {1.1f, 0.6f, 1.1f}, {0.5f, 0.6f, 1.1f}}, {0, 1, 0}, 3},  // shelf
        {{{0.75f, 1.98f, 0.75f}, {1.25f, 1.98f, 0.75f}, {1.25f, 1.98f, 1.25f}, {0.75f, 1.98f, 1.25f}},
         {0, -1, 0}, 4},                                                          // light
    };
    Mesh* light_mesh = nullptr;
    for (Quad const& qd : quads)
    {
        float v[12], n[12];
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 3; ++c)
            {
                v[3 * k + c] = qd.p[k][c];
                n[3 * k + c] = qd.nrm[c];
            }
        int idx[6] = {0, 1, 2, 0, 2, 3};
        int mats[2] = {qd.mat, qd.mat};
        Mesh* mesh = new Mesh(v, 4, 0, n, 4, 0, nullptr, 0, 0, idx, 0, idx, 0, nullptr, 0, mats, (int)sizeof(int), 2);
        world.shapebundles_.push_back(std::unique_ptr<ShapeBundle>(mesh));
        if (qd.mat == 4) light_mesh = mesh;
    }
    for (int s = 0; s < 2; ++s)
    {
        world.lights_.push_back(std::unique_ptr<Light>(new AreaLight(s, *light_mesh, *world.materials_[4])));
        desc.area_lights.push_back(bsh_area_light{(int32_t)world.shapebundles_.size() - 1, s});
    }
    const float eye[3] = {1.f, 1.f, 3.5f}, at[3] = {1.f, 1.f, 0.f}, up[3] = {0.f, 1.f, 0.f};
    const float fovy = 3.14159265f / 4.f;
    world.camera_.reset(new PerscpectiveCamera(float3(eye[0], eye[1], eye[2]), float3(at[0], at[1], at[2]),
                                               float3(up[0], up[1], up[2]), float2(0.01f, 10000.f), fovy, 1.f));
    desc.camera = bsh_camera{{eye[0], eye[1], eye[2]}, {at[0], at[1], at[2]}, {up[0], up[1], up[2]}, 0.01f, 10000.f,
                             fovy, 1.f};
    world.Commit();

    bsh_context* ctx = nullptr;
    if (bsh_context_create(0, &ctx) != BSH_OK)
    {
        std::printf("{\"error\": \"%s\"}\n", bsh_last_error());
        return 2;
    }
    bsh_scene* scene = bsh_world::bsh_scene_from_world(ctx, world, desc);

    // 1. the same frame through ImageRenderer (CPU) and HipImageRenderer (GPU)
    const int W = 32, H = 32, grid = 4, depth = 5;
    MomentsPlane cpu(W, H), gpu(W, H);
    std::srand(1);
    {
        ImageRenderer r(cpu, new GiTracer(depth), new StratifiedSampler(grid, new McRng()),
                        new StratifiedSampler(1, new McRng()), new StratifiedSampler(1, new McRng()));
        r.Render(world);
    }
    bsh_world::HipImageRenderer hr(gpu, ctx, scene, grid, depth, 7u);
    hr.Render(world);
    std::vector<float> const& acc = hr.last_accum();
    int within = 0, px = W * H;
    double max_dev = 0;
    for (int i = 0; i < px; ++i)
        for (int c = 0; c < 3; ++c)
        {
            double n1 = cpu.n[i], n2 = acc[8 * i + 3];
            double m1 = cpu.s[3 * i + c] / n1, m2 = acc[8 * i + c] / n2;
            double v1 = std::max(0.0, cpu.q[3 * i + c] / n1 - m1 * m1) * n1 / (n1 - 1);
            double v2 = std::max(0.0, (double)acc[8 * i + 4 + c] / n2 - m2 * m2) * n2 / (n2 - 1);
            double se = std::sqrt(v1 / n1 + v2 / n2) + 1e-6;
            within += std::fabs(m1 - m2) <= 3 * se;
            max_dev = std::max(max_dev, std::fabs(m1 - m2) / se);
        }
    const double frac3se = within / (3.0 * px);

    // 2. World::Intersect over HipAccelerator against the reference's Bvh, ray for ray
    std::unique_ptr<Intersectable> bvh = std::move(world.accel_);
    bsh_world::HipAccelerator* ha = new bsh_world::HipAccelerator(ctx, scene, world);
    world.accel_.reset(ha);
    McRng rng;
    const int K = 4000;
    int hit_same = 0, occ_same = 0;
    std::vector<bsh_ray> batch(K);
    std::vector<int> ref_hit(K);
    std::vector<float> ref_t(K);
    for (int k = 0; k < K; ++k)
    {
        float3 o(0.1f + 1.8f * rng.NextFloat(), 0.1f + 1.8f * rng.NextFloat(), 0.1f + 1.8f * rng.NextFloat());
        float3 d = normalize(float3(rng.NextFloat() - 0.5f, rng.NextFloat() - 0.5f, rng.NextFloat() - 0.5f));
        ray r;
        r.o = o;
        r.d = d;
        r.t = float2(0.01f, k % 3 == 0 ? 0.7f : 1000000.f);
        ShapeBundle::Hit h0, h1;
        h0.t = r.t.y;
        bool b0 = bvh->Intersect(r, h0);
        bool b1 = world.Intersect(r, h1);   // seeds h1.t = r.t.y, then HipAccelerator
        hit_same += (b0 == b1) && (!b0 || (std::fabs(h0.t - h1.t) <= 1e-5f * h0.t && (h0.p - h1.p).sqnorm() < 1e-10f));
        occ_same += bvh->Intersect(r) == world.Intersect(r);
        batch[k] = bsh_ray{o.x, o.y, o.z, r.t.x, d.x, d.y, d.z, r.t.y};
        ref_hit[k] = b0;
        ref_t[k] = b0 ? h0.t : 0.f;
    }
    std::vector<bsh_hit> bh(K);
    ha->IntersectBatch(batch.data(), K, bh.data());
    int batch_same = 0;
    for (int k = 0; k < K; ++k)
        batch_same += (bh[k].prim >= 0) == (ref_hit[k] != 0) &&
                      (!ref_hit[k] || std::fabs(bh[k].t - ref_t[k]) <= 1e-5f * ref_t[k]);

    // 3. the reference's CPU GiTracer with HipAccelerator installed equals it over the Bvh
    MomentsPlane over_gpu(12, 12), over_bvh(12, 12);
    std::srand(5);
    {
        ImageRenderer r(over_gpu, new GiTracer(depth), new StratifiedSampler(1, new McRng()),
                        new StratifiedSampler(1, new McRng()), new StratifiedSampler(1, new McRng()));
        r.Render(world);
    }
    world.accel_ = std::move(bvh);
    std::srand(5);
    {
        ImageRenderer r(over_bvh, new GiTracer(depth), new StratifiedSampler(1, new McRng()),
                        new StratifiedSampler(1, new McRng()), new StratifiedSampler(1, new McRng()));
        r.Render(world);
    }
    int img_same = 0;
    for (int i = 0; i < 12 * 12 * 3; ++i)
        img_same += std::fabs(over_gpu.s[i] - over_bvh.s[i]) <= 1e-4 * std::fabs(over_bvh.s[i]) + 1e-6;
    const double frac_img = img_same / (12.0 * 12 * 3);

    bsh_scene_destroy(scene);
    bsh_context_destroy(ctx);
    const bool ok = frac3se >= 0.99 && hit_same == K && occ_same == K && batch_same == K && frac_img >= 0.99;
    std::printf("{\"pixels\": %d, \"within_3se\": %.5f, \"max_dev_se\": %.3f, \"rays\": %d, \"closest_same\": %d, "
                "\"occluded_same\": %d, \"batch_same\": %d, \"cpu_tracer_gpu_accel_same\": %.5f, \"ok\": %s}\n",
                px, frac3se, max_dev, K, hit_same, occ_same, batch_same, frac_img, ok ? "true" : "false");
    return ok ? 0 : 1;
}
