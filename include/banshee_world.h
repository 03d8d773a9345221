/* Reference-side adapters over the C ABI (banshee_hip.h), for host code built on the reference's
   World (FireRays/Banshee). Header-only C++; include it where the reference's headers are on
   the include path (world/world.h, primitive/mesh.h, ...). Nothing here is part of
   libbanshee_hip.so.

   - bsh_scene_from_world: a bsh_scene from a committed World whose bundles are Meshes with
     identity transforms (World::Commit, world/world.cpp:5-16).
   - HipImageRenderer: a Renderer (renderer/renderer.h:49) that renders the frame with
     bsh_render and writes each pixel's mean into the caller's ImagePlane, where
     ImageRenderer::Render would add every camera sample (renderer/imagerenderer.cpp:9-59).
   - HipAccelerator: an Intersectable (accelerator/intersectable.h:41-49) over
     bsh_intersect_closest / bsh_intersect_any for World::accel_, with a batch form. */
#pragma once

#include <cstring>
#include <stdexcept>
#include <utility>
#include <vector>

#include "banshee_hip.h"

#include "accelerator/intersectable.h"
#include "imageplane/imageplane.h"
#include "primitive/mesh.h"
#include "renderer/renderer.h"
#include "world/world.h"

namespace bsh_world {

inline void check(bsh_status s)
{
    if (s != BSH_OK) throw std::runtime_error(bsh_last_error());
}

/* What the reference's public interfaces do not expose: the material classes keep their
   parameters private and AreaLight its shape index, so the host code that built the World
   passes them as tables parallel to world.materials_ and world.lights_ (area lights first,
   then PointLight / DirectionalLight / EnvironmentLight rows, the order World::lights_ holds
   them in). */
struct SceneDesc
{
    std::vector<bsh_material> materials;   // one per world.materials_[i]
    std::vector<bsh_area_light> area_lights;
    std::vector<bsh_light> other_lights;
    std::vector<float> env_map;             // environment lights' map (rows of RGB texels), if any
    int env_width = 0, env_height = 0;
    bsh_camera camera{};
};

/* Mesh exposes vertices and faces but not its normal array; FillHit at a face's three corners
   returns the corner normals normalised (mesh.cpp:47-97, identity transform), which is what
   the GPU interpolates. Returns the scene committed with `bvh` (NULL for the defaults). */
inline bsh_scene* bsh_scene_from_world(bsh_context* ctx, World const& world, SceneDesc const& d,
                                       bsh_bvh_params const* bvh = nullptr)
{
    std::vector<float> v, n;
    std::vector<int32_t> tv, tn, tm;
    std::vector<bsh_bundle> bundles;
    for (auto const& sb : world.shapebundles_)
    {
        Mesh const* m = dynamic_cast<Mesh const*>(sb.get());
        if (!m) throw std::runtime_error("bsh_scene_from_world: every shape bundle must be a Mesh");
        bsh_bundle b;
        b.tri_begin = (int32_t)tm.size();
        b.tri_count = (int32_t)m->GetNumFaces();
        b.vert_begin = (int32_t)(v.size() / 3);
        b.vert_count = (int32_t)m->GetNumVertices();
        b.norm_begin = (int32_t)(n.size() / 3);
        b.norm_count = 3 * b.tri_count;
        for (size_t i = 0; i < m->GetNumVertices(); ++i)
        {
            float3 p = m->GetVertices()[i];
            v.insert(v.end(), {p.x, p.y, p.z});
        }
        for (size_t f = 0; f < m->GetNumFaces(); ++f)
        {
            Mesh::Face const& F = m->GetFaces()[f];
            const float corner[3][2] = {{0.f, 0.f}, {1.f, 0.f}, {0.f, 1.f}};
            for (int k = 0; k < 3; ++k)
            {
                ShapeBundle::Hit h;
                m->FillHit(f, 1.f, corner[k][0], corner[k][1], h);
                n.insert(n.end(), {h.n.x, h.n.y, h.n.z});
            }
            const int32_t nb = b.norm_begin + 3 * (int32_t)f;
            tv.insert(tv.end(), {b.vert_begin + F.vi0, b.vert_begin + F.vi1, b.vert_begin + F.vi2});
            tn.insert(tn.end(), {nb, nb + 1, nb + 2});
            tm.push_back(F.m);
        }
        bundles.push_back(b);
    }
    bsh_scene* s = nullptr;
    check(bsh_scene_create(ctx, v.data(), v.size() / 3, n.data(), n.size() / 3, tv.data(), tn.data(), tm.data(),
                           tm.size(), bundles.data(), bundles.size(), &s));
    try
    {
        check(bsh_scene_set_materials(s, d.materials.data(), d.materials.size()));
        check(bsh_scene_set_lights(s, d.area_lights.data(), d.area_lights.size()));
        if (!d.env_map.empty()) check(bsh_scene_set_env_map(s, d.env_map.data(), d.env_width, d.env_height));
        if (!d.other_lights.empty()) check(bsh_scene_set_other_lights(s, d.other_lights.data(), d.other_lights.size()));
        check(bsh_scene_set_camera(s, &d.camera));
        check(bsh_scene_commit(s, bvh));
    }
    catch (...)
    {
        bsh_scene_destroy(s);
        throw;
    }
    return s;
}

/* Renders with GiTracer(maxdepth) semantics on the GPU: image sampler <img_sampler>(img_grid),
   light and BSDF samplers <light_sampler>(1) and <brdf_sampler>(1) (BSH_SAMPLER_*), keyed by
   `seed`. The frame's per-pixel sums (bsh_render's accumulation: sum, n, sum of squares) stay
   readable through last_accum(). */
class HipImageRenderer : public Renderer
{
public:
    HipImageRenderer(ImagePlane& plane, bsh_context* ctx, bsh_scene* scene, int img_grid, int maxdepth, unsigned seed,
                     int img_sampler = BSH_SAMPLER_STRATIFIED, int light_sampler = BSH_SAMPLER_STRATIFIED,
                     int brdf_sampler = BSH_SAMPLER_STRATIFIED, int flags = 0)
        : plane_(plane), ctx_(ctx), scene_(scene), grid_(img_grid), depth_(maxdepth), seed_(seed),
          img_smp_(img_sampler), light_smp_(light_sampler), brdf_smp_(brdf_sampler), flags_(flags)
    {
    }

    void Render(World const& world) const
    {
        const int2 res = plane_.resolution();
        bsh_render_params p;
        std::memset(&p, 0, sizeof(p));
        p.width = res.x;
        p.height = res.y;
        p.img_grid = grid_;
        p.max_depth = depth_;
        p.seed = seed_;
        p.bg[0] = world.bgcolor_.x;
        p.bg[1] = world.bgcolor_.y;
        p.bg[2] = world.bgcolor_.z;
        p.flags = flags_;
        p.img_sampler = img_smp_;
        p.light_sampler = light_smp_;
        p.brdf_sampler = brdf_smp_;
        const size_t px = (size_t)res.x * res.y;
        void* accum = nullptr;
        check(bsh_device_alloc(ctx_, px * 8 * sizeof(float), &accum));
        bsh_status st = bsh_render(scene_, &p, (float*)accum, nullptr, nullptr);
        if (st == BSH_OK) st = bsh_device_sync(ctx_);
        if (st == BSH_OK)
        {
            accum_.resize(px * 8);
            st = bsh_copy_to_host(ctx_, accum_.data(), accum, accum_.size() * sizeof(float));
        }
        bsh_device_free(ctx_, accum);
        check(st);
        plane_.Prepare();
        for (int y = 0; y < res.y; ++y)
            for (int x = 0; x < res.x; ++x)
            {
                float const* a = &accum_[8 * ((size_t)y * res.x + x)];
                plane_.AddSample(float2(x + 0.5f, y + 0.5f), float3(a[0] / a[3], a[1] / a[3], a[2] / a[3]));
            }
        plane_.Finalize();
    }

    std::vector<float> const& last_accum() const { return accum_; }

private:
    ImagePlane& plane_;
    bsh_context* ctx_;
    bsh_scene* scene_;
    int grid_, depth_;
    unsigned seed_;
    int img_smp_, light_smp_, brdf_smp_, flags_;
    mutable std::vector<float> accum_;
};

/* World::accel_ over the GPU. Per-ray calls are kept for the reference's per-ray tracers; each
   is one launch, so batch callers use IntersectBatch / OccludedBatch. The hit record is the
   reference's own: Mesh::FillHit from the returned primitive and barycentrics, which are
   Mesh::IntersectFace's (mesh.cpp:7-45). */
class HipAccelerator : public Intersectable
{
public:
    HipAccelerator(bsh_context* ctx, bsh_scene* s, World const& w) : ctx_(ctx), s_(s), w_(w)
    {
        size_t first = 0;
        for (auto const& sb : w_.shapebundles_)
        {
            starts_.push_back(first);
            first += sb->GetNumShapes();
        }
        starts_.push_back(first);
    }

    // n rays (bsh_ray, host) -> n hits (host), one launch; staged through device buffers.
    void IntersectBatch(bsh_ray const* rays, size_t n, bsh_hit* hits) const
    {
        void *dr = nullptr, *dh = nullptr;
        check(bsh_device_alloc(ctx_, n * sizeof(bsh_ray), &dr));
        bsh_status st = bsh_device_alloc(ctx_, n * sizeof(bsh_hit), &dh);
        if (st == BSH_OK) st = bsh_copy_to_device(ctx_, dr, rays, n * sizeof(bsh_ray));
        if (st == BSH_OK) st = bsh_intersect_closest(s_, (bsh_ray const*)dr, n, (bsh_hit*)dh, nullptr);
        if (st == BSH_OK) st = bsh_device_sync(ctx_);
        if (st == BSH_OK) st = bsh_copy_to_host(ctx_, hits, dh, n * sizeof(bsh_hit));
        bsh_device_free(ctx_, dr);
        bsh_device_free(ctx_, dh);
        check(st);
    }

    // n rays -> n occlusion flags (nonzero: something lies within the ray's interval).
    void OccludedBatch(bsh_ray const* rays, size_t n, int32_t* occluded) const
    {
        void *dr = nullptr, *dq = nullptr;
        check(bsh_device_alloc(ctx_, n * sizeof(bsh_ray), &dr));
        bsh_status st = bsh_device_alloc(ctx_, n * sizeof(int32_t), &dq);
        if (st == BSH_OK) st = bsh_copy_to_device(ctx_, dr, rays, n * sizeof(bsh_ray));
        if (st == BSH_OK) st = bsh_intersect_any(s_, (bsh_ray const*)dr, n, (int32_t*)dq, nullptr);
        if (st == BSH_OK) st = bsh_device_sync(ctx_);
        if (st == BSH_OK) st = bsh_copy_to_host(ctx_, occluded, dq, n * sizeof(int32_t));
        bsh_device_free(ctx_, dr);
        bsh_device_free(ctx_, dq);
        check(st);
    }

    bool Intersect(ray const& r, ShapeBundle::Hit& hit) const
    {
        // World::Intersect seeds hit.t with r.t.y (world.cpp:21): the query's tmax
        bsh_ray q = {r.o.x, r.o.y, r.o.z, r.t.x, r.d.x, r.d.y, r.d.z, hit.t};
        bsh_hit h;
        IntersectBatch(&q, 1, &h);
        if (h.prim < 0) return false;
        std::pair<size_t, size_t> bf = locate(h.prim);
        static_cast<Mesh const&>(*w_.shapebundles_[bf.first]).FillHit(bf.second, h.t, h.b1, h.b2, hit);
        return true;
    }

    bool Intersect(ray const& r) const
    {
        bsh_ray q = {r.o.x, r.o.y, r.o.z, r.t.x, r.d.x, r.d.y, r.d.z, r.t.y};
        int32_t occ = 0;
        OccludedBatch(&q, 1, &occ);
        return occ != 0;
    }

    // Global primitive -> (bundle, face): the bundles are concatenated in World order.
    std::pair<size_t, size_t> locate(int prim) const
    {
        size_t lo = 0, hi = starts_.size() - 1;
        while (hi - lo > 1)
        {
            size_t mid = (lo + hi) / 2;
            if (starts_[mid] <= (size_t)prim) lo = mid;
            else hi = mid;
        }
        return {lo, (size_t)prim - starts_[lo]};
    }

private:
    bsh_context* ctx_;
    bsh_scene* s_;
    World const& w_;
    std::vector<size_t> starts_;
};

}  // namespace bsh_world
