// Host-side objects behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include <mutex>

#include "../../include/banshee_hip.h"
#include "bsh_math.h"
#include "bvh_build.h"
#include "material_dev.h"
#include "sampler_dev.h"
#include "scene_dev.h"


namespace bsh {

void set_error(const std::string& msg);
bsh_status fail(bsh_status code, const std::string& msg);
bsh_status hip_fail(hipError_t e, const char* what);
// BSH_OK, or BSH_ERR_INVALID naming `who` and the field out of range.
bsh_status validate_tuning(const bsh_tuning& t, const char* who);

// Scratch of bsh_sh_project_env_noalloc: 4 MiB per device (tables and block partials).
constexpr size_t kEnvScratchFloats = (size_t)1 << 20;
constexpr int kMaxDevices = 64;
// Looks up the device address of the scratch array on `device` (once per device).
bsh_status env_scratch_init(int device);
// That address, or NULL before env_scratch_init succeeded on `device`.
float* env_scratch_ptr(int device);

#define BSH_HIP(call)                                        \
    do                                                       \
    {                                                        \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return bsh::hip_fail(e_, #call); \
    } while (0)

template <class T>
struct DevBuf
{
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count)
    {
        if (count <= n && p) return hipSuccess;
        release();
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) { p = nullptr; n = 0; return e; }
        n = count;
        return hipSuccess;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

// Camera frame precomputed on the host exactly as the PerscpectiveCamera ctor does
// (camera/perspective_camera.cpp:5-23).
struct CameraDev
{
    float eye[3], fwd[3], right[3], up[3];
    float znear, zfar, dimx, dimy;
};
CameraDev make_camera(const bsh_camera& c);

// Wavefront render buffers, sized for `paths` in-flight camera samples.
struct RenderWork
{
    size_t paths = 0;
    int slots = 0;   // NEE slots per path = 2 * lights at bounce 0
    DevBuf<float4> ray_o, ray_d, thr, thr_nee, rad, out;
    DevBuf<int4> hit;
    DevBuf<uint4> meta;
    DevBuf<unsigned long long> nee_mask;   // per path: NEE slots with a ray in flight (bit 2j shadow, 2j+1 MIS)
    DevBuf<float4> nee_a, nee_b, nee_wi;
    DevBuf<int4> nee_res;
    DevBuf<float4> qc_o[2], qc_d[2], qa_o[2], qa_d[2];   // closest / any-hit queues
    DevBuf<int> qc_own[2], qa_own[2];
    DevBuf<float4> gin, rec;          // guide network inputs (2 per row), training records (per path)
    DevBuf<float4> pdir;              // guided probe jobs: per path, probe ray direction and weight
    DevBuf<_Float16> gout;            // guide network outputs (kGUsed per row)
    DevBuf<int> grow;                 // per path: its guide row this pass
    DevBuf<uint32_t> gseg;            // guide rows per block segment this pass
    DevBuf<uint32_t> rec_count;       // records written so far (running over the chunks)
    DevBuf<uint32_t> rec_blk, rec_off; // per block of paths: records, output offset
    DevBuf<uint32_t> alive[2];     // paths still live after a shade, listed for the next bounce
    DevBuf<uint32_t> counters;     // per launch: count, work, (pad) x 4; then the live-path counts
    DevBuf<unsigned long long> stats;  // closest, any
    DevBuf<uint32_t> stk_ovf;          // deep traversal-stack entries of the trace kernel
};


// One job of a batch: a camera frame (type 0) or a probe set (type 1), with the device
// scene it runs on. Copied to the device as the batch's job table.
struct JobDev
{
    SceneDev s;
    CameraDev cam;
    int type;
    int W, H, cx, cy, cw, ch, grid, spp, s_begin, s_count;
    int img_smp, light_smp, brdf_smp;   // SMP_* kinds (sampler_dev.h); grid is the image sampler's argument
    int depth;
    uint32_t seed;
    float bg[3];
    int accumulate;
    float* accum;
    float* samples;
    const float* probes;
    int nprobes, N, g, lmax;
    float tmin;
    float* coeffs;
    float* radiance;
    float* ray_dirs;       // probe jobs: optional per-ray (direction, projection weight)
    int flags;             // JOB_NO_QUIRKS, JOB_GUIDED, JOB_RECORD, JOB_PROBE_DIR
    float bsdf_frac;       // guided: probability of sampling the BSDF
    int guide_bounces;     // guided: vertices 0 .. guide_bounces-1 use the guide (0 = all)
    float glo[3], ginv[3]; // guided: scene box for the network's position input
};

// JOB_PROBE_DIR: a guided probe job whose probe rays are drawn from the uniform/guide mixture
// (the probe position is the path's vertex -1; its training record takes record slot 0).
enum : int { JOB_NO_QUIRKS = 1, JOB_GUIDED = 2, JOB_RECORD = 4, JOB_PROBE_DIR = 8 };

struct GuideRun
{
    bsh_guide* guide = nullptr;   // network used by JOB_GUIDED jobs
    float* records = nullptr;     // training records out (12 floats each)
    size_t cap = 0;
    size_t written = 0;
    int bounces = 0;              // guided vertices per path (0 = all)
};
bsh_status guide_forward(bsh_guide* g, const float4* in, int stride, uint32_t n, float* out, hipStream_t st);
// Inference for the wavefront: rows of 2-float4 inputs in segments of kGuideSeg rows, of
// which the first seg_counts[s] (device) are live; writes the kGUsed used outputs per row in
// FP16. Tiles past a segment's count are skipped.
constexpr int kGuideSeg = 256;
bsh_status guide_forward_rows(bsh_guide* g, const float4* in, const uint32_t* seg_counts, uint32_t rows,
                              _Float16* out, hipStream_t st);
float guide_bsdf_fraction(const bsh_guide* g);
int guide_bounces(const bsh_guide* g);

bsh_status run_jobs(bsh_context* c, std::vector<JobDev>& jobs, bool count, const bsh_tuning* tune, int use_graph,
                    hipStream_t st, GuideRun* gr = nullptr);
}  // namespace bsh

struct bsh_context
{
    int device = 0;
    int num_cus = 0;
    size_t lds_per_cu = 0;
    std::mutex mu;                 // serialises work on the shared buffers below
    bsh::RenderWork work;
    // further wavefront lanes: a run of several chunks deals them round-robin over the lanes,
    // each extra lane with its own buffers and stream, so one chunk's kernels fill another's
    // trace drain (measured: two lanes 7.7 % on a 1080p frame; a third added nothing there
    // and 1.7 % on a 64-scene batch, for another set of buffers)
    static constexpr int kLanes = 2;
    bsh::RenderWork lane_work[kLanes - 1];
    hipStream_t lane_stream[kLanes - 1] = {};
    // the last batch graph (bsh_batch_run with use_graph), replayed while its key matches
    hipStream_t graph_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint8_t> graph_key;
    void drop_graph()
    {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_key.clear();
    }
    ~bsh_context()
    {
        drop_graph();
        if (graph_stream) (void)hipStreamDestroy(graph_stream);
        for (hipStream_t s : lane_stream)
            if (s) (void)hipStreamDestroy(s);
    }
    bsh::DevBuf<bsh::JobDev> jobs;
    bsh_tuning tuning{1 << 23, 0, 0, 0, 0, 0};
    bsh_render_stats stats{};
    std::vector<bsh_launch_stats> launch_stats;   // per bounce launch, of the last run
};

struct bsh_scene
{
    bsh_context* ctx = nullptr;
    std::vector<float> verts, normals;
    std::vector<int32_t> tri_v, tri_n, tri_m;
    std::vector<bsh_bundle> bundles;
    std::vector<bsh_material> mats;
    std::vector<bsh_area_light> lights;
    std::vector<bsh_light> others;        // after the area lights in the light list
    std::vector<float> xforms;            // per bundle: m then minv, row-major 4x4 (empty: identities)
    std::vector<float> uvs;               // texture coordinates, (u, v) pairs
    std::vector<int32_t> tri_uv;          // per triangle: three indices into uvs (empty: no UVs)
    std::vector<float> tex_texels;        // texture maps' texels
    std::vector<bsh_texture> tex_desc;
    std::vector<int32_t> mat_maps;        // per material row: colour map, kr map (empty: none)
    std::vector<int32_t> nmaps;           // per material row: normal map, -1: none (empty: none)
    std::vector<float> env_tex;           // env_h rows of env_w RGB texels
    int env_w = 0, env_h = 0;
    bsh_camera cam{};
    bool has_cam = false;
    bool committed = false;

    bsh::DevBuf<float4> d_nodes;   // GpuNode (width 2), GpuNode4 or GpuNodeQ (width 4) as float4s
    bsh::DevBuf<float4> d_tris, d_stri;
    bsh::DevBuf<bsh::LightDev> d_lights;
    bsh::DevBuf<float> d_light_cdf;
    bsh::DevBuf<float> d_env;             // map, then the environment distribution's tables
    bsh::DevBuf<float4> d_mats;           // material rows, 2 float4 each
    bsh::DevBuf<float4> d_tri_uv;
    bsh::DevBuf<float> d_tex;
    bsh::DevBuf<int4> d_tex_desc;
    bsh::DevBuf<int2> d_mat_maps;
    bsh::DevBuf<int> d_nmaps;
    int ext_mats = 0;
    bsh::EnvDev env{};
    std::vector<bsh::LightDev> h_lights;
    float box_lo[3] = {0.f, 0.f, 0.f}, box_hi[3] = {0.f, 0.f, 0.f};
    uint64_t nnodes = 0;
    int width = 2;
    bool quantized = false;        // width 4 in GpuNodeQ records
    int max_depth = 0;
    double build_ms = 0.0;


    bsh::SceneDev dev() const
    {
        bsh::SceneDev s;
        s.nodes = d_nodes.p;
        s.width = width;
        s.qnodes = quantized ? 1 : 0;
        s.tris = d_tris.p;
        s.stri = d_stri.p;
        s.lights = d_lights.p;
        s.light_cdf = d_light_cdf.p;
        s.env = env;
        s.nlights = (int)(lights.size() + others.size());
        s.narea = (int)lights.size();
        s.mats = d_mats.p;
        s.ext_mats = ext_mats;
        s.tri_uv = tri_uv.empty() ? nullptr : d_tri_uv.p;
        s.tex = d_tex.p;
        s.tex_desc = d_tex_desc.p;
        s.mat_maps = mat_maps.empty() ? nullptr : d_mat_maps.p;
        s.nmaps = nmaps.empty() ? nullptr : d_nmaps.p;
        s.ntris = (int)tri_v.size() / 3;
        return s;
    }
};

namespace bsh {
bsh_status launch_intersect(bsh_scene* s, const bsh_ray* rays, size_t n, int any, bsh_hit* hits, int32_t* occ,
                            uint32_t* visits, hipStream_t st);
bsh_status make_render_job(bsh_scene* s, const bsh_render_params* p, float* accum, float* samples, JobDev& J);
bsh_status make_probe_job(bsh_scene* s, const float* probes, size_t nprobes, const bsh_probe_params* p,
                          float* coeffs, float* radiance, JobDev& J);
}  // namespace bsh
