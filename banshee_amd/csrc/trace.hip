// Ray-query kernels: the Intersectable slot (closest / any hit) over caller ray buffers.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace bsh {

constexpr int kTraceBlock = 256;

template <bool kAny, bool kCount>
__global__ __launch_bounds__(kTraceBlock) void k_intersect(SceneDev s, const bsh_ray* __restrict__ rays, uint32_t n,
                                                           bsh_hit* __restrict__ hits, int32_t* __restrict__ occ,
                                                           uint32_t* __restrict__ visits)
{
    // whole stack in LDS: 48 entries x 4 B x 256 threads = 48 KB, enough for either layout
    __shared__ uint32_t stk[kMaxStack4 * kTraceBlock];
    uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    const float4* r4 = reinterpret_cast<const float4*>(rays);
    float4 a = r4[2 * i], b = r4[2 * i + 1];
    TraceResult res;
    TraceCounts cnt{0u, 0u};
    bool h = s.width == 4 ? traverse4<kAny, kCount, kMaxStack4>(s, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, b.w,
                                                                 stk + threadIdx.x, kTraceBlock, res, cnt, nullptr, 0)
                          : traverse<kAny, kCount>(s, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, b.w,
                                                   stk + threadIdx.x, kTraceBlock, res, cnt);
    if (kAny && occ) occ[i] = h ? 1 : 0;
    if (hits)
    {
        bsh_hit o;
        o.prim = h ? res.prim : -1;
        o.t = h ? res.t : 0.f;
        o.b1 = h ? res.b1 : 0.f;
        o.b2 = h ? res.b2 : 0.f;
        hits[i] = o;
    }
    if (kCount && visits)
    {
        visits[2 * i] = cnt.nodes;
        visits[2 * i + 1] = cnt.tris;
    }
}

bsh_status launch_intersect(bsh_scene* s, const bsh_ray* rays, size_t n, int any, bsh_hit* hits, int32_t* occ,
                            uint32_t* visits, hipStream_t st)
{
    if (!s) return fail(BSH_ERR_INVALID, "intersect: NULL scene");
    if (!s->committed) return fail(BSH_ERR_STATE, "intersect: scene not committed");
    if (n == 0) return BSH_OK;
    if (!rays || (!hits && !occ)) return fail(BSH_ERR_INVALID, "intersect: NULL buffer");
    if (n > 0xffffffffull - kTraceBlock) return fail(BSH_ERR_INVALID, "intersect: too many rays");
    BSH_HIP(hipSetDevice(s->ctx->device));
    dim3 grid((uint32_t)((n + kTraceBlock - 1) / kTraceBlock));
    SceneDev d = s->dev();
    if (visits)
    {
        if (any) hipLaunchKernelGGL((k_intersect<true, true>), grid, dim3(kTraceBlock), 0, st, d, rays, (uint32_t)n, hits, occ, visits);
        else hipLaunchKernelGGL((k_intersect<false, true>), grid, dim3(kTraceBlock), 0, st, d, rays, (uint32_t)n, hits, occ, visits);
    }
    else
    {
        if (any) hipLaunchKernelGGL((k_intersect<true, false>), grid, dim3(kTraceBlock), 0, st, d, rays, (uint32_t)n, hits, occ, visits);
        else hipLaunchKernelGGL((k_intersect<false, false>), grid, dim3(kTraceBlock), 0, st, d, rays, (uint32_t)n, hits, occ, visits);
    }
    BSH_HIP(hipGetLastError());
    return BSH_OK;
}

}  // namespace bsh
