// Host-side BVH construction for the GPU layout.
//
// Replaces Bvh::Build/BuildImpl/BuildNode/FindSahSplit (accelerator/bvh.cpp:16-355):
// a binned-SAH build over triangle bounds, but with multi-triangle leaves, near-first
// child ordering decided at traversal time, and a hard depth cap so the traversal
// stack has a fixed size. The tree differs from the reference's; closest-hit results
// do not depend on the tree except for exact ties in t (which the reference resolves
// by test order, mesh.cpp:33).
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <vector>

namespace bsh {

// GPU node: 4 x 16 B.
//   n0 = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//   n1 = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   n2 = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   n3 = (c0.ref, c1.ref, c0.count, c1.count) as int32
// count == 0: ref is an internal node index; count > 0: ref is the first triangle
// of a leaf of `count` triangles; count < 0: empty child (never entered).
struct GpuNode
{
    float n0[4], n1[4], n2[4];
    int32_t n3[4];
};
static_assert(sizeof(GpuNode) == 64, "node is 64 B");

// 4-wide node, 128 B = 8 x float4: per child slab bounds, then refs and counts (same
// meaning as in GpuNode). Collapsed from the binned-SAH BVH2 by repeatedly opening the
// internal child of largest surface area, so one visit tests up to four boxes.
struct GpuNode4
{
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t ref[4];
    int32_t cnt[4];
};
static_assert(sizeof(GpuNode4) == 128, "node is 128 B");

// Quantised 4-wide node, 64 B = 4 x 16 B, the same tree as GpuNode4 with each child box stored
// as 8-bit planes on a per-node, per-axis power-of-two grid (after Ylitie, Karras and Laine,
// "Efficient incoherent ray traversal on GPUs through compressed wide BVHs", HPG 2017):
//   w0 = (origin.x, origin.y, origin.z, exponents as signed bytes: e_x | e_y << 8 | e_z << 16)
//   w1 = (qlo.x, qhi.x, qlo.y, qhi.y), each 4 bytes, byte c for child c
//   w2 = (qlo.z, qhi.z, counts as 4 signed bytes, 0)
//   w3 = refs (as in GpuNode4)
// Child c's box on axis a is [origin + qlo 2^e, origin + qhi 2^e]. The builder rounds every plane
// outward by at least half a grid step (a lower plane of 0 is the node's own padded bound), and
// the grid step is at least 2^-17 of the scene's largest coordinate, so the traversal's rounding
// of these planes stays inside the margin and the boxes contain the float boxes they replace.
// e stays within [-100, 60].
struct GpuNodeQ
{
    float origin[3];
    uint32_t exps;
    uint32_t qlox, qhix, qloy, qhiy, qloz, qhiz;
    uint32_t cnts;
    uint32_t pad;
    int32_t ref[4];
};
static_assert(sizeof(GpuNodeQ) == 64, "node is 64 B");

// Stack entries the traversal kernel reserves per lane. The builder caps the depth of
// internal nodes at kMaxStack - 2 so that the near-first traversal, which pushes at
// most one far child per internal ancestor, never needs more.
constexpr int kMaxStack = 32;
// The 4-wide traversal pushes at most 3 children per level; the builder falls back to the
// binary layout when 3 x (4-wide depth) + 1 would exceed this.
constexpr int kMaxStack4 = 48;

struct BuildParams
{
    int max_leaf = 4;
    int bins = 16;
    int threads = 0;
    int width = 4;      // 2: GpuNode, 4: GpuNode4 (the faster layout on the 3M-triangle terrain)
    // width 4 only: 0 (the default) or 1 GpuNodeQ, -1 GpuNode4. Quantised nodes halve the node
    // bytes at about 20 more instructions per visit; they traced faster on every workload
    // measured (3.08M-triangle frame +5 %, 64-scene batch +3.6 %, 2M-triangle probe bake +9.6 %).
    int quantized = 0;
};

struct BuildOutput
{
    std::vector<GpuNode> nodes;       // width 2
    std::vector<GpuNode4> nodes4;     // width 4
    std::vector<GpuNodeQ> nodesq;     // width 4, quantized (nodes4 is kept: the same tree in floats)
    int width = 2;                    // the layout actually built
    bool quantized = false;           // nodesq holds the tree
    std::vector<int32_t> tri_order;   // leaf order -> global primitive index
    int max_depth = 0;
    double build_ms = 0.0;
};

// tris: per triangle the three vertices, 9 floats, global primitive order.
void build_bvh(const float* tri_verts, size_t ntris, const BuildParams& p, BuildOutput& out);

}  // namespace bsh
