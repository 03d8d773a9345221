// Binned-SAH BVH2 builder for the GPU traversal layout (see bvh_build.h).
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <thread>

namespace bsh {
namespace {

// GpuNode4 -> GpuNodeQ, node for node (refs and counts unchanged). Per node and axis the grid
// step 2^e is the smallest power of two with 254 steps covering the children's extent and at
// least 2^-17 of the scene's largest coordinate (the root node's bounds), so a plane moved
// outward by half a step keeps a margin far above the slab test's rounding error at any
// coordinate in the scene. Lower planes round down to floor(x - 1/2) (clamped at the node's
// origin, which is a child's own padded bound), upper planes up to ceil(x + 1/2) <= 255.
// Returns false (no quantised layout) when an exponent leaves [-100, 60], where 2^e times an
// inverse direction could leave the normal float range.
bool quantize4(const std::vector<GpuNode4>& in, std::vector<GpuNodeQ>& out)
{
    out.assign(in.size(), GpuNodeQ{});
    if (in.empty()) return false;
    double mag = 0.0;
    {
        const GpuNode4& r = in[0];
        for (int c = 0; c < 4; ++c)
            if (r.cnt[c] >= 0)
                for (float v : {r.lox[c], r.hix[c], r.loy[c], r.hiy[c], r.loz[c], r.hiz[c]})
                    mag = std::max(mag, std::fabs((double)v));
    }
    const double floor_step = std::max(mag * std::ldexp(1.0, -17), std::ldexp(1.0, -100));
    for (size_t i = 0; i < in.size(); ++i)
    {
        const GpuNode4& g = in[i];
        GpuNodeQ& q = out[i];
        const float* lo[3] = {g.lox, g.loy, g.loz};
        const float* hi[3] = {g.hix, g.hiy, g.hiz};
        uint32_t* qlo[3] = {&q.qlox, &q.qloy, &q.qloz};
        uint32_t* qhi[3] = {&q.qhix, &q.qhiy, &q.qhiz};
        q.exps = 0;
        q.cnts = 0;
        q.pad = 0;
        for (int c = 0; c < 4; ++c)
        {
            q.ref[c] = g.ref[c];
            q.cnts |= (uint32_t)(uint8_t)(int8_t)g.cnt[c] << (8 * c);
        }
        for (int a = 0; a < 3; ++a)
        {
            float o = std::numeric_limits<float>::max(), h = -std::numeric_limits<float>::max();
            for (int c = 0; c < 4; ++c)
                if (g.cnt[c] >= 0)
                {
                    o = std::min(o, lo[a][c]);
                    h = std::max(h, hi[a][c]);
                }
            if (!(o <= h)) return false;   // a node has at least two children
            const double ext = (double)h - (double)o;
            int e;
            std::frexp(std::max(ext / 254.0, floor_step), &e);   // 2^(e-1) <= x < 2^e
            if (std::ldexp(1.0, e - 1) >= std::max(ext / 254.0, floor_step)) --e;
            if (e < -100 || e > 60) return false;
            const double step = std::ldexp(1.0, e);
            q.origin[a] = o;
            q.exps |= (uint32_t)(uint8_t)(int8_t)e << (8 * a);
            uint32_t wl = 0, wh = 0;
            for (int c = 0; c < 4; ++c)
            {
                uint32_t l = 255, u = 0;   // empty child: an inverted box (its count is -1 too)
                if (g.cnt[c] >= 0)
                {
                    const double xl = ((double)lo[a][c] - (double)o) / step;
                    const double xh = ((double)hi[a][c] - (double)o) / step;
                    l = (uint32_t)std::max(0.0, std::floor(xl - 0.5));
                    u = (uint32_t)std::ceil(xh + 0.5);
                    if (u > 255u) return false;   // ext <= 254 steps makes this unreachable
                }
                wl |= l << (8 * c);
                wh |= u << (8 * c);
            }
            *qlo[a] = wl;
            *qhi[a] = wh;
        }
    }
    return true;
}

struct Aabb
{
    float lo[3], hi[3];
    void reset()
    {
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = std::numeric_limits<float>::max();
            hi[k] = -std::numeric_limits<float>::max();
        }
    }
    void grow(const Aabb& b)
    {
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float* p)
    {
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    float area() const
    {
        float e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
        if (e0 < 0.f || e1 < 0.f || e2 < 0.f) return 0.f;
        return 2.f * (e0 * e1 + e0 * e2 + e1 * e2);
    }
};

struct TNode
{
    Aabb box;
    int left = -1, right = -1;
    int first = 0, count = 0;
};

int ceil_log2(int n)
{
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

class Builder
{
public:
    Builder(const float* tv, size_t n, const BuildParams& p) : tv_(tv), n_(n), p_(p)
    {
        if (p_.max_leaf < 1) p_.max_leaf = 1;
        if (p_.max_leaf > 8) p_.max_leaf = 8;
        if (p_.bins < 4) p_.bins = 4;
        if (p_.bins > 64) p_.bins = 64;
        threads_ = p_.threads > 0 ? p_.threads : (int)std::max(1u, std::thread::hardware_concurrency());
    }

    void run(BuildOutput& out)
    {
        tb_.resize(n_);
        cen_.resize(3 * n_);
        idx_.resize(n_);
        Aabb root, croot;
        root.reset();
        croot.reset();
        for (size_t i = 0; i < n_; ++i)
        {
            Aabb b;
            b.reset();
            b.grow(tv_ + 9 * i);
            b.grow(tv_ + 9 * i + 3);
            b.grow(tv_ + 9 * i + 6);
            tb_[i] = b;
            for (int k = 0; k < 3; ++k) cen_[3 * i + k] = 0.5f * (b.lo[k] + b.hi[k]);
            idx_[i] = (int)i;
            root.grow(b);
            croot.grow(&cen_[3 * i]);
        }
        nodes_.resize(std::max<size_t>(1, 2 * n_));
        ncount_ = 0;
        int r = -1;
        if (n_ > 0) r = build(0, (int)n_, root, croot, 0);
        out.width = 2;
        if (p_.width == 4 && r >= 0 && nodes_[r].left >= 0)
        {
            flatten4(r, out);
            if (3 * out.max_depth + 1 <= kMaxStack4)
            {
                out.width = 4;
                out.quantized = p_.quantized >= 0 && quantize4(out.nodes4, out.nodesq);
                return;
            }
            out.nodes4.clear();
        }
        flatten(r, root, out);
    }

private:
    int alloc() { return ncount_.fetch_add(1); }

    void bounds(int first, int count, Aabb& box, Aabb& cbox) const
    {
        box.reset();
        cbox.reset();
        for (int i = first; i < first + count; ++i)
        {
            box.grow(tb_[idx_[i]]);
            cbox.grow(&cen_[3 * idx_[i]]);
        }
    }

    int build(int first, int count, const Aabb& box, const Aabb& cbox, int depth)
    {
        int id = alloc();
        TNode& nd = nodes_[id];
        nd.box = box;
        nd.first = first;
        nd.count = count;
        if (count == 1) return id;

        const float ci = 1.0f, ct = 1.0f;
        float ext[3] = {cbox.hi[0] - cbox.lo[0], cbox.hi[1] - cbox.lo[1], cbox.hi[2] - cbox.lo[2]};
        bool degenerate = !(ext[0] > 0.f) && !(ext[1] > 0.f) && !(ext[2] > 0.f);
        bool forced = depth + ceil_log2(count) >= kMaxStack - 2;

        int mid = -1;
        if (!degenerate && !forced)
        {
            const int B = p_.bins;
            struct Bin { Aabb b; int n; };
            Bin bins[3][64];
            for (int a = 0; a < 3; ++a)
                for (int i = 0; i < B; ++i) { bins[a][i].b.reset(); bins[a][i].n = 0; }
            float scale[3];
            for (int a = 0; a < 3; ++a) scale[a] = ext[a] > 0.f ? (float)B / ext[a] : 0.f;
            for (int i = first; i < first + count; ++i)
            {
                int t = idx_[i];
                for (int a = 0; a < 3; ++a)
                {
                    if (!(ext[a] > 0.f)) continue;
                    int bi = (int)((cen_[3 * t + a] - cbox.lo[a]) * scale[a]);
                    bi = std::min(std::max(bi, 0), B - 1);
                    bins[a][bi].b.grow(tb_[t]);
                    bins[a][bi].n++;
                }
            }
            float best = std::numeric_limits<float>::max();
            int best_axis = -1, best_bin = -1;
            float inv_area = 1.f / std::max(box.area(), 1e-30f);
            for (int a = 0; a < 3; ++a)
            {
                if (!(ext[a] > 0.f)) continue;
                float rarea[64];
                int rcnt[64];
                Aabb acc;
                acc.reset();
                int c = 0;
                for (int i = B - 1; i > 0; --i)
                {
                    acc.grow(bins[a][i].b);
                    c += bins[a][i].n;
                    rarea[i] = acc.area();
                    rcnt[i] = c;
                }
                acc.reset();
                c = 0;
                for (int i = 0; i < B - 1; ++i)
                {
                    acc.grow(bins[a][i].b);
                    c += bins[a][i].n;
                    if (c == 0 || rcnt[i + 1] == 0) continue;
                    float cost = ct + ci * (c * acc.area() + rcnt[i + 1] * rarea[i + 1]) * inv_area;
                    if (cost < best)
                    {
                        best = cost;
                        best_axis = a;
                        best_bin = i;
                    }
                }
            }
            float leaf_cost = ci * (float)count;
            if (count <= p_.max_leaf && (best_axis < 0 || leaf_cost <= best)) return id;
            if (best_axis >= 0)
            {
                const int a = best_axis;
                const float lo = cbox.lo[a], sc = scale[a];
                const int bb = best_bin;
                auto it = std::partition(idx_.begin() + first, idx_.begin() + first + count, [&](int t) {
                    int bi = (int)((cen_[3 * t + a] - lo) * sc);
                    bi = std::min(std::max(bi, 0), B - 1);
                    return bi <= bb;
                });
                mid = (int)(it - idx_.begin());
                if (mid == first || mid == first + count) mid = -1;
            }
        }
        else if (count <= p_.max_leaf)
        {
            return id;
        }
        if (mid < 0)
        {
            // object median on the widest centroid axis (or by index when all centroids coincide)
            mid = first + count / 2;
            int a = 0;
            if (ext[1] > ext[a]) a = 1;
            if (ext[2] > ext[a]) a = 2;
            if (!degenerate)
                std::nth_element(idx_.begin() + first, idx_.begin() + mid, idx_.begin() + first + count,
                                 [&](int x, int y) { return cen_[3 * x + a] < cen_[3 * y + a]; });
        }
        Aabb lb, lcb, rb, rcb;
        bounds(first, mid - first, lb, lcb);
        bounds(mid, first + count - mid, rb, rcb);
        int l, r;
        if (count > 65536 && active_.load() < threads_)
        {
            active_++;
            auto fut = std::async(std::launch::async, [&]() { return build(first, mid - first, lb, lcb, depth + 1); });
            r = build(mid, first + count - mid, rb, rcb, depth + 1);
            l = fut.get();
            active_--;
        }
        else
        {
            l = build(first, mid - first, lb, lcb, depth + 1);
            r = build(mid, first + count - mid, rb, rcb, depth + 1);
        }
        nodes_[id].left = l;
        nodes_[id].right = r;
        return id;
    }

    static void pad(const Aabb& b, float* lo, float* hi)
    {
        // Outward rounding by a relative 2^-20 keeps the slab test conservative for
        // triangles lying in a box face (axis-aligned walls).
        for (int k = 0; k < 3; ++k)
        {
            float m = std::max(std::fabs(b.lo[k]), std::fabs(b.hi[k]));
            float e = m * 9.5367431640625e-7f + 1e-30f;
            lo[k] = b.lo[k] - e;
            hi[k] = b.hi[k] + e;
        }
    }

    struct ChildRef { int ref; int count; Aabb box; };

    void flatten(int root, const Aabb& rootbox, BuildOutput& out)
    {
        out.nodes.clear();
        out.tri_order.assign(idx_.begin(), idx_.end());
        out.max_depth = 0;
        ChildRef empty;
        empty.ref = 0;
        empty.count = -1;
        empty.box.reset();
        if (root < 0 || nodes_[root].left < 0)
        {
            // root is a leaf (or the scene is empty): wrap it in one internal node
            out.nodes.resize(1);
            ChildRef c0 = empty;
            if (root >= 0) { c0.ref = nodes_[root].first; c0.count = nodes_[root].count; c0.box = rootbox; }
            write(out.nodes[0], c0, empty);
            return;
        }
        out.nodes.reserve(ncount_.load());
        emit(root, 0, out);
    }

    int emit(int t, int depth, BuildOutput& out)
    {
        out.max_depth = std::max(out.max_depth, depth);
        int me = (int)out.nodes.size();
        out.nodes.emplace_back();
        const TNode& nd = nodes_[t];
        ChildRef c[2];
        int kids[2] = {nd.left, nd.right};
        for (int k = 0; k < 2; ++k)
        {
            const TNode& ch = nodes_[kids[k]];
            c[k].box = ch.box;
            if (ch.left < 0)
            {
                c[k].ref = ch.first;
                c[k].count = ch.count;
            }
            else
            {
                c[k].ref = emit(kids[k], depth + 1, out);
                c[k].count = 0;
            }
        }
        write(out.nodes[me], c[0], c[1]);
        return me;
    }

    void flatten4(int root, BuildOutput& out)
    {
        out.nodes4.clear();
        out.tri_order.assign(idx_.begin(), idx_.end());
        out.max_depth = 0;
        out.nodes4.reserve(ncount_.load() / 2 + 1);
        emit4(root, 0, out);
    }

    // Children of a 4-wide node: start from the BVH2 node's two children and open the
    // internal one of largest area until there are four or only leaves remain.
    int emit4(int t, int depth, BuildOutput& out)
    {
        out.max_depth = std::max(out.max_depth, depth);
        int me = (int)out.nodes4.size();
        out.nodes4.emplace_back();
        int kids[4] = {nodes_[t].left, nodes_[t].right, -1, -1};
        int nk = 2;
        while (nk < 4)
        {
            int best = -1;
            float ba = -1.f;
            for (int k = 0; k < nk; ++k)
                if (nodes_[kids[k]].left >= 0 && nodes_[kids[k]].box.area() > ba)
                {
                    ba = nodes_[kids[k]].box.area();
                    best = k;
                }
            if (best < 0) break;
            int open = kids[best];
            kids[best] = nodes_[open].left;
            kids[nk++] = nodes_[open].right;
        }
        ChildRef c[4];
        for (int k = 0; k < 4; ++k)
        {
            if (k >= nk)
            {
                c[k].ref = 0;
                c[k].count = -1;
                c[k].box.reset();
                continue;
            }
            const TNode& ch = nodes_[kids[k]];
            c[k].box = ch.box;
            if (ch.left < 0)
            {
                c[k].ref = ch.first;
                c[k].count = ch.count;
            }
            else
            {
                c[k].ref = emit4(kids[k], depth + 1, out);
                c[k].count = 0;
            }
        }
        GpuNode4& g = out.nodes4[me];
        for (int k = 0; k < 4; ++k)
        {
            float lo[3], hi[3];
            if (c[k].count < 0) { for (int a = 0; a < 3; ++a) { lo[a] = 1.f; hi[a] = -1.f; } }
            else pad(c[k].box, lo, hi);
            g.lox[k] = lo[0]; g.hix[k] = hi[0];
            g.loy[k] = lo[1]; g.hiy[k] = hi[1];
            g.loz[k] = lo[2]; g.hiz[k] = hi[2];
            g.ref[k] = c[k].ref;
            g.cnt[k] = c[k].count;
        }
        return me;
    }

    static void write(GpuNode& g, const ChildRef& a, const ChildRef& b)
    {
        float alo[3], ahi[3], blo[3], bhi[3];
        if (a.count < 0) { for (int k = 0; k < 3; ++k) { alo[k] = 1.f; ahi[k] = -1.f; } }
        else pad(a.box, alo, ahi);
        if (b.count < 0) { for (int k = 0; k < 3; ++k) { blo[k] = 1.f; bhi[k] = -1.f; } }
        else pad(b.box, blo, bhi);
        g.n0[0] = alo[0]; g.n0[1] = ahi[0]; g.n0[2] = alo[1]; g.n0[3] = ahi[1];
        g.n1[0] = blo[0]; g.n1[1] = bhi[0]; g.n1[2] = blo[1]; g.n1[3] = bhi[1];
        g.n2[0] = alo[2]; g.n2[1] = ahi[2]; g.n2[2] = blo[2]; g.n2[3] = bhi[2];
        g.n3[0] = a.ref; g.n3[1] = b.ref; g.n3[2] = a.count; g.n3[3] = b.count;
    }

    const float* tv_;
    size_t n_;
    BuildParams p_;
    int threads_;
    std::vector<Aabb> tb_;
    std::vector<float> cen_;
    std::vector<int> idx_;
    std::vector<TNode> nodes_;
    std::atomic<int> ncount_{0};
    std::atomic<int> active_{1};
};

}  // namespace

void build_bvh(const float* tri_verts, size_t ntris, const BuildParams& p, BuildOutput& out)
{
    auto t0 = std::chrono::steady_clock::now();
    Builder b(tri_verts, ntris, p);
    b.run(out);
    out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace bsh
