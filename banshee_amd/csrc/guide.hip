// Learned path guiding: FP16-weight / FP32-accumulate MFMA kernels for the guide MLP
// (inference, training forward, backward, weight gradients, Adam). See guide_dev.h for the
// network and the vMF mixture. No reference counterpart.
//
// Matrix tiles use the CDNA operand layouts
//   v_mfma_f32_16x16x16_f16 (backward, weight gradients):
//     A (16x16, MxK): lane l holds A[l%16][4*(l/16) + i], i = 0..3
//     B (16x16, KxN): lane l holds B[4*(l/16) + i][l%16]
//     D (16x16):      lane l holds D[4*(l/16) + i][l%16]
//   v_mfma_f32_16x16x32_f16 (forward; twice the work per issue on gfx950):
//     A (16x32): lane l holds A[l%16][8*(l/16) + i], i = 0..7; B (32x16): B[8*(l/16) + i][l%16];
//     D as above
// (the first checked on the device by bsh_diag_mfma, both by the forward tests).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"
#include "guide_dev.h"

namespace bsh {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 hp2 __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));

// ReLU of an FP32 accumulator fragment, rounded to FP16: two paired conversions, then two
// packed FP16 maxima (max(round(x), 0) = round(max(x, 0)) for every x).
__device__ __forceinline__ h4 relu_h4(f4v a)
{
    const hp2 z = {(_Float16)0.f, (_Float16)0.f};
    const hp2 lo = __builtin_elementwise_max((hp2){(_Float16)a[0], (_Float16)a[1]}, z);
    const hp2 hi = __builtin_elementwise_max((hp2){(_Float16)a[2], (_Float16)a[3]}, z);
    return (h4){lo[0], lo[1], hi[0], hi[1]};
}

__device__ __forceinline__ f4v mfma16(h4 a, h4 b, f4v c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f4v mfma32(h8 a, h8 b, f4v c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ h8 cat8(h4 lo, h4 hi) { return (h8){lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]}; }

// ---- forward weight layout in FP16 (out-major rows Wt[n][k]) ----
// In LDS each row is padded to a 16-byte multiple, and within every 32-column block the
// input features are permuted (perm32) so that the 8 halves lane group g reads for a 16x32
// A fragment, columns 32 kb + 8g .. 8g+7, are the features 32 kb + 4g + 0..3 and
// 32 kb + 16 + 4g + 0..3: exactly what the lane holds of the two 16-feature output fragments
// 2 kb and 2 kb + 1 of the previous layer (or of bands 2 kb, 2 kb + 1 of the encoding), so
// the B operand is their concatenation.
constexpr int kPad = 8;
constexpr int kL1S = kGIn + kPad, kLHS = kGHid + kPad;             // LDS row strides (halves)
static_assert(kGIn % 32 == 0 && kGHid % 32 == 0, "16x32 tiles");
__host__ __device__ constexpr int perm32(int k)
{
    return (k & ~31) | (((k >> 2) & 3) << 3) | (((k >> 4) & 1) << 2) | (k & 3);
}
constexpr int kF1 = 0;                                             // 64 rows x 96
constexpr int kF2 = kF1 + kGHid * kGIn;                            // 64 x 64
constexpr int kF3 = kF2 + kGHid * kGHid;
constexpr int kF4 = kF3 + kGHid * kGHid;
constexpr int kF5 = kF4 + kGHid * kGHid;                           // 48 x 64
constexpr int kFTotal = kF5 + kGOut * kGHid;
// ---- backward layout (in-major rows W[n_in][k_out]) for layers 5..2 ----
constexpr int kB5 = 0;                                             // 64 rows x 48
constexpr int kB4 = kB5 + kGHid * kGOut;                           // 64 x 64
constexpr int kB3 = kB4 + kGHid * kGHid;
constexpr int kB2 = kB3 + kGHid * kGHid;
constexpr int kBTotal = kB2 + kGHid * kGHid;
constexpr int kBiasTotal = 4 * kGHid + kGOut;
// LDS image of the forward weights (padded rows)
constexpr int kLF1 = 0, kLF2 = kLF1 + kGHid * kL1S, kLF3 = kLF2 + kGHid * kLHS, kLF4 = kLF3 + kGHid * kLHS,
              kLF5 = kLF4 + kGHid * kLHS, kLFTotal = kLF5 + kGOut * kLHS;

struct GuideNet
{
    const _Float16* wf;   // forward FP16 weights (kFTotal)
    const _Float16* wb;   // backward FP16 weights (kBTotal)
    const float* bias;    // b1..b5 (kBiasTotal)
};

__device__ __forceinline__ void load_fwd_weights(const GuideNet& g, _Float16* lw, float* lb)
{
    // unpadded global rows -> padded, column-permuted LDS rows, 4 halves at a time (perm32
    // keeps runs of 4 consecutive features together)
    const h4* src = reinterpret_cast<const h4*>(g.wf);
    for (int u = threadIdx.x; u < kFTotal / 4; u += blockDim.x)
    {
        const int t = 4 * u;
        int dst;
        if (t < kF2) { int n = t / kGIn, k = t % kGIn; dst = kLF1 + n * kL1S + perm32(k); }
        else if (t < kF5)
        {
            int v = (t - kF2), l = v / (kGHid * kGHid), r = v % (kGHid * kGHid);
            int n = r / kGHid, k = r % kGHid;
            dst = kLF2 + l * kGHid * kLHS + n * kLHS + perm32(k);
        }
        else { int v = t - kF5; int n = v / kGHid, k = v % kGHid; dst = kLF5 + n * kLHS + perm32(k); }
        *reinterpret_cast<h4*>(&lw[dst]) = src[u];
    }
    for (int t = threadIdx.x; t < kBiasTotal; t += blockDim.x) lb[t] = g.bias[t];
}

// Encoded input of one row for the lane holding k-positions 4q..4q+3 of every 16-feature
// block (the B-fragment layout): block b holds band b, position p < 14 the sin (p even) or cos
// (p odd) of 2^b pi v_{p/2}, positions 14 and 15 the raw values v_{2b}, v_{2b+1} (zero past
// v6). So lane group q < 3 owns dimensions 2q and 2q+1 in every band and q = 3 dimension 6
// plus the raw slots. Band 0 uses the hardware sine/cosine on the exact revolution argument
// fract(v/2); each further band doubles the angle: sin 2x = 2 sin x cos x,
// cos 2x = (cos x - sin x)(cos x + sin x).
template <int B>
__device__ __forceinline__ float raw_slot(float4 a, float4 c)
{
    return B == 0 ? a.x : B == 1 ? a.y : B == 2 ? a.z : B == 3 ? a.w : B == 4 ? c.x : B == 5 ? c.y : B == 6 ? c.z : 0.f;
}

template <int BAND>
__device__ __forceinline__ void encode_band(float4 a, float4 c, int q, float& sA, float& cA, float& sB, float& cB,
                                            h4 (&x)[kGIn / 16])
{
    const bool tail = q == 3;
    x[BAND][0] = (_Float16)sA;
    x[BAND][1] = (_Float16)cA;
    x[BAND][2] = (_Float16)(tail ? raw_slot<2 * BAND>(a, c) : sB);
    x[BAND][3] = (_Float16)(tail ? raw_slot<2 * BAND + 1>(a, c) : cB);
    float s2 = 2.f * sA * cA, c2 = (cA - sA) * (cA + sA);
    sA = s2;
    cA = c2;
    s2 = 2.f * sB * cB;
    c2 = (cB - sB) * (cB + sB);
    sB = s2;
    cB = c2;
}

__device__ __forceinline__ void encode_row(float4 a, float4 c, int q, bool live, h4 (&x)[kGIn / 16])
{
    // the lane's two dimensions (q = 3: dimension 6 and an unused second one)
    float vA = q == 0 ? a.x : q == 1 ? a.z : q == 2 ? c.x : c.z;
    float vB = q == 0 ? a.y : q == 1 ? a.w : q == 2 ? c.y : 0.f;
    float rA = __builtin_amdgcn_fractf(0.5f * vA), rB = __builtin_amdgcn_fractf(0.5f * vB);
    float sA = __builtin_amdgcn_sinf(rA), cA = __builtin_amdgcn_cosf(rA);
    float sB = __builtin_amdgcn_sinf(rB), cB = __builtin_amdgcn_cosf(rB);
    encode_band<0>(a, c, q, sA, cA, sB, cB, x);
    encode_band<1>(a, c, q, sA, cA, sB, cB, x);
    encode_band<2>(a, c, q, sA, cA, sB, cB, x);
    encode_band<3>(a, c, q, sA, cA, sB, cB, x);
    encode_band<4>(a, c, q, sA, cA, sB, cB, x);
    encode_band<5>(a, c, q, sA, cA, sB, cB, x);
    static_assert(kGIn / 16 == kGBands, "one 16-feature block per band");
    if (!live)
#pragma unroll
        for (int k = 0; k < kGIn / 16; ++k) x[k] = h4{0, 0, 0, 0};
}

// One dense layer for kT 16-row tiles on the transposed product Y^T = W^T X^T with 16x32 tiles:
// the weights are the A operand (LDS, out-major rows, columns permuted by perm32), the
// activations the B operand. Each weight fragment is read from LDS once and used for all kT
// tiles. The MFMA result fragment (lane l holds Y^T[16 mt + 4(l/16) + i][row l%16]) is half of
// the next layer's B fragment for k-block mt/2, so activations never leave the registers.
template <int K, int N, int kT>
__device__ __forceinline__ void dense_t(const h4 (&x)[kT][K / 16], const _Float16* Wt, int ws, const float* b,
                                        h4 (&y)[kT][N / 16], int r, int q)
{
#pragma unroll
    for (int mt = 0; mt < N / 16; ++mt)
    {
        // the accumulators start at the bias (lane rows 16 mt + 4q + i), so no add follows
        const float4 bb = *reinterpret_cast<const float4*>(&b[mt * 16 + 4 * q]);
        f4v acc[kT];
#pragma unroll
        for (int t = 0; t < kT; ++t) acc[t] = f4v{bb.x, bb.y, bb.z, bb.w};
#pragma unroll
        for (int kb = 0; kb < K / 32; ++kb)
        {
            const h8 w = *reinterpret_cast<const h8*>(&Wt[(mt * 16 + r) * ws + kb * 32 + 8 * q]);
#pragma unroll
            for (int t = 0; t < kT; ++t) acc[t] = mfma32(w, cat8(x[t][2 * kb], x[t][2 * kb + 1]), acc[t]);
        }
#pragma unroll
        for (int t = 0; t < kT; ++t) y[t][mt] = relu_h4(acc[t]);
    }
}

constexpr int kFwdWaves = 8;
// 16-row tiles per wave per step, sharing each weight read: 4 halves the LDS reads per FLOP
// but needs 124 VGPRs (4 waves per SIMD instead of 6), and the guided run's inference took
// 75.0 ms against 72.1 with 2
constexpr int kFwdTiles = 2;

// Fused forward: rows of raw inputs (`stride` float4 per row, the first two holding
// v0..v6) -> the network's outputs. Weights and biases are staged in LDS once per block; each
// wave runs kFwdTiles 16-row tiles at a time with all activations in registers. kHalf: write only the kGUsed outputs the mixture uses, in FP16 (the wavefront's
// form), else all kGOut in FP32. seg (when set): rows come in segments of kGuideSeg of which
// the first seg[s] are live; tile groups past a segment's rows are skipped.
template <bool kHalf>
__global__ __launch_bounds__(64 * kFwdWaves) void k_guide_forward(GuideNet g, const float4* in, int stride,
                                                                 uint32_t nrows, const uint32_t* seg, void* outp)
{
    constexpr int kGroup = 16 * kFwdTiles;
    static_assert(kGuideSeg % kGroup == 0, "tile groups must not straddle segments");
    __shared__ __align__(16) _Float16 lw[kLFTotal];
    __shared__ __align__(16) float lb[kBiasTotal];
    load_fwd_weights(g, lw, lb);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, q = lane >> 4;
    const uint32_t ngroups = (nrows + kGroup - 1) / kGroup;
    for (uint32_t grp = blockIdx.x * kFwdWaves + wave; grp < ngroups; grp += gridDim.x * kFwdWaves)
    {
        // an offset the compiler cannot see through keeps the weight reads inside the loop:
        // hoisted, all five layers' fragments would take ~180 registers per lane
        int zoff = 0;
        asm volatile("" : "+v"(zoff));
        const _Float16* W = lw + zoff;
        const uint32_t row0 = grp * kGroup;
        uint32_t seg_live = kGuideSeg;
        if (seg)
        {
            seg_live = seg[row0 / kGuideSeg];                 // the same for the whole group
            if (row0 % kGuideSeg >= seg_live) continue;       // group past the segment's rows
        }
        uint32_t row[kFwdTiles];
        bool live[kFwdTiles];
        h4 x[kFwdTiles][kGIn / 16];
#pragma unroll
        for (int t = 0; t < kFwdTiles; ++t)
        {
            row[t] = row0 + 16 * t + r;
            live[t] = row[t] < nrows && row[t] % kGuideSeg < seg_live;
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f), c = a;
            if (live[t])
            {
                a = in[(size_t)stride * row[t]];
                c = in[(size_t)stride * row[t] + 1];
            }
            encode_row(a, c, q, live[t], x[t]);
        }
        h4 h1[kFwdTiles][kGHid / 16], h2[kFwdTiles][kGHid / 16];
        dense_t<kGIn, kGHid, kFwdTiles>(x, W + kLF1, kL1S, lb, h1, r, q);
        dense_t<kGHid, kGHid, kFwdTiles>(h1, W + kLF2, kLHS, lb + kGHid, h2, r, q);
        dense_t<kGHid, kGHid, kFwdTiles>(h2, W + kLF3, kLHS, lb + 2 * kGHid, h1, r, q);
        dense_t<kGHid, kGHid, kFwdTiles>(h1, W + kLF4, kLHS, lb + 3 * kGHid, h2, r, q);
        // output layer, FP32 straight from the accumulators: lane holds outputs 16 mt + 4q + i of its row
#pragma unroll
        for (int mt = 0; mt < kGOut / 16; ++mt)
        {
            const float4 bb = *reinterpret_cast<const float4*>(&lb[4 * kGHid + mt * 16 + 4 * q]);
            f4v acc[kFwdTiles];
#pragma unroll
            for (int t = 0; t < kFwdTiles; ++t) acc[t] = f4v{bb.x, bb.y, bb.z, bb.w};
#pragma unroll
            for (int kb = 0; kb < kGHid / 32; ++kb)
            {
                const h8 w = *reinterpret_cast<const h8*>(&W[kLF5 + (mt * 16 + r) * kLHS + kb * 32 + 8 * q]);
#pragma unroll
                for (int t = 0; t < kFwdTiles; ++t) acc[t] = mfma32(w, cat8(h2[t][2 * kb], h2[t][2 * kb + 1]), acc[t]);
            }
#pragma unroll
            for (int t = 0; t < kFwdTiles; ++t)
            {
                if (!live[t]) continue;
                if (kHalf)
                {
                    if (mt * 16 + 4 * q < kGUsed)
                    {
                        h4 o = {(_Float16)acc[t][0], (_Float16)acc[t][1], (_Float16)acc[t][2], (_Float16)acc[t][3]};
                        *reinterpret_cast<h4*>(&static_cast<_Float16*>(outp)[(size_t)row[t] * kGUsed + mt * 16 + 4 * q]) = o;
                    }
                }
                else
                {
                    *reinterpret_cast<float4*>(&static_cast<float*>(outp)[(size_t)row[t] * kGOut + mt * 16 + 4 * q]) =
                        make_float4(acc[t][0], acc[t][1], acc[t][2], acc[t][3]);
                }
            }
        }
    }
}

constexpr float kMaxRecordWeight = 256.f;

// Per-row loss -w log q(w_dir) and its gradient w.r.t. the 48 outputs. Weights are
// normalised to mean 1 over the batch (s = w * rows / sum w) so FP16 gradients stay O(1);
// the weight-gradient reduction divides by rows. rec: 3 float4 per row: (x, r), (n, active),
// (dir, w). T: the gradient's type (FP16 for the MFMA backward, FP32 in FP32 mode).
template <class T>
__global__ void k_guide_loss(const float* out, const float4* rec, int srow, uint32_t nrows, const float* wsum,
                             T* dout, float* loss_acc)
{
    uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    const float* o = out + (size_t)row * kGOut;
    float4 c = rec[(size_t)srow * row + 2];
    f3 w = mk(c.x, c.y, c.z);
    float wt = c.w;
    float scale = (float)nrows / fmaxf(wsum[0], 1e-30f);
    Lobes L = guide_lobes(o);
    float v[kGLobes], q = 0.f;
    for (int k = 0; k < kGLobes; ++k)
    {
        v[k] = vmf_pdf(L.mu[k], L.kappa[k], w);
        q += L.w[k] * v[k];
    }
    float g[kGOut];
    for (int k = 0; k < kGOut; ++k) g[k] = 0.f;
    // weights normalised to mean 1, capped so one rare record cannot overflow the FP16
    // gradients (this shapes the proposal only; the rendered estimate stays unbiased)
    float s = fminf(wt * scale, kMaxRecordWeight);
    if (q > 0.f && s > 0.f && isfinite(q) && isfinite(s))
    {
        for (int k = 0; k < kGLobes; ++k)
        {
            float gam = L.w[k] * v[k] / q;
            g[k] = s * (L.w[k] - gam);
            // d(-log q)/d mu_k = -gam kappa w, projected through mu = r / |r|
            f3 r = mk(o[8 + 3 * k], o[9 + 3 * k], o[10 + 3 * k]);
            float rn = sqrtf(sqnorm(r));
            if (rn > 0.f)
            {
                f3 gm = w * (-s * gam * L.kappa[k]);
                f3 gr = (gm - L.mu[k] * dot(L.mu[k], gm)) * (1.f / rn);
                g[8 + 3 * k] = gr.x;
                g[9 + 3 * k] = gr.y;
                g[10 + 3 * k] = gr.z;
            }
            float lk = o[32 + k];
            if (lk > kLogKappaLo && lk < kLogKappaHi)
            {
                float kap = L.kappa[k];
                float e = expf(-2.f * kap);
                float dlogv = 1.f / kap - 2.f * e / (1.f - e) + (dot(L.mu[k], w) - 1.f);
                g[32 + k] = -s * gam * kap * dlogv;
            }
        }
        atomicAdd(loss_acc, -s * logf(q));
    }
    for (int k = 0; k < kGOut; ++k) dout[(size_t)row * kGOut + k] = (T)g[k];
}

// Sum of the record weights in a fixed order, so that training is reproducible bit for bit:
// kWsumParts blocks each reduce a contiguous range with a fixed tree (pass 1), then one block
// adds the partial sums the same way (pass 2).
constexpr int kWsumBlock = 256, kWsumParts = 256;
__device__ __forceinline__ float block_sum(float s, float* red)
{
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = kWsumBlock / 2; st > 0; st >>= 1)
    {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    return red[0];
}

__global__ __launch_bounds__(kWsumBlock) void k_guide_wsum(const float4* rec, int srow, uint32_t nrows, float* part)
{
    __shared__ float red[kWsumBlock];
    const uint32_t per = (nrows + kWsumParts - 1) / kWsumParts;
    const uint32_t r0 = blockIdx.x * per, r1 = min(r0 + per, nrows);
    float s = 0.f;
    for (uint32_t r = r0 + threadIdx.x; r < r1; r += kWsumBlock) s += rec[(size_t)srow * r + 2].w;
    float t = block_sum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kWsumBlock) void k_guide_wsum_final(const float* part, float* wsum)
{
    __shared__ float red[kWsumBlock];
    float s = 0.f;
    for (int k = threadIdx.x; k < kWsumParts; k += kWsumBlock) s += part[k];
    float t = block_sum(s, red);
    if (threadIdx.x == 0) wsum[0] = t;
}

// ---- fused training step ----
// One kernel computes a batch's loss gradient for every parameter: blocks of 8 waves (two per
// SIMD) take rounds of 128 rows, 16 per wave. In a round each wave runs the forward pass of its
// rows with the activations in registers, the loss and its gradient with respect to the 48
// outputs (4 lanes per row, two lobes each), then the backward pass layer by layer as the
// transposed products dH^T = W dOut^T masked by H > 0, so every activation and gradient stays in
// the forward's register layout (lane r + 16 q: row r, features 16 j + 4 q + i). For each layer
// the waves stage its inputs In_l and output gradients G_l feature-major in LDS (one LDS row per
// feature, one column per row of the round), compute the next layer's gradient, and after a
// block barrier add the round's 128 rows to the weight-gradient tiles they own with 16x32
// products whose operands are 8 rows of one feature, one 16-byte LDS read. Two layers are
// staged at a time (five layers, six barriers per round; LDS 157 KB). The backward products read
// only the weight image, so they fill the waits at the barriers: dH3 and dH1 are computed after
// a staging, dH2 after a pair's weight gradients, and each is computed as late as its staging
// allows, which keeps the kernel inside 256 VGPRs without spills. The tiles stay in registers over all the block's
// rounds (bias gradients as a product with a ones operand); the blocks' partial sums are added
// in a fixed order by k_guide_reduce_part / k_guide_reduce_train, so a step's result does not
// depend on scheduling.
constexpr int kTrWaves = 8;                                    // waves per block, 16 rows each
constexpr int kTrRows = 16 * kTrWaves;                         // rows per round
constexpr int kRp = kTrRows + kPad;                            // staging pitch (halves) per feature
constexpr int kStIn = 0;                                       // a layer's inputs^T: up to 96 features
constexpr int kStG = kStIn + kGIn * kRp;                       // its output gradients^T: up to 64
constexpr int kStLayer = kStG + kGHid * kRp;                   // one staged layer
constexpr int kStTotal = 2 * kStLayer;                         // two layers at a time
constexpr int kPartStride = kGParams + 4;                      // a block's partial sums, then its loss
static_assert((kRp * 2) % 16 == 0, "16-byte aligned feature rows");

// ReLU derivative bits of a layer's four fragments: bit 4 j + i = (h[j][i] > 0).
__device__ __forceinline__ uint32_t relu_bits(const h4 (&h)[kGHid / 16])
{
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < kGHid / 16; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) m |= ((float)h[j][i] > 0.f ? 1u : 0u) << (4 * j + i);
    return m;
}

// dense_t for one tile with the k-blocks outer: each k-block's weight fragments for all output
// blocks are read together and feed independent accumulators (the training kernel runs two
// waves per SIMD, so little else hides the LDS latency of a dependent chain).
template <int K, int N>
__device__ __forceinline__ void dense1(const h4 (&x)[K / 16], const _Float16* Wt, int ws, const float* b,
                                       h4 (&y)[N / 16], int r, int q)
{
    f4v acc[N / 16];
#pragma unroll
    for (int mt = 0; mt < N / 16; ++mt)
    {
        const float4 bb = *reinterpret_cast<const float4*>(&b[mt * 16 + 4 * q]);
        acc[mt] = f4v{bb.x, bb.y, bb.z, bb.w};
    }
#pragma unroll
    for (int kb = 0; kb < K / 32; ++kb)
    {
        h8 w[N / 16];
#pragma unroll
        for (int mt = 0; mt < N / 16; ++mt) w[mt] = *reinterpret_cast<const h8*>(&Wt[(mt * 16 + r) * ws + kb * 32 + 8 * q]);
#pragma unroll
        for (int mt = 0; mt < N / 16; ++mt) acc[mt] = mfma32(w[mt], cat8(x[2 * kb], x[2 * kb + 1]), acc[mt]);
    }
#pragma unroll
    for (int mt = 0; mt < N / 16; ++mt) y[mt] = relu_h4(acc[mt]);
}

// dH^T = W dOut^T masked by the ReLU bits, with A[m = in][k = out] fragments of W^T read from the
// forward LDS image (out-major rows, perm32 columns): four rows apart. Out-blocks outer, four
// independent accumulators.
template <int NOUT, int NIN>
__device__ __forceinline__ void bwd_t(const _Float16* Wt, int ws, const h4 (&dout)[NOUT / 16], uint32_t mask,
                                      h4 (&dh)[NIN / 16], int r, int q)
{
    f4v acc[NIN / 16];
#pragma unroll
    for (int nt = 0; nt < NIN / 16; ++nt) acc[nt] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < NOUT / 16; ++ks)
    {
        h4 a[NIN / 16];
#pragma unroll
        for (int nt = 0; nt < NIN / 16; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i) a[nt][i] = Wt[(16 * ks + 4 * q + i) * ws + perm32(16 * nt + r)];
#pragma unroll
        for (int nt = 0; nt < NIN / 16; ++nt) acc[nt] = mfma16(a[nt], dout[ks], acc[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < NIN / 16; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) dh[nt][i] = (mask >> (4 * nt + i)) & 1u ? (_Float16)acc[nt][i] : (_Float16)0.f;
}

// Feature-major staging of the lane's fragments: T[feature][lr].
template <int N>
__device__ __forceinline__ void stage_t(_Float16* T, int lr, const h4 (&v)[N / 16], int q)
{
#pragma unroll
    for (int j = 0; j < N / 16; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) T[(16 * j + 4 * q + i) * kRp + lr] = v[j][i];
}

// One wave's 16 rows (raw inputs a, c and record (dir, weight) d already loaded): the forward
// pass into x and h1..h4, the loss gradient with respect to the outputs into dy (the backward's
// B fragments). O: the wave's 16 x 48 exchange area. Returns the wave's summed loss (every lane
// the same value).
__device__ __forceinline__ float train_forward(const _Float16* lw, const float* lb, float* O, float4 a, float4 c,
                                               float4 d, bool live, float scale, int r, int q, h4 (&x)[kGIn / 16],
                                               h4 (&h1)[kGHid / 16], h4 (&h2)[kGHid / 16], h4 (&h3)[kGHid / 16],
                                               h4 (&hh)[kGHid / 16], h4 (&dy)[kGOut / 16])
{
    encode_row(a, c, q, live, x);
    dense1<kGIn, kGHid>(x, lw + kLF1, kL1S, lb, h1, r, q);
    dense1<kGHid, kGHid>(h1, lw + kLF2, kLHS, lb + kGHid, h2, r, q);
    dense1<kGHid, kGHid>(h2, lw + kLF3, kLHS, lb + 2 * kGHid, h3, r, q);
    dense1<kGHid, kGHid>(h3, lw + kLF4, kLHS, lb + 3 * kGHid, hh, r, q);
    {
        f4v acc[kGOut / 16];
#pragma unroll
        for (int mt = 0; mt < kGOut / 16; ++mt)
        {
            const float4 bb = *reinterpret_cast<const float4*>(&lb[4 * kGHid + mt * 16 + 4 * q]);
            acc[mt] = f4v{bb.x, bb.y, bb.z, bb.w};
        }
#pragma unroll
        for (int kb = 0; kb < kGHid / 32; ++kb)
        {
            h8 w[kGOut / 16];
#pragma unroll
            for (int mt = 0; mt < kGOut / 16; ++mt)
                w[mt] = *reinterpret_cast<const h8*>(&lw[kLF5 + (mt * 16 + r) * kLHS + kb * 32 + 8 * q]);
#pragma unroll
            for (int mt = 0; mt < kGOut / 16; ++mt) acc[mt] = mfma32(w[mt], cat8(hh[2 * kb], hh[2 * kb + 1]), acc[mt]);
        }
#pragma unroll
        for (int mt = 0; mt < kGOut / 16; ++mt)
            *reinterpret_cast<float4*>(&O[r * kGOut + mt * 16 + 4 * q]) =
                make_float4(acc[mt][0], acc[mt][1], acc[mt][2], acc[mt][3]);
    }
    __builtin_amdgcn_wave_barrier();   // the wave's LDS accesses complete in order
    // loss -s log q(w) and its gradient (k_guide_loss's formulas); lane q takes lobes 2q, 2q+1
    const float* o = O + r * kGOut;
    float mx = o[0];
#pragma unroll
    for (int k = 1; k < kGLobes; ++k) mx = fmaxf(mx, o[k]);
    float se = 0.f;
#pragma unroll
    for (int k = 0; k < kGLobes; ++k) se += gexp(o[k] - mx);
    const float is = grcp(se);
    const f3 wd = mk(d.x, d.y, d.z);
    const float sw = fminf(d.w * scale, kMaxRecordWeight);
    float wk[2], vk[2], kap[2];
    f3 mu[2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
    {
        const int k = 2 * q + j;
        wk[j] = gexp(o[k] - mx) * is;
        mu[j] = lobe_mu(o, k);
        kap[j] = gexp(clampf(o[32 + k], kLogKappaLo, kLogKappaHi));
        vk[j] = vmf_pdf(mu[j], kap[j], wd);
    }
    float qm = wk[0] * vk[0] + wk[1] * vk[1];
    qm += __shfl_xor(qm, 16);
    qm += __shfl_xor(qm, 32);
    const bool ok = live && qm > 0.f && sw > 0.f && isfinite(qm) && isfinite(sw);
    float gl[2] = {0.f, 0.f}, gk[2] = {0.f, 0.f};
    f3 gmu[2] = {mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f)};
    if (ok)
    {
#pragma unroll
        for (int j = 0; j < 2; ++j)
        {
            const int k = 2 * q + j;
            const float gam = wk[j] * vk[j] / qm;
            gl[j] = sw * (wk[j] - gam);
            const f3 rr = mk(o[8 + 3 * k], o[9 + 3 * k], o[10 + 3 * k]);
            const float rn = sqrtf(sqnorm(rr));
            if (rn > 0.f)
            {
                const f3 gm = wd * (-sw * gam * kap[j]);
                gmu[j] = (gm - mu[j] * dot(mu[j], gm)) * (1.f / rn);
            }
            const float lk = o[32 + k];
            if (lk > kLogKappaLo && lk < kLogKappaHi)
            {
                const float e = expf(-2.f * kap[j]);
                const float dlogv = 1.f / kap[j] - 2.f * e / (1.f - e) + (dot(mu[j], wd) - 1.f);
                gk[j] = -sw * gam * kap[j] * dlogv;
            }
        }
    }
    float loss = (ok && q == 0) ? -sw * logf(qm) : 0.f;
    __builtin_amdgcn_wave_barrier();   // every output read before the area holds the gradient
    // the row's 48 output gradients as halves over the exchange area, this lane's 12 of them
    _Float16* Y = reinterpret_cast<_Float16*>(O) + r * kGOut;
    Y[2 * q] = (_Float16)gl[0];
    Y[2 * q + 1] = (_Float16)gl[1];
    Y[8 + 6 * q + 0] = (_Float16)gmu[0].x;
    Y[8 + 6 * q + 1] = (_Float16)gmu[0].y;
    Y[8 + 6 * q + 2] = (_Float16)gmu[0].z;
    Y[8 + 6 * q + 3] = (_Float16)gmu[1].x;
    Y[8 + 6 * q + 4] = (_Float16)gmu[1].y;
    Y[8 + 6 * q + 5] = (_Float16)gmu[1].z;
    Y[32 + 2 * q] = (_Float16)gk[0];
    Y[33 + 2 * q] = (_Float16)gk[1];
    Y[40 + 2 * q] = (_Float16)0.f;
    Y[41 + 2 * q] = (_Float16)0.f;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int ks = 0; ks < kGOut / 16; ++ks) dy[ks] = *reinterpret_cast<const h4*>(&Y[16 * ks + 4 * q]);
    // the wave's loss, summed over its lanes in a fixed butterfly
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) loss += __shfl_xor(loss, m);
    return loss;
}

// One layer's weight-gradient tiles of a wave over the round's staged rows: acc[A0 + t] +=
// In[row][16 (ib0 + t ibstep) + in]^T G[row][16 ob + out] for t < NT (when `tiles`), and with
// `bias` acc[B0] += sum_rows G[row][16 ob + out] in every row (a product with a ones operand).
// A[m = in][k = row] and B[k = row][n = out] are 8 consecutive rows of one feature per lane.
constexpr int kAccTiles = 16;
template <int NT, int A0, int B0>
__device__ __forceinline__ void wgrad_layer(const _Float16* In, const _Float16* G, int ib0, int ibstep, int ob,
                                            bool tiles, bool bias, f4v (&acc)[kAccTiles], int r, int q)
{
    const h8 ones = {(_Float16)1.f, (_Float16)1.f, (_Float16)1.f, (_Float16)1.f,
                     (_Float16)1.f, (_Float16)1.f, (_Float16)1.f, (_Float16)1.f};
#pragma unroll
    for (int g = 0; g < kTrRows / 32; ++g)
    {
        const h8 b = *reinterpret_cast<const h8*>(&G[(16 * ob + r) * kRp + 32 * g + 8 * q]);
        if (bias) acc[B0] = mfma32(ones, b, acc[B0]);
        if (tiles)
        {
            h8 a[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t)
                a[t] = *reinterpret_cast<const h8*>(&In[(16 * (ib0 + t * ibstep) + r) * kRp + 32 * g + 8 * q]);
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[A0 + t] = mfma32(a[t], b, acc[A0 + t]);
        }
    }
}

template <int NT, int A0, int B0>
__device__ __forceinline__ void wgrad_store(float* P, int woff, int boff, int nout, int ib0, int ibstep, int ob,
                                            bool tiles, bool bias, const f4v (&acc)[kAccTiles], int r, int q)
{
    if (tiles)
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                P[woff + (16 * (ib0 + t * ibstep) + 4 * q + i) * nout + 16 * ob + r] = acc[A0 + t][i];
    if (bias && q == 0) P[boff + 16 * ob + r] = acc[B0][0];
}

// Tiles of wave w (the same shapes for every wave, so one code path): out-block w % 4 of layers
// 1-4, in-blocks w / 4 + 2 t (layer 1: three, layers 2-4: two), and waves 0-5 out-block w % 3
// of layer 5, in-blocks w / 3 + 2 t (two); the bias tiles go to waves 0-3 (layer 5: 0-2).
// acc: layer 1 [0, 3) + bias 3, layer 2 [4, 6) + 6, layer 3 [7, 9) + 9, layer 4 [10, 12) + 12,
// layer 5 [13, 15) + 15.
__global__ __launch_bounds__(64 * kTrWaves, 1) void k_guide_train_fused(GuideNet g, const float4* rec, int srow,
                                                                        uint32_t nrows, const float* wsum,
                                                                        float* part)
{
    __shared__ __align__(16) _Float16 lw[kLFTotal];
    __shared__ __align__(16) float lb[kBiasTotal];
    __shared__ __align__(16) _Float16 st[kStTotal];
    __shared__ __align__(16) float ox[kTrWaves * 16 * kGOut];
    __shared__ float red[kTrWaves];
    load_fwd_weights(g, lw, lb);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    const float scale = (float)nrows / fmaxf(wsum[0], 1e-30f);
    const int ob = wave & 3, ib = wave >> 2, ob5 = wave % 3, ib5 = wave / 3;
    const bool bias4 = wave < 4, tiles5 = wave < 6, bias5 = wave < 3;
    f4v acc[kAccTiles];
#pragma unroll
    for (int t = 0; t < kAccTiles; ++t) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
    float loss = 0.f;
    const uint32_t rounds = (nrows + kTrRows - 1) / kTrRows;
    const int lr = 16 * wave + r;
    _Float16* sA = st;
    _Float16* sB = st + kStLayer;
    float* O = ox + wave * 16 * kGOut;
    // the lane's record of the next round, loaded during the current one
    float4 na = make_float4(0.f, 0.f, 0.f, 0.f), nc = na, nd = na;
    uint32_t nrow = blockIdx.x * kTrRows + lr;
    if (nrow < nrows)
    {
        na = rec[(size_t)srow * nrow];
        nc = rec[(size_t)srow * nrow + 1];
        nd = rec[(size_t)srow * nrow + 2];
    }
    for (uint32_t rd = blockIdx.x; rd < rounds; rd += gridDim.x)
    {
        const float4 a = na, c = nc, d = nd;
        const bool live = nrow < nrows;
        nrow += gridDim.x * kTrRows;
        // an offset the compiler cannot see through keeps the weight reads inside the loop
        int zoff = 0;
        asm volatile("" : "+v"(zoff));
        const _Float16* W = lw + zoff;
        h4 x[kGIn / 16], h1[kGHid / 16], h2[kGHid / 16], h3[kGHid / 16], hh[kGHid / 16], dy[kGOut / 16];
        loss += train_forward(W, lb, O, a, c, d, live, scale, r, q, x, h1, h2, h3, hh, dy);
        // dH4 before the first barrier (its weight reads are not in the staging area)
        h4 g4[kGHid / 16], g3[kGHid / 16], g2[kGHid / 16], g1[kGHid / 16];
        bwd_t<kGOut, kGHid>(W + kLF5, kLHS, dy, relu_bits(hh), g4, r, q);
        const uint32_t m3 = relu_bits(h3);
        // the previous round's last weight gradients have read the staging area
        __syncthreads();
        // layers 5 and 4
        stage_t<kGHid>(sA + kStIn, lr, hh, q);
        stage_t<kGOut>(sA + kStG, lr, dy, q);
        stage_t<kGHid>(sB + kStIn, lr, h3, q);
        stage_t<kGHid>(sB + kStG, lr, g4, q);
        bwd_t<kGHid, kGHid>(W + kLF4, kLHS, g4, m3, g3, r, q);
        // the next round's record, issued here where fewer activations are live
        if (rd + gridDim.x < rounds && nrow < nrows)
        {
            na = rec[(size_t)srow * nrow];
            nc = rec[(size_t)srow * nrow + 1];
            nd = rec[(size_t)srow * nrow + 2];
        }
        __syncthreads();
        wgrad_layer<2, 13, 15>(sA + kStIn, sA + kStG, ib5, 2, ob5, tiles5, bias5, acc, r, q);
        wgrad_layer<2, 10, 12>(sB + kStIn, sB + kStG, ib, 2, ob, true, bias4, acc, r, q);
        // dH2 while the other waves finish reading layers 5 and 4
        bwd_t<kGHid, kGHid>(W + kLF3, kLHS, g3, relu_bits(h2), g2, r, q);
        const uint32_t m1 = relu_bits(h1);
        __syncthreads();
        // layers 3 and 2
        stage_t<kGHid>(sA + kStIn, lr, h2, q);
        stage_t<kGHid>(sA + kStG, lr, g3, q);
        stage_t<kGHid>(sB + kStIn, lr, h1, q);
        stage_t<kGHid>(sB + kStG, lr, g2, q);
        bwd_t<kGHid, kGHid>(W + kLF2, kLHS, g2, m1, g1, r, q);
        __syncthreads();
        wgrad_layer<2, 7, 9>(sA + kStIn, sA + kStG, ib, 2, ob, true, bias4, acc, r, q);
        wgrad_layer<2, 4, 6>(sB + kStIn, sB + kStG, ib, 2, ob, true, bias4, acc, r, q);
        __syncthreads();
        // layer 1
        stage_t<kGIn>(sA + kStIn, lr, x, q);
        stage_t<kGHid>(sA + kStG, lr, g1, q);
        __syncthreads();
        wgrad_layer<3, 0, 3>(sA + kStIn, sA + kStG, ib, 2, ob, true, bias4, acc, r, q);
    }
    float* P = part + (size_t)blockIdx.x * kPartStride;
    wgrad_store<3, 0, 3>(P, kGW1, kGB1, kGHid, ib, 2, ob, true, bias4, acc, r, q);
    wgrad_store<2, 4, 6>(P, kGW2, kGB2, kGHid, ib, 2, ob, true, bias4, acc, r, q);
    wgrad_store<2, 7, 9>(P, kGW3, kGB3, kGHid, ib, 2, ob, true, bias4, acc, r, q);
    wgrad_store<2, 10, 12>(P, kGW4, kGB4, kGHid, ib, 2, ob, true, bias4, acc, r, q);
    wgrad_store<2, 13, 15>(P, kGW5, kGB5, kGOut, ib5, 2, ob5, tiles5, bias5, acc, r, q);
    if (lane == 0) red[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        float s = 0.f;
        for (int w = 0; w < kTrWaves; ++w) s += red[w];
        P[kGParams] = s;
    }
}

// Sums the fused kernel's block partials in a fixed order, in two passes so that no thread walks
// every block: pass 1 (blockIdx.y = group) adds groups of kRedGroup blocks, pass 2 the groups,
// giving grad = sum / rows and loss = the summed per-row losses.
constexpr int kRedGroup = 16;
__global__ void k_guide_reduce_part(const float* part, int nblocks, float* out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > kGParams) return;
    const int b0 = blockIdx.y * kRedGroup, b1 = min(b0 + kRedGroup, nblocks);
    float s = 0.f;
    for (int b = b0; b < b1; ++b) s += part[(size_t)b * kPartStride + t];
    out[(size_t)blockIdx.y * kPartStride + t] = s;
}

__global__ void k_guide_reduce_train(const float* part, int ngroups, float inv_rows, float* grad, float* loss)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > kGParams) return;
    float s = 0.f;
    for (int b = 0; b < ngroups; ++b) s += part[(size_t)b * kPartStride + t];
    if (t < kGParams) grad[t] = s * inv_rows;
    else loss[0] = s;
}

__global__ void k_guide_zero(float* p, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.f;
}

// Adam on the FP32 master parameters, then refresh the FP16 forward / backward copies.
__global__ void k_guide_adam(float* p, float* m, float* v, const float* grad, int n, float lr, float b1, float b2,
                             float eps, float c1, float c2, _Float16* wf, _Float16* wb, float* bias)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float g = grad[t];
    if (!isfinite(g)) g = 0.f;   // an overflowed entry leaves its parameter as it is
    float mm = b1 * m[t] + (1.f - b1) * g;
    float vv = b2 * v[t] + (1.f - b2) * g * g;
    m[t] = mm;
    v[t] = vv;
    float x = p[t] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    p[t] = x;
    // scatter into the FP16 layouts
    _Float16 hx = (_Float16)x;
    if (t < kGB1) { int k = t / kGHid, nn = t % kGHid; wf[kF1 + nn * kGIn + k] = hx; }
    else if (t < kGW2) bias[t - kGB1] = x;
    else if (t < kGB2) { int u = t - kGW2, k = u / kGHid, nn = u % kGHid; wf[kF2 + nn * kGHid + k] = hx; wb[kB2 + k * kGHid + nn] = hx; }
    else if (t < kGW3) bias[kGHid + t - kGB2] = x;
    else if (t < kGB3) { int u = t - kGW3, k = u / kGHid, nn = u % kGHid; wf[kF3 + nn * kGHid + k] = hx; wb[kB3 + k * kGHid + nn] = hx; }
    else if (t < kGW4) bias[2 * kGHid + t - kGB3] = x;
    else if (t < kGB4) { int u = t - kGW4, k = u / kGHid, nn = u % kGHid; wf[kF4 + nn * kGHid + k] = hx; wb[kB4 + k * kGHid + nn] = hx; }
    else if (t < kGW5) bias[3 * kGHid + t - kGB4] = x;
    else if (t < kGB5) { int u = t - kGW5, k = u / kGOut, nn = u % kGOut; wf[kF5 + nn * kGHid + k] = hx; wb[kB5 + k * kGOut + nn] = hx; }
    else bias[4 * kGHid + t - kGB5] = x;
}

// ---- FP32 mode (verification) ----
// The same network, loss and backward computed from the FP32 master parameters (input-major
// W[k][n], then the bias) with FP32 features, activations and gradients and plain FMA
// arithmetic: one thread per output element, every sum in a fixed order. It checks the FP16
// kernels' feature order, layer wiring and gradient formulas without their rounding and is
// not a fast path. The features use the accurate sinpi/cospi of 2^b v (an exact scaling)
// rather than the FP16 encoder's double-angle recurrence.
__global__ void k_f32_encode(const float4* in, int stride, uint32_t nrows, float* X)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nrows * kGIn) return;
    const size_t row = t / kGIn;
    const int f = (int)(t % kGIn), b = f / 16, p = f % 16;
    const float4 a = in[(size_t)stride * row], c = in[(size_t)stride * row + 1];
    const float v[kGRaw] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z};
    float x = 0.f;
    if (p < 14)
    {
        const float arg = ldexpf(v[p / 2], b);
        x = (p & 1) ? cospif(arg) : sinpif(arg);
    }
    else if (2 * b + p - 14 < kGRaw)
        x = v[2 * b + p - 14];
    X[t] = x;
}

// Y[row][n] = act(b[n] + sum_k X[row][k] W[k][n]), k ascending; act = ReLU or identity.
__global__ void k_f32_dense(const float* X, int K, const float* W, const float* b, int N, uint32_t nrows, int relu,
                            float* Y)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nrows * N) return;
    const size_t row = t / N;
    const int n = (int)(t % N);
    const float* x = X + row * K;
    float acc = b[n];
    for (int k = 0; k < K; ++k) acc = fmaf(x[k], W[(size_t)k * N + n], acc);
    Y[t] = relu ? fmaxf(acc, 0.f) : acc;
}

// dX[row][i] = (H[row][i] > 0) sum_j dY[row][j] W[i][j]: the gradient through a layer whose
// input-major weights are W[NIN][KOUT] and through the ReLU that produced its input H.
__global__ void k_f32_bwd(const float* dY, int KOUT, const float* W, const float* H, int NIN, uint32_t nrows,
                          float* dX)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nrows * NIN) return;
    const size_t row = t / NIN;
    const int i = (int)(t % NIN);
    float s = 0.f;
    if (H[t] > 0.f)
        for (int j = 0; j < KOUT; ++j) s = fmaf(dY[row * KOUT + j], W[(size_t)i * KOUT + j], s);
    dX[t] = s;
}

// One layer's gradient block in the parameter layout: W[k][n] = inv_rows sum_rows X[row][k] dY[row][n]
// (threads 0 .. K*N-1), then b[n] = inv_rows sum_rows dY[row][n], rows ascending.
__global__ void k_f32_wgrad(const float* X, int K, const float* dY, int N, uint32_t nrows, float inv_rows, float* grad)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= K * N + N) return;
    float s = 0.f;
    if (t < K * N)
    {
        const int k = t / N, n = t % N;
        for (uint32_t r = 0; r < nrows; ++r) s = fmaf(X[(size_t)r * K + k], dY[(size_t)r * N + n], s);
    }
    else
    {
        const int n = t - K * N;
        for (uint32_t r = 0; r < nrows; ++r) s += dY[(size_t)r * N + n];
    }
    grad[t] = s * inv_rows;
}

__global__ void k_diag_mfma(const _Float16* a, const _Float16* b, float* d)
{
    // a: 16x16 row-major (M x K), b: 16x16 row-major (K x N), d: 16x16 row-major
    const int lane = threadIdx.x, r = lane & 15, q = lane >> 4;
    h4 af, bf;
    for (int i = 0; i < 4; ++i)
    {
        af[i] = a[r * 16 + 4 * q + i];
        bf[i] = b[(4 * q + i) * 16 + r];
    }
    f4v acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma16(af, bf, acc);
    for (int i = 0; i < 4; ++i) d[(4 * q + i) * 16 + r] = acc[i];
}

// Lobe decoding, mixture pdf and Jakob sampling on given network outputs (diagnostic).
// out row: sampled direction (3), pdf at the sample, pdf at dirs[t], lobe weight sum, 2 unused.
__global__ void k_diag_guide(uint32_t n, const float* o, const float* u, const float* dirs, float* out)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    Lobes L = guide_lobes(o + (size_t)t * kGOut);
    f3 w = guide_sample(L, o + (size_t)t * kGOut, u[3 * t], u[3 * t + 1], u[3 * t + 2]);
    f3 d = mk(dirs[3 * t], dirs[3 * t + 1], dirs[3 * t + 2]);
    float ws = 0.f;
    for (int k = 0; k < kGLobes; ++k) ws += L.w[k];
    float* r = out + (size_t)t * 8;
    r[0] = w.x;
    r[1] = w.y;
    r[2] = w.z;
    r[3] = guide_pdf(L, w);
    r[4] = guide_pdf(L, d);
    r[5] = ws;
    r[6] = 0.f;
    r[7] = 0.f;
}

}  // namespace bsh

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
using namespace bsh;

struct bsh_guide
{
    bsh_context* ctx = nullptr;
    bsh_guide_params prm{};
    std::vector<float> host_params;
    DevBuf<float> p, m, v, grad, bias, part, part2, loss, wsum;
    DevBuf<_Float16> wf, wb;
    int step = 0;
    GuideNet net() const { return GuideNet{wf.p, wb.p, bias.p}; }
};

namespace bsh {

// PCG32 (M. O'Neill, 2014) for the deterministic He initialisation.
static uint32_t pcg32(uint64_t& st)
{
    uint64_t old = st;
    st = old * 6364136223846793005ULL + 1442695040888963407ULL;
    uint32_t x = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (x >> rot) | (x << ((32 - rot) & 31));
}

static bsh_status upload_params(bsh_guide* g)
{
    // fp32 master + derived fp16 copies, built on the host in the same layouts k_guide_adam writes
    const std::vector<float>& P = g->host_params;
    std::vector<_Float16> wf(kFTotal), wb(kBTotal);
    std::vector<float> bias(kBiasTotal);
    for (int k = 0; k < kGIn; ++k)
        for (int n = 0; n < kGHid; ++n) wf[kF1 + n * kGIn + k] = (_Float16)P[kGW1 + k * kGHid + n];
    const int wo[3] = {kGW2, kGW3, kGW4}, fo[3] = {kF2, kF3, kF4}, bo[3] = {kB2, kB3, kB4};
    for (int l = 0; l < 3; ++l)
        for (int k = 0; k < kGHid; ++k)
            for (int n = 0; n < kGHid; ++n)
            {
                _Float16 h = (_Float16)P[wo[l] + k * kGHid + n];
                wf[fo[l] + n * kGHid + k] = h;
                wb[bo[l] + k * kGHid + n] = h;
            }
    for (int k = 0; k < kGHid; ++k)
        for (int n = 0; n < kGOut; ++n)
        {
            _Float16 h = (_Float16)P[kGW5 + k * kGOut + n];
            wf[kF5 + n * kGHid + k] = h;
            wb[kB5 + k * kGOut + n] = h;
        }
    const int bsrc[5] = {kGB1, kGB2, kGB3, kGB4, kGB5};
    for (int l = 0; l < 4; ++l)
        for (int n = 0; n < kGHid; ++n) bias[l * kGHid + n] = P[bsrc[l] + n];
    for (int n = 0; n < kGOut; ++n) bias[4 * kGHid + n] = P[kGB5 + n];
    BSH_HIP(hipMemcpy(g->p.p, P.data(), kGParams * 4, hipMemcpyHostToDevice));
    BSH_HIP(hipMemcpy(g->wf.p, wf.data(), kFTotal * 2, hipMemcpyHostToDevice));
    BSH_HIP(hipMemcpy(g->wb.p, wb.data(), kBTotal * 2, hipMemcpyHostToDevice));
    BSH_HIP(hipMemcpy(g->bias.p, bias.data(), kBiasTotal * 4, hipMemcpyHostToDevice));
    return BSH_OK;
}

float guide_bsdf_fraction(const bsh_guide* g) { return g->prm.bsdf_fraction; }
int guide_bounces(const bsh_guide* g) { return g->prm.guided_bounces; }

bsh_status guide_forward(bsh_guide* g, const float4* in, int stride, uint32_t n, float* out, hipStream_t st)
{
    if (n == 0) return BSH_OK;
    uint32_t groups = (n + 16 * kFwdTiles - 1) / (16 * kFwdTiles);
    uint32_t blocks = std::min<uint32_t>((groups + kFwdWaves - 1) / kFwdWaves, (uint32_t)g->ctx->num_cus * 3);
    hipLaunchKernelGGL((k_guide_forward<false>), dim3(blocks), dim3(64 * kFwdWaves), 0, st, g->net(), in,
                       stride, n, (const uint32_t*)nullptr, (void*)out);
    BSH_HIP(hipGetLastError());
    return BSH_OK;
}

bsh_status guide_forward_rows(bsh_guide* g, const float4* in, const uint32_t* seg_counts, uint32_t rows,
                              _Float16* out, hipStream_t st)
{
    if (rows == 0) return BSH_OK;
    uint32_t groups = (rows + 16 * kFwdTiles - 1) / (16 * kFwdTiles);
    uint32_t blocks = std::min<uint32_t>((groups + kFwdWaves - 1) / kFwdWaves, (uint32_t)g->ctx->num_cus * 3);
    hipLaunchKernelGGL((k_guide_forward<true>), dim3(blocks), dim3(64 * kFwdWaves), 0, st, g->net(), in, 2,
                       rows, seg_counts, (void*)out);
    BSH_HIP(hipGetLastError());
    return BSH_OK;
}

// Forward + backward on `rec` (3 float4 per row); leaves the batch gradient in g->grad
// and the summed loss (times rows) in g->loss.
static bsh_status guide_gradients(bsh_guide* g, const float4* rec, int srow, uint32_t n, hipStream_t st)
{
    const int ncu = g->ctx->num_cus;
    const uint32_t nb = std::min<uint32_t>((n + kTrRows - 1) / kTrRows, (uint32_t)ncu);
    const uint32_t ngroups = (nb + kRedGroup - 1) / kRedGroup;
    if (!g->part.p) BSH_HIP(g->part.alloc((size_t)ncu * kPartStride));
    if (!g->part2.p) BSH_HIP(g->part2.alloc((size_t)((ncu + kRedGroup - 1) / kRedGroup) * kPartStride));
    hipLaunchKernelGGL(k_guide_wsum, dim3(kWsumParts), dim3(kWsumBlock), 0, st, rec, srow, n, g->wsum.p + 1);
    hipLaunchKernelGGL(k_guide_wsum_final, dim3(1), dim3(kWsumBlock), 0, st, g->wsum.p + 1, g->wsum.p);
    hipLaunchKernelGGL(k_guide_train_fused, dim3(nb), dim3(64 * kTrWaves), 0, st, g->net(), rec, srow, n,
                       (const float*)g->wsum.p, g->part.p);
    const dim3 rg((kGParams + 1 + 255) / 256, ngroups);
    hipLaunchKernelGGL(k_guide_reduce_part, rg, dim3(256), 0, st, (const float*)g->part.p, (int)nb, g->part2.p);
    hipLaunchKernelGGL(k_guide_reduce_train, dim3(rg.x), dim3(256), 0, st, (const float*)g->part2.p, (int)ngroups,
                       1.f / (float)n, g->grad.p, g->loss.p);
    BSH_HIP(hipGetLastError());
    return BSH_OK;
}

// FP32 mode: the buffers of one call (features, the four hidden activations, outputs, the
// output gradient, two ping-pong hidden gradients, the parameter gradient). Allocated per
// call: FP32 mode is a verification path.
constexpr size_t kF32MaxRows = (size_t)1 << 22;
struct F32Work
{
    DevBuf<float> X, H, out, dY, dA, dB, grad;
};

static dim3 f32_grid(size_t threads) { return dim3((uint32_t)((threads + 255) / 256)); }

static bsh_status f32_forward(bsh_guide* g, const float4* in, int stride, uint32_t n, F32Work& w, float* out,
                              hipStream_t st)
{
    BSH_HIP(w.X.alloc((size_t)n * kGIn));
    BSH_HIP(w.H.alloc((size_t)n * kGHid * 4));
    const float* P = g->p.p;
    const size_t hs = (size_t)n * kGHid;
    hipLaunchKernelGGL(k_f32_encode, f32_grid((size_t)n * kGIn), dim3(256), 0, st, in, stride, n, w.X.p);
    hipLaunchKernelGGL(k_f32_dense, f32_grid(hs), dim3(256), 0, st, (const float*)w.X.p, kGIn, P + kGW1, P + kGB1,
                       kGHid, n, 1, w.H.p);
    const int wo[3] = {kGW2, kGW3, kGW4}, bo[3] = {kGB2, kGB3, kGB4};
    for (int l = 0; l < 3; ++l)
        hipLaunchKernelGGL(k_f32_dense, f32_grid(hs), dim3(256), 0, st, (const float*)(w.H.p + l * hs), kGHid,
                           P + wo[l], P + bo[l], kGHid, n, 1, w.H.p + (l + 1) * hs);
    hipLaunchKernelGGL(k_f32_dense, f32_grid((size_t)n * kGOut), dim3(256), 0, st, (const float*)(w.H.p + 3 * hs),
                       kGHid, P + kGW5, P + kGB5, kGOut, n, 0, out);
    BSH_HIP(hipGetLastError());
    return BSH_OK;
}

// FP32-mode forward + backward on `rec` (3 float4 per row): the batch gradient in w.grad,
// the weighted loss (times rows) in g->loss. The loss weights and their normalisation are
// the FP16 path's (same kernels).
static bsh_status f32_gradients(bsh_guide* g, const float4* rec, uint32_t n, F32Work& w, hipStream_t st)
{
    BSH_HIP(w.out.alloc((size_t)n * kGOut));
    BSH_HIP(w.dY.alloc((size_t)n * kGOut));
    BSH_HIP(w.dA.alloc((size_t)n * kGHid));
    BSH_HIP(w.dB.alloc((size_t)n * kGHid));
    BSH_HIP(w.grad.alloc(kGParams));
    hipLaunchKernelGGL(k_guide_zero, dim3(1), dim3(64), 0, st, g->wsum.p, 1);
    hipLaunchKernelGGL(k_guide_zero, dim3(1), dim3(64), 0, st, g->loss.p, 1);
    hipLaunchKernelGGL(k_guide_wsum, dim3(kWsumParts), dim3(kWsumBlock), 0, st, rec, 3, n, g->wsum.p + 1);
    hipLaunchKernelGGL(k_guide_wsum_final, dim3(1), dim3(kWsumBlock), 0, st, g->wsum.p + 1, g->wsum.p);
    bsh_status r = f32_forward(g, rec, 3, n, w, w.out.p, st);
    if (r != BSH_OK) return r;
    hipLaunchKernelGGL(k_guide_loss<float>, dim3((n + 255) / 256), dim3(256), 0, st, (const float*)w.out.p, rec, 3, n,
                       (const float*)g->wsum.p, w.dY.p, g->loss.p);
    const float* P = g->p.p;
    const size_t hs = (size_t)n * kGHid;
    const float* H[4] = {w.H.p, w.H.p + hs, w.H.p + 2 * hs, w.H.p + 3 * hs};
    const float inv = 1.f / (float)n;
    auto wgrad = [&](const float* X, int K, const float* dY, int N, int off) {
        hipLaunchKernelGGL(k_f32_wgrad, f32_grid((size_t)K * N + N), dim3(256), 0, st, X, K, dY, N, n, inv,
                           w.grad.p + off);
    };
    auto bwd = [&](const float* dY, int KOUT, int woff, const float* Hin, float* dX) {
        hipLaunchKernelGGL(k_f32_bwd, f32_grid(hs), dim3(256), 0, st, dY, KOUT, P + woff, Hin, kGHid, n, dX);
    };
    wgrad(H[3], kGHid, w.dY.p, kGOut, kGW5);
    bwd(w.dY.p, kGOut, kGW5, H[3], w.dA.p);    // layer 4's pre-activation gradient
    wgrad(H[2], kGHid, w.dA.p, kGHid, kGW4);
    bwd(w.dA.p, kGHid, kGW4, H[2], w.dB.p);    // layer 3
    wgrad(H[1], kGHid, w.dB.p, kGHid, kGW3);
    bwd(w.dB.p, kGHid, kGW3, H[1], w.dA.p);    // layer 2
    wgrad(H[0], kGHid, w.dA.p, kGHid, kGW2);
    bwd(w.dA.p, kGHid, kGW2, H[0], w.dB.p);    // layer 1
    wgrad(w.X.p, kGIn, w.dB.p, kGHid, kGW1);
    BSH_HIP(hipGetLastError());
    return BSH_OK;
}

}  // namespace bsh

extern "C" {

size_t bsh_guide_num_params(void) { return (size_t)kGParams; }

bsh_status bsh_guide_create(bsh_context* ctx, const bsh_guide_params* prm, bsh_guide** out)
{
    if (!ctx || !out) return fail(BSH_ERR_INVALID, "bsh_guide_create: NULL argument");
    BSH_HIP(hipSetDevice(ctx->device));
    bsh_guide* g = new bsh_guide();
    g->ctx = ctx;
    bsh_guide_params d{1u, 5e-3f, 0.9f, 0.999f, 1e-8f, 0.5f, 0, 0};
    g->prm = prm ? *prm : d;
    if (!(g->prm.bsdf_fraction > 0.f && g->prm.bsdf_fraction <= 1.f) || g->prm.guided_bounces < 0)
    {
        delete g;
        return fail(BSH_ERR_INVALID, "bsh_guide_create: the BSDF fraction must be in (0, 1] (unbiasedness) and "
                                     "guided_bounces >= 0");
    }
    g->host_params.assign(kGParams, 0.f);
    // He-uniform weights (limit sqrt(6 / fan_in)), zero biases, from PCG32(seed)
    uint64_t st = 0x853c49e6748fea9bULL ^ ((uint64_t)g->prm.seed << 1);
    auto U = [&]() { return (pcg32(st) >> 8) * (1.f / 16777216.f); };
    const int wo[5] = {kGW1, kGW2, kGW3, kGW4, kGW5}, fin[5] = {kGIn, kGHid, kGHid, kGHid, kGHid},
              fout[5] = {kGHid, kGHid, kGHid, kGHid, kGOut};
    for (int l = 0; l < 5; ++l)
    {
        float lim = std::sqrt(6.f / (l == 0 ? 91.f : (float)fin[l]));
        for (int k = 0; k < fin[l] * fout[l]; ++k) g->host_params[wo[l] + k] = (2.f * U() - 1.f) * lim;
    }
    // the padding features are always zero; keep their weights zero
    for (int k = 0; k < kGIn; ++k)
        if (guide_pad_feature(k))
            for (int n = 0; n < kGHid; ++n) g->host_params[kGW1 + k * kGHid + n] = 0.f;
    bsh_status r = BSH_OK;
    auto alloc = [&]() -> bsh_status {
        BSH_HIP(g->p.alloc(kGParams));
        BSH_HIP(g->m.alloc(kGParams));
        BSH_HIP(g->v.alloc(kGParams));
        BSH_HIP(g->grad.alloc(kGParams));
        BSH_HIP(g->bias.alloc(kBiasTotal));
        BSH_HIP(g->wf.alloc(kFTotal));
        BSH_HIP(g->wb.alloc(kBTotal));
        BSH_HIP(g->loss.alloc(1));
        BSH_HIP(g->wsum.alloc(1 + kWsumParts));   // the sum, then the partial sums
        BSH_HIP(hipMemset(g->m.p, 0, kGParams * 4));
        BSH_HIP(hipMemset(g->v.p, 0, kGParams * 4));
        return BSH_OK;
    };
    r = alloc();
    if (r == BSH_OK) r = upload_params(g);
    if (r != BSH_OK)
    {
        delete g;
        return r;
    }
    *out = g;
    return BSH_OK;
}

bsh_status bsh_guide_destroy(bsh_guide* g)
{
    delete g;
    return BSH_OK;
}

bsh_status bsh_guide_get_params(const bsh_guide* g, float* host)
{
    if (!g || !host) return fail(BSH_ERR_INVALID, "bsh_guide_get_params: NULL argument");
    BSH_HIP(hipMemcpy(host, g->p.p, kGParams * 4, hipMemcpyDeviceToHost));
    return BSH_OK;
}

bsh_status bsh_guide_set_params(bsh_guide* g, const float* host)
{
    if (!g || !host) return fail(BSH_ERR_INVALID, "bsh_guide_set_params: NULL argument");
    g->host_params.assign(host, host + kGParams);
    BSH_HIP(hipMemset(g->m.p, 0, kGParams * 4));
    BSH_HIP(hipMemset(g->v.p, 0, kGParams * 4));
    g->step = 0;
    return upload_params(g);
}

bsh_status bsh_guide_get_state(const bsh_guide* g, float* params, float* m1, float* m2, int64_t* step)
{
    if (!g || !params || !m1 || !m2 || !step) return fail(BSH_ERR_INVALID, "bsh_guide_get_state: NULL argument");
    BSH_HIP(hipSetDevice(g->ctx->device));
    BSH_HIP(hipMemcpy(params, g->p.p, kGParams * 4, hipMemcpyDeviceToHost));
    BSH_HIP(hipMemcpy(m1, g->m.p, kGParams * 4, hipMemcpyDeviceToHost));
    BSH_HIP(hipMemcpy(m2, g->v.p, kGParams * 4, hipMemcpyDeviceToHost));
    *step = g->step;
    return BSH_OK;
}

bsh_status bsh_guide_set_state(bsh_guide* g, const float* params, const float* m1, const float* m2, int64_t step)
{
    if (!g || !params || !m1 || !m2) return fail(BSH_ERR_INVALID, "bsh_guide_set_state: NULL argument");
    if (step < 0 || step > (int64_t)1 << 30) return fail(BSH_ERR_INVALID, "bsh_guide_set_state: step out of range");
    BSH_HIP(hipSetDevice(g->ctx->device));
    g->host_params.assign(params, params + kGParams);
    BSH_HIP(hipMemcpy(g->m.p, m1, kGParams * 4, hipMemcpyHostToDevice));
    BSH_HIP(hipMemcpy(g->v.p, m2, kGParams * 4, hipMemcpyHostToDevice));
    g->step = (int)step;
    return upload_params(g);
}

bsh_status bsh_guide_set_bsdf_fraction(bsh_guide* g, float bsdf_fraction)
{
    if (!g) return fail(BSH_ERR_INVALID, "bsh_guide_set_bsdf_fraction: NULL guide");
    if (!(bsdf_fraction > 0.f && bsdf_fraction <= 1.f))
        return fail(BSH_ERR_INVALID, "bsh_guide_set_bsdf_fraction: the BSDF fraction must be in (0, 1]");
    g->prm.bsdf_fraction = bsdf_fraction;
    return BSH_OK;
}

bsh_status bsh_guide_forward(bsh_guide* g, const float* inputs, size_t n, float* out, void* stream)
{
    if (!g || (n && (!inputs || !out))) return fail(BSH_ERR_INVALID, "bsh_guide_forward: NULL argument");
    if (n >= ((size_t)1 << 31)) return fail(BSH_ERR_UNSUPPORTED, "bsh_guide_forward: too many rows");
    BSH_HIP(hipSetDevice(g->ctx->device));
    return guide_forward(g, reinterpret_cast<const float4*>(inputs), 2, (uint32_t)n, out, (hipStream_t)stream);
}

bsh_status bsh_guide_gradient(bsh_guide* g, const float* records, size_t n, float* grad, float* loss)
{
    if (!g || !records || n == 0) return fail(BSH_ERR_INVALID, "bsh_guide_gradient: bad argument");
    if (n >= ((size_t)1 << 30)) return fail(BSH_ERR_UNSUPPORTED, "bsh_guide_gradient: too many rows");
    BSH_HIP(hipSetDevice(g->ctx->device));
    bsh_status r = guide_gradients(g, reinterpret_cast<const float4*>(records), 3, (uint32_t)n, nullptr);
    if (r != BSH_OK) return r;
    BSH_HIP(hipDeviceSynchronize());
    if (grad) BSH_HIP(hipMemcpy(grad, g->grad.p, kGParams * 4, hipMemcpyDeviceToHost));
    if (loss)
    {
        float l = 0.f;
        BSH_HIP(hipMemcpy(&l, g->loss.p, 4, hipMemcpyDeviceToHost));
        *loss = l / (float)n;
    }
    return BSH_OK;
}

bsh_status bsh_guide_forward_fp32(bsh_guide* g, const float* inputs, size_t n, float* out, void* stream)
{
    if (!g || (n && (!inputs || !out))) return fail(BSH_ERR_INVALID, "bsh_guide_forward_fp32: NULL argument");
    if (n == 0) return BSH_OK;
    if (n > kF32MaxRows) return fail(BSH_ERR_UNSUPPORTED, "bsh_guide_forward_fp32: more than 2^22 rows");
    BSH_HIP(hipSetDevice(g->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    F32Work w;
    bsh_status r = f32_forward(g, reinterpret_cast<const float4*>(inputs), 2, (uint32_t)n, w, out, st);
    BSH_HIP(hipStreamSynchronize(st));   // the work buffers are released on return
    return r;
}

bsh_status bsh_guide_gradient_fp32(bsh_guide* g, const float* records, size_t n, float* grad, float* loss)
{
    if (!g || !records || n == 0) return fail(BSH_ERR_INVALID, "bsh_guide_gradient_fp32: bad argument");
    if (n > kF32MaxRows) return fail(BSH_ERR_UNSUPPORTED, "bsh_guide_gradient_fp32: more than 2^22 rows");
    BSH_HIP(hipSetDevice(g->ctx->device));
    F32Work w;
    bsh_status r = f32_gradients(g, reinterpret_cast<const float4*>(records), (uint32_t)n, w, nullptr);
    BSH_HIP(hipDeviceSynchronize());
    if (r != BSH_OK) return r;
    if (grad) BSH_HIP(hipMemcpy(grad, w.grad.p, kGParams * 4, hipMemcpyDeviceToHost));
    if (loss)
    {
        float l = 0.f;
        BSH_HIP(hipMemcpy(&l, g->loss.p, 4, hipMemcpyDeviceToHost));
        *loss = l / (float)n;
    }
    return BSH_OK;
}

// One Adam step on n records read `stride` records apart.
static bsh_status train_step(bsh_guide* g, const float* records, size_t stride, size_t n, float* loss,
                             hipStream_t st)
{
    bsh_status r = guide_gradients(g, reinterpret_cast<const float4*>(records), (int)(3 * stride), (uint32_t)n, st);
    if (r != BSH_OK) return r;
    g->step++;
    float c1 = 1.f - std::pow(g->prm.beta1, (float)g->step), c2 = 1.f - std::pow(g->prm.beta2, (float)g->step);
    hipLaunchKernelGGL(k_guide_adam, dim3((kGParams + 255) / 256), dim3(256), 0, st, g->p.p, g->m.p, g->v.p, g->grad.p,
                       kGParams, g->prm.lr, g->prm.beta1, g->prm.beta2, g->prm.eps, c1, c2, g->wf.p, g->wb.p,
                       g->bias.p);
    BSH_HIP(hipGetLastError());
    if (loss)
    {
        float l = 0.f;
        BSH_HIP(hipMemcpyAsync(&l, g->loss.p, 4, hipMemcpyDeviceToHost, st));
        BSH_HIP(hipStreamSynchronize(st));
        *loss = l / (float)n;
    }
    return BSH_OK;
}

bsh_status bsh_guide_train_step(bsh_guide* g, const float* records, size_t n, float* loss, void* stream)
{
    if (!g || !records) return fail(BSH_ERR_INVALID, "bsh_guide_train_step: NULL argument");
    if (n == 0) return BSH_OK;
    if (n >= ((size_t)1 << 30)) return fail(BSH_ERR_UNSUPPORTED, "bsh_guide_train_step: too many rows");
    BSH_HIP(hipSetDevice(g->ctx->device));
    return train_step(g, records, 1, n, loss, (hipStream_t)stream);
}

bsh_status bsh_guide_train(bsh_guide* g, const float* records, size_t n, int steps, size_t batch, float* loss,
                           void* stream)
{
    if (!g || !records) return fail(BSH_ERR_INVALID, "bsh_guide_train: NULL argument");
    if (steps < 0 || batch == 0) return fail(BSH_ERR_INVALID, "bsh_guide_train: bad steps or batch");
    if (n == 0 || steps == 0) return BSH_OK;
    if (n >= ((size_t)1 << 34)) return fail(BSH_ERR_UNSUPPORTED, "bsh_guide_train: too many records");
    BSH_HIP(hipSetDevice(g->ctx->device));
    // step s takes every stride-th record from offset s mod stride: each batch spans the whole
    // record array (records are in path order, i.e. image order), and successive steps see
    // different records
    const size_t stride = std::max<size_t>(1, n / batch);
    for (int s = 0; s < steps; ++s)
    {
        const size_t off = (size_t)s % stride;
        const size_t rows = std::min(batch, (n - off + stride - 1) / stride);
        bsh_status r = train_step(g, records + off * 12, stride, rows, s + 1 == steps ? loss : nullptr,
                                  (hipStream_t)stream);
        if (r != BSH_OK) return r;
    }
    return BSH_OK;
}

bsh_status bsh_diag_mfma(bsh_context* ctx, const uint16_t* a, const uint16_t* b, float* d)
{
    if (!ctx || !a || !b || !d) return fail(BSH_ERR_INVALID, "bsh_diag_mfma: NULL argument");
    BSH_HIP(hipSetDevice(ctx->device));
    DevBuf<_Float16> da, db;
    DevBuf<float> dd;
    BSH_HIP(da.alloc(256));
    BSH_HIP(db.alloc(256));
    BSH_HIP(dd.alloc(256));
    BSH_HIP(hipMemcpy(da.p, a, 512, hipMemcpyHostToDevice));
    BSH_HIP(hipMemcpy(db.p, b, 512, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_diag_mfma, dim3(1), dim3(64), 0, 0, da.p, db.p, dd.p);
    BSH_HIP(hipGetLastError());
    BSH_HIP(hipMemcpy(d, dd.p, 1024, hipMemcpyDeviceToHost));
    return BSH_OK;
}

bsh_status bsh_diag_guide(bsh_context* ctx, size_t n, const float* outputs, const float* u, const float* dirs,
                          float* out)
{
    if (!ctx || (n && (!outputs || !u || !dirs || !out))) return fail(BSH_ERR_INVALID, "bsh_diag_guide: NULL argument");
    if (n == 0) return BSH_OK;
    if (n >= ((size_t)1 << 31)) return fail(BSH_ERR_UNSUPPORTED, "bsh_diag_guide: too many rows");
    BSH_HIP(hipSetDevice(ctx->device));
    DevBuf<float> d_o, d_u, d_d, d_r;
    BSH_HIP(d_o.alloc(n * kGOut));
    BSH_HIP(d_u.alloc(n * 3));
    BSH_HIP(d_d.alloc(n * 3));
    BSH_HIP(d_r.alloc(n * 8));
    BSH_HIP(hipMemcpy(d_o.p, outputs, n * kGOut * 4, hipMemcpyHostToDevice));
    BSH_HIP(hipMemcpy(d_u.p, u, n * 12, hipMemcpyHostToDevice));
    BSH_HIP(hipMemcpy(d_d.p, dirs, n * 12, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_diag_guide, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0, (uint32_t)n, d_o.p, d_u.p,
                       d_d.p, d_r.p);
    BSH_HIP(hipGetLastError());
    BSH_HIP(hipMemcpy(out, d_r.p, n * 32, hipMemcpyDeviceToHost));
    return BSH_OK;
}

}  // extern "C"
