// kpax_mlp.cuh — device pieces of the learned control proposal: the
// network's input features, bf16 conversion, the MFMA tile primitive and the
// Gaussian sampler that turns a node's (mean, sigma) into a control.
//
// Network: 80 inputs (79 features + one zero pad) -> 128 GELU -> 128 GELU ->
// 8 outputs (mean of the 4 normalised controls, log sigma of each).  Weights
// are bf16 on the matrix cores with fp32 accumulation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "kpax_device.cuh"

namespace kpax {

constexpr int kMlpIn = 80, kMlpHidden = 128, kMlpOut = 8;
constexpr int kMlpFeatures = 79;
// flat fp32 parameter layout: W1 [80][128], b1 [128], W2 [128][128], b2 [128],
// W3 [128][8], b3 [8] (W[k][n]: input k, output n)
constexpr int kMlpW1 = 0, kMlpB1 = kMlpW1 + kMlpIn * kMlpHidden, kMlpW2 = kMlpB1 + kMlpHidden,
              kMlpB2 = kMlpW2 + kMlpHidden * kMlpHidden, kMlpW3 = kMlpB2 + kMlpHidden,
              kMlpB3 = kMlpW3 + kMlpHidden * kMlpOut, kMlpParams = kMlpB3 + kMlpOut;
constexpr int kOccDim = 16;  // occupancy bitmap: 16^3 cells over the workspace

// bf16 copies of the network on the device.  The part up to `b3` is the
// inference layout (transposed weights, rows padded so the matrix-core reads
// of 32 lanes spread over the LDS banks) and is copied verbatim into LDS;
// the rest is what the backward pass reads.
constexpr int kLdW1 = 88, kLdH = 136, kLdW3b = 16;

// Inference weights as matrix-core operand fragments (v_mfma_f32_16x16x32_bf16,
// weights as the A operand): fragment [t][s] holds, for lane l = j + 16 g, the
// eight bf16 weights of output unit 16 t + j over inputs in(s, g, 0..7), so a
// wave's 64 lanes read 1 KB of consecutive bytes.  Layer 1 reads the features
// in natural order, in(s, g, e) = 32 s + 8 g + e (80 padded to 96 with zero
// weights); layers 2 and 3 read the previous layer's units in the order the
// accumulators leave them, frag_in_hidden(s, g, e).  Layer 3 has one tile of
// 16 output rows, rows 8..15 zero.
constexpr int kFragT = kMlpHidden / 16;   // 16-unit tiles of a hidden layer
constexpr int kFragS1 = 3;                 // 32-input steps of layer 1
constexpr int kFragS2 = kMlpHidden / 32;   // 32-input steps of layers 2 and 3
struct MlpFrag {
    uint4 w1[kFragT][kFragS1][64];
    uint4 w2[kFragT][kFragS2][64];
    uint4 w3[kFragS2][64];
};

// Hidden unit fed into k-slot (g, e) of step s: the units 16 t + 4 g + (e % 4)
// that lane group g holds after a layer, of tiles t = 2 s (e < 4) and 2 s + 1.
__host__ __device__ constexpr int frag_in_hidden(int s, int g, int e)
{
    return 32 * s + 16 * (e / 4) + 4 * g + e % 4;
}

struct MlpDev {
    uint16_t w1t[kMlpHidden][kLdW1];  // W1^T [out][in], in < 80
    uint16_t w2t[kMlpHidden][kLdH];   // W2^T [out][in]
    uint16_t w3t[32][kLdH];           // W3^T [out][in], out < 8
    float b1[kMlpHidden], b2[kMlpHidden], b3[32];
    uint16_t w2[kMlpHidden][kLdH];    // W2 [in][out]
    uint16_t w3[kMlpHidden][kLdW3b];  // W3 [in][out], out < 8
    MlpFrag frag;                     // inference fragments (with b1, b2, b3 above)
};
constexpr size_t kMlpLdsWeights = offsetof(MlpDev, w2);

// Feature scaling, folded on the host from the configuration.
struct FeatParams {
    float mid[3], half[3];  // workspace centre and half extent
    float occ_mn[3], occ_cell[3];
    float inv_pi, inv_vel, rate_scale;
    int32_t S;
};

namespace dev {

// ---------------------------------------------------------------------------
// bf16 as raw 16-bit patterns (round to nearest even; NaN stays NaN)
// ---------------------------------------------------------------------------
// fp32 -> bf16, round to nearest even (NaN stays NaN): one conversion
// instruction on gfx950, the same rounding written out for the host build.
__device__ __forceinline__ uint16_t f2bf(float f)
{
#ifndef KPAX_HOST_EMULATION
    return __builtin_bit_cast(uint16_t, (__bf16)f);
#else
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
#endif
}

__device__ __forceinline__ float bf2f(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

// ---------------------------------------------------------------------------
// Matrix-core tile: D(32x32, fp32) = A(32x16) * B(16x32) + C, bf16 operands.
// Lane l holds A row (l % 32), k = 8 (l / 32) .. +7, and B column (l % 32),
// the same k; its 16 results are column (l % 32), rows
// 8 (i / 4) + 4 (l / 32) + (i % 4), i = 0..15.
// ---------------------------------------------------------------------------
#ifndef KPAX_HOST_EMULATION
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f32x16 mfma_32x32x16(uint4 a, uint4 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}
// D(16x16, fp32) = A(16x32) * B(32x16) + C, bf16 operands.  Lane l = j + 16 g
// holds A row j and B column j over the eight k-slots of group g, and receives
// D column j, rows 4 g + e (e = 0..3).
__device__ __forceinline__ f32x4 mfma_16x16x32(uint4 a, uint4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}
#endif

__device__ __forceinline__ int mfma_row(int i, int lane) { return 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3); }

// GELU, tanh form: 0.5 x (1 + tanh u) written as x / (1 + e^(-2u)), with the hardware's
// base-2 exponential and reciprocal (about 1 ulp each; the result is rounded
// to bf16 right after, and the training gradient uses the tanh form).
__device__ __forceinline__ float gelu(float x)
{
    const float k = 0.7978845608028654f;  // sqrt(2 / pi)
    const float u = k * (x + 0.044715f * x * x * x);
#ifndef KPAX_HOST_EMULATION
    const float e = __builtin_amdgcn_exp2f(-2.8853900817779268f * u);  // e^(-2u) = 2^(-2u log2 e)
    return x * __builtin_amdgcn_rcpf(1.0f + e);
#else
    return x / (1.0f + exp2f(-2.8853900817779268f * u));
#endif
}

__device__ __forceinline__ float gelu_grad(float x)
{
    const float k = 0.7978845608028654f;
    const float u = k * (x + 0.044715f * x * x * x);
    const float th = tanhf(u);
    return 0.5f * (1.0f + th) + 0.5f * x * (1.0f - th * th) * k * (1.0f + 3.0f * 0.044715f * x * x);
}

// ---------------------------------------------------------------------------
// Features of one node: scaled state, goal-relative position, 4x4x4
// occupancy window around the position (cells outside the workspace count
// as occupied).  f[79] = 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int occ_cell_index(float x, float mn, float cell)
{
    float v = (x - mn) / cell + 0.5f;
    if(!(v >= -8.0f)) v = -8.0f;  // NaN too
    if(v > 24.0f) v = 24.0f;
    return (int)floorf(v) - 2;
}

__device__ __forceinline__ float occ_bit(const uint32_t* occ, int a, int b, int c)
{
    if(a < 0 || b < 0 || c < 0 || a >= kOccDim || b >= kOccDim || c >= kOccDim) return 1.0f;
    const int i = (a * kOccDim + b) * kOccDim + c;
    return (occ[i >> 5] >> (i & 31)) & 1u ? 1.0f : 0.0f;
}

// x: the node's state (S values); writes 80 features.
__device__ __forceinline__ void features(const FeatParams& F, const uint32_t* occ, const float* x, const float* goal,
                                         float* f)
{
#pragma unroll
    for(int d = 0; d < 3; ++d) f[d] = (x[d] - F.mid[d]) / F.half[d];
#pragma unroll
    for(int d = 0; d < 3; ++d) f[3 + d] = x[3 + d] * F.inv_pi;
#pragma unroll
    for(int d = 0; d < 3; ++d) f[6 + d] = x[6 + d] * F.inv_vel;
#pragma unroll
    for(int d = 0; d < 3; ++d) f[9 + d] = x[9 + d] * F.rate_scale;
#pragma unroll
    for(int d = 0; d < 3; ++d) f[12 + d] = (goal[d] - x[d]) / F.half[d];
    const int a0 = occ_cell_index(x[0], F.occ_mn[0], F.occ_cell[0]);
    const int b0 = occ_cell_index(x[1], F.occ_mn[1], F.occ_cell[1]);
    const int c0 = occ_cell_index(x[2], F.occ_mn[2], F.occ_cell[2]);
    for(int i = 0; i < 4; ++i)
        for(int j = 0; j < 4; ++j)
            for(int k = 0; k < 4; ++k) f[15 + 16 * i + 4 * j + k] = occ_bit(occ, a0 + i, b0 + j, c0 + k);
    f[79] = 0.0f;
}

// ---------------------------------------------------------------------------
// Network inference on 16 nodes per wave.  The weights are the A operand and
// the nodes the B operand, so after a layer lane (j, g) = j + 16 g holds units
// 16 t + 4 g + e (e = 0..3) of node j for every tile t: exactly the k-slots
// (g, 0..7) of the next layer's steps in the order frag_in_hidden() gives, so
// the activations never leave the lane's registers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pack_bf16x2(uint16_t lo, uint16_t hi) { return lo | (uint32_t)hi << 16; }

// The lane's share of its node's features as layer-1 operands: features
// 32 s + 8 g + e (e = 0..7) of node j, s = 0..2, as bf16, equal to what
// features() gives (features 79..95 zero).  Lane group g reads the occupancy
// plane a0 + g of the window (four 16-cell rows); the four groups of a node
// exchange their planes.  Every lane of the wave calls it.
__device__ __forceinline__ void features16(const FeatParams& F, const uint32_t* occ, const float* x,
                                           const float* goal, uint4 (&xb)[kFragS1])
{
    const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    float sf[15];
#pragma unroll
    for(int d = 0; d < 3; ++d) sf[d] = (x[d] - F.mid[d]) / F.half[d];
#pragma unroll
    for(int d = 0; d < 3; ++d) sf[3 + d] = x[3 + d] * F.inv_pi;
#pragma unroll
    for(int d = 0; d < 3; ++d) sf[6 + d] = x[6 + d] * F.inv_vel;
#pragma unroll
    for(int d = 0; d < 3; ++d) sf[9 + d] = x[9 + d] * F.rate_scale;
#pragma unroll
    for(int d = 0; d < 3; ++d) sf[12 + d] = (goal[d] - x[d]) / F.half[d];
    const int a  = occ_cell_index(x[0], F.occ_mn[0], F.occ_cell[0]) + g;
    const int b0 = occ_cell_index(x[1], F.occ_mn[1], F.occ_cell[1]);
    const int c0 = occ_cell_index(x[2], F.occ_mn[2], F.occ_cell[2]);
    uint32_t plane = 0;  // bit 4 jj + k: cell (a, b0 + jj, c0 + k)
#pragma unroll
    for(int jj = 0; jj < 4; ++jj)
    {
        const int b      = b0 + jj;
        const bool in    = a >= 0 && a < kOccDim && b >= 0 && b < kOccDim;
        const int rb     = in ? a * kOccDim + b : 0;  // row of 16 cells along c
        const uint32_t w = occ[rb >> 1];
        const uint32_t row = in ? (w >> ((rb & 1) * 16)) & 0xffffu : 0xffffu;
        // cells c < 0 and c >= 16 lie outside the workspace and count as occupied
        const uint64_t ext = 0xffff00000000ull | (uint64_t)row << 16 | 0xffffull;
        plane |= ((uint32_t)(ext >> (c0 + 16)) & 0xfu) << (4 * jj);
    }
    const uint32_t p0 = __shfl(plane, j), p1 = __shfl(plane, j + 16), p2 = __shfl(plane, j + 32),
                   p3 = __shfl(plane, j + 48);
    const uint64_t M = (uint64_t)(p0 | p1 << 16) | (uint64_t)(p2 | p3 << 16) << 32;  // bit b: feature 15 + b
#pragma unroll
    for(int s = 0; s < kFragS1; ++s)
    {
        const int sh      = 32 * s + 8 * g - 15;  // bit of M that is the group's first feature
        const uint32_t by = sh >= 64 ? 0u : sh >= 0 ? (uint32_t)(M >> sh) & 0xffu : (uint32_t)(M << -sh) & 0xffu;
        uint16_t v[8];
#pragma unroll
        for(int e = 0; e < 8; ++e)
        {
            v[e] = (by >> e) & 1u ? (uint16_t)0x3f80u : (uint16_t)0;  // bf16 1.0 or 0.0
            if(s == 0)
            {
                if(g == 0) v[e] = f2bf(sf[e]);
                else if(g == 1 && e < 7) v[e] = f2bf(sf[8 + e]);
            }
        }
        xb[s] = make_uint4(pack_bf16x2(v[0], v[1]), pack_bf16x2(v[2], v[3]), pack_bf16x2(v[4], v[5]),
                           pack_bf16x2(v[6], v[7]));
    }
}

// One hidden layer over the lane's operands x (S steps): GELU(W x + b) of the
// 128 units, rounded to bf16 and packed as the next layer's operands.
template<int S>
__device__ __forceinline__ void hidden16(const uint4 (*w)[S][64], const float* bias, const uint4 (&x)[S],
                                         uint4 (&y)[kFragS2])
{
    const int lane = threadIdx.x & 63, g = lane >> 4;
    f32x4 acc[kFragT];
#pragma unroll
    for(int t = 0; t < kFragT; ++t)
    {
        acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for(int s = 0; s < S; ++s) acc[t] = mfma_16x16x32(w[t][s][lane], x[s], acc[t]);
    }
#pragma unroll
    for(int s = 0; s < kFragS2; ++s)
    {
        const float4 b0 = *(const float4*)&bias[32 * s + 4 * g];       // tile 2 s
        const float4 b1 = *(const float4*)&bias[32 * s + 16 + 4 * g];  // tile 2 s + 1
        const f32x4 u = acc[2 * s], v = acc[2 * s + 1];
        y[s] = make_uint4(pack_bf16x2(f2bf(gelu(u[0] + b0.x)), f2bf(gelu(u[1] + b0.y))),
                          pack_bf16x2(f2bf(gelu(u[2] + b0.z)), f2bf(gelu(u[3] + b0.w))),
                          pack_bf16x2(f2bf(gelu(v[0] + b1.x)), f2bf(gelu(v[1] + b1.y))),
                          pack_bf16x2(f2bf(gelu(v[2] + b1.z)), f2bf(gelu(v[3] + b1.w))));
    }
}

// The network on the wave's 16 nodes: lane (j, g) returns raw outputs
// 4 g .. 4 g + 3 of node j (g < 2; groups 2 and 3 hold zero rows).
__device__ __forceinline__ f32x4 mlp_forward16(const MlpDev* W, const uint4 (&x)[kFragS1])
{
    const int lane = threadIdx.x & 63, g = lane >> 4;
    uint4 h1[kFragS2], h2[kFragS2];
    hidden16<kFragS1>(W->frag.w1, W->b1, x, h1);
    hidden16<kFragS2>(W->frag.w2, W->b2, h1, h2);
    f32x4 o = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for(int s = 0; s < kFragS2; ++s) o = mfma_16x16x32(W->frag.w3[s][lane], h2[s], o);
    const float4 b = *(const float4*)&W->b3[4 * g];  // b3 is zero past output 8
    o[0] += b.x;
    o[1] += b.y;
    o[2] += b.z;
    o[3] += b.w;
    return o;
}

// ---------------------------------------------------------------------------
// Sampler.  ln_f32: natural log of x in (0, 1] from its exponent and
// 2 atanh((m - 1) / (m + 1)) of the mantissa m in [sqrt(1/2), sqrt(2)) (odd
// series to s^11, |s| <= 0.172), every operation written out so the host
// restatement is bit-identical.  Box-Muller turns two 24-bit uniforms into
// two standard normals.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float ln_f32(float x)
{
    const uint32_t bits = __builtin_bit_cast(uint32_t, x);
    int e               = (int)(bits >> 23) - 127;
    float m             = __builtin_bit_cast(float, (bits & 0x007fffffu) | 0x3f800000u);
    if(m > 1.41421356f)
    {
        m = m * 0.5f;
        e = e + 1;
    }
    const float s  = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p        = __builtin_fmaf(s2, 0.09090909f, 0.11111111f);
    p              = __builtin_fmaf(s2, p, 0.14285714f);
    p              = __builtin_fmaf(s2, p, 0.2f);
    p              = __builtin_fmaf(s2, p, 0.33333333f);
    const float lm = 2.0f * __builtin_fmaf(s * s2, p, s);
    return __builtin_fmaf((float)e, 0.69314718f, lm);
}

__device__ __forceinline__ void box_muller(uint32_t w0, uint32_t w1, float* z0, float* z1)
{
    const float u1 = 1.0f - u01(w0);  // (0, 1]
    const float u2 = u01(w1);
    const float r  = __builtin_sqrtf(-2.0f * ln_f32(u1));
    float sn, cs;
    sincos_f32(6.28318531f * u2, &sn, &cs);
    *z0 = r * cs;
    *z1 = r * sn;
}

// Control from the node's Gaussian: normalised u = clamp(mean + sigma z,
// -1, 1) mapped onto [ctrl_min, ctrl_min + ctrl_rng].
template<int C>
__device__ __forceinline__ void propose_control(const float* mean, const float* sigma, uint4 w, const ModelParams& P,
                                                float* u)
{
    float z[4];
    box_muller(w.x, w.y, &z[0], &z[1]);
    box_muller(w.z, w.w, &z[2], &z[3]);
#pragma unroll
    for(int j = 0; j < C; ++j)
    {
        float un = mean[j] + sigma[j] * z[j];
        un       = un < -1.0f ? -1.0f : (un > 1.0f ? 1.0f : un);
        u[j]     = P.ctrl_min[j] + (un + 1.0f) * 0.5f * P.ctrl_rng[j];
    }
}

}  // namespace dev
}  // namespace kpax
