// kpax_mlp.hip — the learned control proposal on the matrix cores: occupancy
// bitmaps of the environments, the fused feature + 3-layer network pass that
// gives every new tree node a Gaussian over its children's controls, and the
// training kernels (forward, Gaussian negative log-likelihood, backward,
// Adam) with their C ABI.
//
// No counterpart in the reference, which draws every control uniformly
// (curand_uniform per dimension, statePropagatorSpatialHash.cu:236-241).  The
// proposal is an addition: a child takes the network's control with
// probability `mix` and the reference's uniform control otherwise, so every
// control the plain planner can draw keeps a positive density.
//
// Inference runs once per tree node, when the node enters the tree, not once
// per child: the 32 children of a repeat entry share their parent's input,
// so the network's output is the parent's and only the Gaussian draw differs.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "kpax_handle.h"
#include "kpax_mlp.cuh"

using namespace kpax;
using namespace kpax::dev;

namespace {

// Wave-local LDS exchange: every lane's LDS writes are visible to the
// wave's other lanes after this.
__device__ __forceinline__ void wave_lds_sync()
{
#ifndef KPAX_HOST_EMULATION
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#else
    emu_wave_barrier();
#endif
}

// ---------------------------------------------------------------------------
// Occupancy bitmap of an environment: cell (a, b, c) of the 16^3 lattice over
// the workspace is set when an obstacle overlaps its open box.
// ---------------------------------------------------------------------------
constexpr uint32_t kOccChunk = 32;  // obstacles per work-group row of the grid

// Grid (cells / 256, obstacle chunks): a lane tests its cell against chunks
// blockIdx.y, blockIdx.y + gridDim.y, ... of the obstacles and ORs the wave's
// verdicts into the zeroed bitmap.
__global__ __launch_bounds__(256) void k_occupancy(FeatParams F, const float4* lo, const float4* hi, uint32_t n,
                                                   uint32_t* bits)
{
    constexpr int kCells = kOccDim * kOccDim * kOccDim;
    const int cidx       = blockIdx.x * blockDim.x + threadIdx.x;
    bool hit             = false;
    if(cidx < kCells)
    {
        const int a = cidx / (kOccDim * kOccDim), b = (cidx / kOccDim) % kOccDim, c = cidx % kOccDim;
        const float cl[3] = {F.occ_mn[0] + (float)a * F.occ_cell[0], F.occ_mn[1] + (float)b * F.occ_cell[1],
                             F.occ_mn[2] + (float)c * F.occ_cell[2]};
        const float ch[3] = {F.occ_mn[0] + (float)(a + 1) * F.occ_cell[0],
                             F.occ_mn[1] + (float)(b + 1) * F.occ_cell[1],
                             F.occ_mn[2] + (float)(c + 1) * F.occ_cell[2]};
        for(uint32_t o0 = blockIdx.y * kOccChunk; o0 < n && !hit; o0 += gridDim.y * kOccChunk)
            for(uint32_t o = o0, oe = n - o0 < kOccChunk ? n : o0 + kOccChunk; o < oe && !hit; ++o)
            {
                const float4 p = lo[o], q = hi[o];
                hit = p.x < ch[0] && q.x > cl[0] && p.y < ch[1] && q.y > cl[1] && p.z < ch[2] && q.z > cl[2];
            }
    }
    const uint64_t m = __ballot(hit);
    const int lane   = threadIdx.x & 63;
    const uint32_t w = (uint32_t)(lane == 0 ? m : m >> 32);
    if(cidx < kCells && (lane == 0 || lane == 32) && w) atomicOr(&bits[cidx / 32], w);
}

// fp32 parameters (flat layout of kpax_mlp.cuh) -> bf16 device layouts.
__global__ __launch_bounds__(256) void k_pack_weights(const float* w, MlpDev* d)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for(int i = t; i < kMlpHidden * kLdW1; i += stride)
    {
        const int n = i / kLdW1, k = i % kLdW1;
        d->w1t[n][k] = k < kMlpIn ? f2bf(w[kMlpW1 + k * kMlpHidden + n]) : (uint16_t)0;
    }
    for(int i = t; i < kMlpHidden * kLdH; i += stride)
    {
        const int n = i / kLdH, k = i % kLdH;
        d->w2t[n][k] = k < kMlpHidden ? f2bf(w[kMlpW2 + k * kMlpHidden + n]) : (uint16_t)0;  // W2^T[out][in]
        d->w2[n][k]  = k < kMlpHidden ? f2bf(w[kMlpW2 + n * kMlpHidden + k]) : (uint16_t)0;  // W2[in][out]
    }
    for(int i = t; i < 32 * kLdH; i += stride)
    {
        const int n = i / kLdH, k = i % kLdH;
        d->w3t[n][k] = (n < kMlpOut && k < kMlpHidden) ? f2bf(w[kMlpW3 + k * kMlpOut + n]) : (uint16_t)0;
    }
    for(int i = t; i < kMlpHidden * kLdW3b; i += stride)
    {
        const int k = i / kLdW3b, n = i % kLdW3b;
        d->w3[k][n] = n < kMlpOut ? f2bf(w[kMlpW3 + k * kMlpOut + n]) : (uint16_t)0;
    }
    for(int i = t; i < kMlpHidden; i += stride)
    {
        d->b1[i] = w[kMlpB1 + i];
        d->b2[i] = w[kMlpB2 + i];
    }
    for(int i = t; i < 32; i += stride) d->b3[i] = i < kMlpOut ? w[kMlpB3 + i] : 0.0f;
    // inference fragments: element e of lane l = j + 16 g of fragment [tile][step]
    uint16_t* f1 = (uint16_t*)d->frag.w1;
    for(int i = t; i < kFragT * kFragS1 * 64 * 8; i += stride)
    {
        const int e = i & 7, l = (i >> 3) & 63, st = i >> 9;
        const int n = 16 * (st / kFragS1) + (l & 15), k = 32 * (st % kFragS1) + 8 * (l >> 4) + e;
        f1[i] = k < kMlpIn ? f2bf(w[kMlpW1 + k * kMlpHidden + n]) : (uint16_t)0;
    }
    uint16_t* f2 = (uint16_t*)d->frag.w2;
    for(int i = t; i < kFragT * kFragS2 * 64 * 8; i += stride)
    {
        const int e = i & 7, l = (i >> 3) & 63, st = i >> 9;
        const int n = 16 * (st / kFragS2) + (l & 15), k = frag_in_hidden(st % kFragS2, l >> 4, e);
        f2[i] = f2bf(w[kMlpW2 + k * kMlpHidden + n]);
    }
    uint16_t* f3 = (uint16_t*)d->frag.w3;
    for(int i = t; i < kFragS2 * 64 * 8; i += stride)
    {
        const int e = i & 7, l = (i >> 3) & 63, s = i >> 9;
        const int n = l & 15, k = frag_in_hidden(s, l >> 4, e);
        f3[i] = n < kMlpOut ? f2bf(w[kMlpW3 + k * kMlpOut + n]) : (uint16_t)0;
    }
}

// ---------------------------------------------------------------------------
// The training forward pass on one 32-row tile held by one wave.  xa holds
// the tile's inputs (bf16, 80 columns of kLdH); xb is scratch.  Returns, per
// lane, the 16 raw outputs of column (lane % 32) (< 8 meaningful) at the rows
// mfma_row(i, lane), and stores the pre-activations (fp32) and activations
// (bf16) of both hidden layers for the backward pass.
// ---------------------------------------------------------------------------
struct Keep {
    float* h1;      // [rows][128] pre-activations
    float* h2;
    uint16_t* a1;   // [rows][128] activations
    uint16_t* a2;
    uint32_t row0;  // first row of this tile in the arrays above
    uint32_t nrows; // valid rows of this tile
};

__device__ __forceinline__ void mlp_tile(const MlpDev& W, uint16_t (*xa)[kLdH], uint16_t (*xb)[kLdH], float* out,
                                         const Keep& keep)
{
    const int lane = threadIdx.x & 63, r = lane & 31, kh = (lane >> 5) * 8;
    f32x16 acc[4];
    // layer 1: 80 -> 128
#pragma unroll
    for(int nt = 0; nt < 4; ++nt)
#pragma unroll
        for(int i = 0; i < 16; ++i) acc[nt][i] = 0.0f;
#pragma unroll
    for(int ks = 0; ks < kMlpIn / 16; ++ks)
    {
        const uint4 a = *(const uint4*)&xa[r][ks * 16 + kh];
#pragma unroll
        for(int nt = 0; nt < 4; ++nt)
            acc[nt] = mfma_32x32x16(a, *(const uint4*)&W.w1t[nt * 32 + r][ks * 16 + kh], acc[nt]);
    }
#pragma unroll
    for(int nt = 0; nt < 4; ++nt)
#pragma unroll
        for(int i = 0; i < 16; ++i)
        {
            const int row = mfma_row(i, lane), col = nt * 32 + r;
            const float h = acc[nt][i] + W.b1[col];
            const uint16_t a = f2bf(gelu(h));
            xb[row][col] = a;
            if((uint32_t)row < keep.nrows)
                {
                    keep.h1[(size_t)(keep.row0 + row) * kMlpHidden + col] = h;
                    keep.a1[(size_t)(keep.row0 + row) * kMlpHidden + col] = a;
                }
        }
    wave_lds_sync();
    // layer 2: 128 -> 128
#pragma unroll
    for(int nt = 0; nt < 4; ++nt)
#pragma unroll
        for(int i = 0; i < 16; ++i) acc[nt][i] = 0.0f;
#pragma unroll
    for(int ks = 0; ks < kMlpHidden / 16; ++ks)
    {
        const uint4 a = *(const uint4*)&xb[r][ks * 16 + kh];
#pragma unroll
        for(int nt = 0; nt < 4; ++nt)
            acc[nt] = mfma_32x32x16(a, *(const uint4*)&W.w2t[nt * 32 + r][ks * 16 + kh], acc[nt]);
    }
#pragma unroll
    for(int nt = 0; nt < 4; ++nt)
#pragma unroll
        for(int i = 0; i < 16; ++i)
        {
            const int row = mfma_row(i, lane), col = nt * 32 + r;
            const float h = acc[nt][i] + W.b2[col];
            const uint16_t a = f2bf(gelu(h));
            xa[row][col] = a;
            if((uint32_t)row < keep.nrows)
                {
                    keep.h2[(size_t)(keep.row0 + row) * kMlpHidden + col] = h;
                    keep.a2[(size_t)(keep.row0 + row) * kMlpHidden + col] = a;
                }
        }
    wave_lds_sync();
    // layer 3: 128 -> 8 (one 32-column tile, columns 8..31 zero)
    f32x16 o;
#pragma unroll
    for(int i = 0; i < 16; ++i) o[i] = 0.0f;
#pragma unroll
    for(int ks = 0; ks < kMlpHidden / 16; ++ks)
        o = mfma_32x32x16(*(const uint4*)&xa[r][ks * 16 + kh], *(const uint4*)&W.w3t[r][ks * 16 + kh], o);
#pragma unroll
    for(int i = 0; i < 16; ++i) out[i] = o[i] + W.b3[r];
    wave_lds_sync();
}

// Shared memory of the training forward pass: the weights it reads (the
// part of MlpDev up to b3), then two 32-row activation tiles per wave.
constexpr int kMlpWaves = 4;
constexpr size_t kMlpTileBytes = 32 * kLdH * sizeof(uint16_t);
constexpr size_t kMlpLdsBytes  = kMlpLdsWeights + 2 * kMlpWaves * kMlpTileBytes;

__device__ __forceinline__ void load_weights_lds(const MlpDev* W, char* lds)
{
    const uint4* src = (const uint4*)W;
    uint4* dst       = (uint4*)lds;
    for(uint32_t i = threadIdx.x; i < kMlpLdsWeights / 16; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
}

__device__ __forceinline__ void wave_tiles(char* lds, uint16_t (**xa)[kLdH], uint16_t (**xb)[kLdH])
{
    const int w = threadIdx.x >> 6;
    *xa = (uint16_t (*)[kLdH])(lds + kMlpLdsWeights + (2 * w) * kMlpTileBytes);
    *xb = (uint16_t (*)[kLdH])(lds + kMlpLdsWeights + (2 * w + 1) * kMlpTileBytes);
}

// Raw network outputs of n given feature rows (fp32 [n][80]) -> out [n][8],
// 16 rows per wave through the same inference code as the planner's pass.
__global__ __launch_bounds__(256) void k_mlp_infer(const MlpDev* W, const float* X, uint32_t n, float* out)
{
    const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    const uint32_t ntiles = (n + 15) / 16;
    for(uint32_t tile = blockIdx.x * kMlpWaves + (threadIdx.x >> 6); tile < ntiles; tile += gridDim.x * kMlpWaves)
    {
        const uint32_t row = tile * 16 + j, rr = row < n ? row : n - 1;
        uint4 xb[kFragS1];
#pragma unroll
        for(int s = 0; s < kFragS1; ++s)
        {
            uint16_t v[8];
#pragma unroll
            for(int e = 0; e < 8; ++e)
            {
                const int k = 32 * s + 8 * g + e;
                v[e] = k < kMlpIn ? f2bf(X[(size_t)rr * kMlpIn + k]) : (uint16_t)0;
            }
            xb[s] = make_uint4(pack_bf16x2(v[0], v[1]), pack_bf16x2(v[2], v[3]), pack_bf16x2(v[4], v[5]),
                               pack_bf16x2(v[6], v[7]));
        }
        const f32x4 o = mlp_forward16(W, xb);
        if(g < 2 && row < n)
#pragma unroll
            for(int e = 0; e < 4; ++e) out[(size_t)row * kMlpOut + 4 * g + e] = o[e];
    }
}

// Features of n states (SoA, S x n) with one goal per state (3 x n).
__global__ __launch_bounds__(256) void k_features(FeatParams F, const uint32_t* occ, const float* st,
                                                  const float* goals, uint32_t n, float* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    float x[12], g[3], f[kMlpIn];
    for(int d = 0; d < F.S; ++d) x[d] = st[(size_t)d * n + i];
    for(int d = F.S; d < 12; ++d) x[d] = 0.0f;
    for(int d = 0; d < 3; ++d) g[d] = goals[(size_t)d * n + i];
    features(F, occ, x, g, f);
    for(int k = 0; k < kMlpIn; ++k) out[(size_t)i * kMlpIn + k] = f[k];
}

// Controls from given (mean, sigma) and Philox words (parity operator).
__global__ __launch_bounds__(256) void k_propose_sample(ModelParams P, const float* ms, const uint32_t* words,
                                                        uint32_t n, float* u)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const uint4 w = make_uint4(words[4 * (size_t)i], words[4 * (size_t)i + 1], words[4 * (size_t)i + 2],
                               words[4 * (size_t)i + 3]);
    float c[4];
    propose_control<4>(ms + 8 * (size_t)i, ms + 8 * (size_t)i + 4, w, P, c);
    for(int j = 0; j < 4; ++j) u[4 * (size_t)i + j] = c[j];
}

// ---------------------------------------------------------------------------
// Proposal pass of the planner: features of the nodes appended this
// iteration (or of the roots of newly admitted slots), the network, and the
// node's Gaussian (means, then sigmas = exp(clamp(log sigma, -4, 1/2))) into
// the per-node proposal arrays.  One wave per 16-node tile; each block lays
// out the tiles of all running slots from the append kernel's counts.
// ---------------------------------------------------------------------------
struct ProposeArgs {
    const Ctl* ctl;
    const QueryDev* qd;
    const GridView* envs;
    const float* tree_st;
    float* prop;
    FeatParams F;
    uint32_t T, Q;
    const uint32_t* init_slots;
    uint32_t n_init;
};

// Gaussian of every node appended this iteration (or of every root being
// initialised, n_init > 0), 16 nodes per wave: each block first lays out the
// slots' 16-node tiles from the counts the append kernel left (ctl[q].Fn), so a
// block with no tile exits at once and the tile loop reads only LDS for it.
// The weights are read as operand fragments straight from global memory (L2).
__global__ __launch_bounds__(256) void k_propose(const MlpDev* W, ProposeArgs A)
{
    __shared__ uint32_t s_scan[4];
    __shared__ uint32_t s_off[kMaxSlots + 1], s_node[kMaxSlots], s_nv[kMaxSlots];
    __shared__ const uint32_t* s_occ[kMaxSlots];
    __shared__ float s_goal[kMaxSlots][3];
    const uint32_t n = A.n_init ? A.n_init : A.Q;  // table entries (1 <= n <= kMaxSlots = 4 x 256)
    uint32_t cnt[4], mine = 0;
#pragma unroll
    for(int j = 0; j < 4; ++j)
    {
        // every entry's loads go out together (indices clamped), then the selects
        const uint32_t k  = threadIdx.x * 4 + j;
        const uint32_t kc = k < n ? k : 0u;
        const uint32_t q  = A.n_init ? A.init_slots[kc] : kc;
        const QueryDev& d = A.qd[q];
        const float mix = d.mix;
        const uint32_t base = d.base, env = d.env;
        const float g0 = d.goal[0], g1 = d.goal[1], g2 = d.goal[2];
        uint32_t nv, node0;
        if(A.n_init)
        {
            nv    = mix > 0.0f ? 1u : 0u;  // queries under the budget policy sample uniformly
            node0 = base;
        }
        else
        {
            const Ctl& c = A.ctl[kc];
            const int32_t st = c.cur.status;
            const uint32_t fn = c.Fn, ts = c.cur.ts;
            nv    = st == kRunning && mix > 0.0f ? fn : 0u;
            node0 = base + ts;
        }
        if(k >= n) nv = 0;
        if(k < n)
        {
            s_node[k]    = node0;
            s_nv[k]      = nv;
            s_occ[k]     = nv ? A.envs[env].occ : nullptr;
            s_goal[k][0] = g0;
            s_goal[k][1] = g1;
            s_goal[k][2] = g2;
        }
        cnt[j] = (nv + 15) / 16;
        mine += cnt[j];
    }
    uint32_t ntiles;
    uint32_t run = block_excl_scan<4>(mine, &ntiles, s_scan);
#pragma unroll
    for(int j = 0; j < 4; ++j)
    {
        const uint32_t k = threadIdx.x * 4 + j;
        if(k < n) s_off[k] = run;
        run += cnt[j];
    }
    if(threadIdx.x == 0) s_off[n] = ntiles;
    __syncthreads();
    if(blockIdx.x * kMlpWaves >= ntiles) return;  // the same for every thread of the block
    const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    for(uint32_t tile = blockIdx.x * kMlpWaves + (threadIdx.x >> 6); tile < ntiles; tile += gridDim.x * kMlpWaves)
    {
        const uint32_t k     = segment_of(s_off, n, tile);
        const uint32_t lt    = tile - s_off[k];
        const uint32_t left  = s_nv[k] - lt * 16;
        const uint32_t nv    = left < 16 ? left : 16u;
        const uint32_t node0 = s_node[k] + lt * 16;
        const uint32_t node  = node0 + ((uint32_t)j < nv ? (uint32_t)j : nv - 1);  // rows past the last repeat it
        float x[12];  // the proposal is defined for the quadcopter: 12 state values
#pragma unroll
        for(int d = 0; d < 12; ++d) x[d] = A.tree_st[(size_t)d * A.T + node];
        const float goal[3] = {s_goal[k][0], s_goal[k][1], s_goal[k][2]};
        uint4 xb[kFragS1];
        features16(A.F, s_occ[k], x, goal, xb);
        const f32x4 o = mlp_forward16(W, xb);
        if(g < 2 && (uint32_t)j < nv)
#pragma unroll
            for(int e = 0; e < 4; ++e)
            {
                float v = o[e];
                if(g == 1) v = expf(v < -4.0f ? -4.0f : (v > 0.5f ? 0.5f : v));
                A.prop[(size_t)(4 * g + e) * A.T + node0 + j] = v;
            }
    }
}

// ---------------------------------------------------------------------------
// Training.  One wave per 32-row tile of the batch.
// ---------------------------------------------------------------------------
struct TrainBufs {
    const float* X;        // dataset features [n][80]
    const float* Y;        // dataset targets [n][4] (normalised controls)
    const uint32_t* idx;   // batch rows -> dataset rows
    uint16_t* xb;          // batch inputs bf16 [B][80]
    float* h1;             // [B][128]
    float* h2;
    uint16_t* a1;
    uint16_t* a2;
    float* out;            // [B][8]
    uint16_t* dout;        // bf16 [B][16] (columns 8..15 zero)
    uint16_t* dh2;         // bf16 [B][128]
    uint16_t* dh1;
    float* dh2f;           // fp32 [B][128] (bias gradients)
    float* dh1f;
    float* loss_part;      // per tile
};

// Forward over the batch, keeping activations; then the Gaussian NLL
// (mean over rows of sum_j 0.5 ((y - mu) / sigma)^2 + log sigma, log sigma
// floored at -4) and its gradient with respect to the 8 outputs.
__global__ __launch_bounds__(256) void k_train_forward(const MlpDev* Wd, TrainBufs T, uint32_t B)
{
    HIP_DYNAMIC_SHARED(char, lds)
    load_weights_lds(Wd, lds);
    const MlpDev& W = *(const MlpDev*)lds;
    uint16_t(*xa)[kLdH];
    uint16_t(*xb)[kLdH];
    wave_tiles(lds, &xa, &xb);
    const int lane = threadIdx.x & 63, r = lane & 31;
    const uint32_t ntiles = (B + 31) / 32;
    const float invB      = 1.0f / (float)B;
    for(uint32_t tile = blockIdx.x * kMlpWaves + (threadIdx.x >> 6); tile < ntiles; tile += gridDim.x * kMlpWaves)
    {
        const uint32_t row0 = tile * 32;
        for(int k = lane; k < 32 * kMlpIn; k += 64)
        {
            const int rr = k / kMlpIn, cc = k % kMlpIn;
            uint16_t v = 0;
            if(row0 + rr < B)
            {
                v = f2bf(T.X[(size_t)T.idx[row0 + rr] * kMlpIn + cc]);
                T.xb[(size_t)(row0 + rr) * kMlpIn + cc] = v;
            }
            xa[rr][cc] = v;
        }
        wave_lds_sync();
        float o[16];
        const Keep keep{T.h1, T.h2, T.a1, T.a2, row0, B - row0 < 32 ? B - row0 : 32u};
        mlp_tile(W, xa, xb, o, keep);
        // outputs to LDS (fp32 view of xb: 32 rows x 8) for the row-wise loss
        float* ot = (float*)&xb[0][0];
        if(r < kMlpOut)
            for(int i = 0; i < 16; ++i) ot[mfma_row(i, lane) * kMlpOut + r] = o[i];
        wave_lds_sync();
        float l = 0.0f;
        if(lane < 32 && row0 + lane < B)
        {
            const uint32_t row = row0 + lane;
            const float* y     = T.Y + (size_t)T.idx[row] * 4;
            for(int j = 0; j < 4; ++j)
            {
                const float mu = ot[lane * kMlpOut + j], s = ot[lane * kMlpOut + 4 + j];
                const float sc = s < -4.0f ? -4.0f : s;
                const float iv = expf(-2.0f * sc);  // 1 / sigma^2
                const float d  = y[j] - mu;
                l += 0.5f * d * d * iv + sc;
                const float gmu = -d * iv * invB;
                const float gs  = s < -4.0f ? 0.0f : (1.0f - d * d * iv) * invB;
                T.out[(size_t)row * kMlpOut + j]     = mu;
                T.out[(size_t)row * kMlpOut + 4 + j] = s;
                T.dout[(size_t)row * 16 + j]         = f2bf(gmu);
                T.dout[(size_t)row * 16 + 4 + j]     = f2bf(gs);
            }
            for(int j = 8; j < 16; ++j) T.dout[(size_t)row * 16 + j] = 0;
        }
        const float ls = wave_sum_f(l);
        if(lane == 0) T.loss_part[tile] = ls;
        wave_lds_sync();
    }
}

// Backward through the rows: dA2 = dOut W3^T, dH2 = dA2 gelu'(H2),
// dA1 = dH2 W2^T, dH1 = dA1 gelu'(H1).  One wave per 32-row tile; W2 and W3
// are read from HBM in their [in][out] layout.
__global__ __launch_bounds__(256) void k_train_backward(const MlpDev* W, TrainBufs T, uint32_t B)
{
    HIP_DYNAMIC_SHARED(char, lds)
    const int w = threadIdx.x >> 6;
    uint16_t(*ta)[kLdH] = (uint16_t (*)[kLdH])(lds + (2 * w) * kMlpTileBytes);
    const int lane = threadIdx.x & 63, r = lane & 31, kh = (lane >> 5) * 8;
    const uint32_t ntiles = (B + 31) / 32;
    for(uint32_t tile = blockIdx.x * kMlpWaves + w; tile < ntiles; tile += gridDim.x * kMlpWaves)
    {
        const uint32_t row0 = tile * 32;
        const uint32_t nr   = B - row0 < 32 ? B - row0 : 32u;
        // dA2 (32 x 128) = dOut (32 x 16) * W3^T (16 x 128): B[k][n] = W3[n][k]
        f32x16 acc[4];
        {
            const uint4 a = r < (int)nr ? *(const uint4*)&T.dout[(size_t)(row0 + r) * 16 + kh] : make_uint4(0, 0, 0, 0);
#pragma unroll
            for(int nt = 0; nt < 4; ++nt)
            {
#pragma unroll
                for(int i = 0; i < 16; ++i) acc[nt][i] = 0.0f;
                acc[nt] = mfma_32x32x16(a, *(const uint4*)&W->w3[nt * 32 + r][kh], acc[nt]);
            }
        }
#pragma unroll
        for(int nt = 0; nt < 4; ++nt)
#pragma unroll
            for(int i = 0; i < 16; ++i)
            {
                const int row = mfma_row(i, lane), col = nt * 32 + r;
                float g       = 0.0f;
                if(row < (int)nr)
                {
                    g = acc[nt][i] * gelu_grad(T.h2[(size_t)(row0 + row) * kMlpHidden + col]);
                    T.dh2f[(size_t)(row0 + row) * kMlpHidden + col] = g;
                    T.dh2[(size_t)(row0 + row) * kMlpHidden + col]  = f2bf(g);
                }
                ta[row][col] = f2bf(g);
            }
        wave_lds_sync();
        // dA1 (32 x 128) = dH2 (32 x 128) * W2^T: B[k][n] = W2[n][k]
#pragma unroll
        for(int nt = 0; nt < 4; ++nt)
#pragma unroll
            for(int i = 0; i < 16; ++i) acc[nt][i] = 0.0f;
#pragma unroll
        for(int ks = 0; ks < kMlpHidden / 16; ++ks)
        {
            const uint4 a = *(const uint4*)&ta[r][ks * 16 + kh];
#pragma unroll
            for(int nt = 0; nt < 4; ++nt)
                acc[nt] = mfma_32x32x16(a, *(const uint4*)&W->w2[nt * 32 + r][ks * 16 + kh], acc[nt]);
        }
#pragma unroll
        for(int nt = 0; nt < 4; ++nt)
#pragma unroll
            for(int i = 0; i < 16; ++i)
            {
                const int row = mfma_row(i, lane), col = nt * 32 + r;
                if(row < (int)nr)
                {
                    const float g = acc[nt][i] * gelu_grad(T.h1[(size_t)(row0 + row) * kMlpHidden + col]);
                    T.dh1f[(size_t)(row0 + row) * kMlpHidden + col] = g;
                    T.dh1[(size_t)(row0 + row) * kMlpHidden + col]  = f2bf(g);
                }
            }
        wave_lds_sync();
    }
}

// Weight gradients C[m][n] = sum_b P[b][m] Q[b][n] (P: bf16 [B][ldp],
// Q: bf16 [B][ldq]) over row chunks of 256: each wave owns one 32 x 32
// output tile of one chunk and writes its partial sum; k_reduce_partials adds
// the chunks in order, so the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_grad_tiles(const uint16_t* P, int ldp, int M, const uint16_t* Q, int ldq,
                                                    int N, uint32_t B, float* partial)
{
    const int lane = threadIdx.x & 63, r = lane & 31, kh = (lane >> 5) * 8;
    const int mt = (M + 31) / 32, ntl = (N + 31) / 32;
    const uint32_t chunks = (B + 255) / 256;
    const uint32_t tasks  = chunks * mt * ntl;
    for(uint32_t task = blockIdx.x * kMlpWaves + (threadIdx.x >> 6); task < tasks; task += gridDim.x * kMlpWaves)
    {
        const uint32_t ch = task / (mt * ntl), tt = task % (mt * ntl);
        const int m0 = (int)(tt / ntl) * 32, n0 = (int)(tt % ntl) * 32;
        f32x16 acc;
#pragma unroll
        for(int i = 0; i < 16; ++i) acc[i] = 0.0f;
        const uint32_t b0 = ch * 256, b1 = b0 + 256 < B ? b0 + 256 : B;
        for(uint32_t bk = b0; bk < b1; bk += 16)
        {
            uint16_t av[8], bv[8];
#pragma unroll
            for(int j = 0; j < 8; ++j)
            {
                const uint32_t b = bk + kh + j;
                const int m = m0 + r, n = n0 + r;
                av[j] = (b < b1 && m < M) ? P[(size_t)b * ldp + m] : (uint16_t)0;
                bv[j] = (b < b1 && n < N) ? Q[(size_t)b * ldq + n] : (uint16_t)0;
            }
            uint4 a, bb;
            memcpy(&a, av, 16);
            memcpy(&bb, bv, 16);
            acc = mfma_32x32x16(a, bb, acc);
        }
        float* out = partial + ((size_t)ch * M * N);
#pragma unroll
        for(int i = 0; i < 16; ++i)
        {
            const int m = m0 + mfma_row(i, lane), n = n0 + r;
            if(m < M && n < N) out[(size_t)m * N + n] = acc[i];
        }
    }
}

__global__ __launch_bounds__(256) void k_reduce_partials(const float* partial, uint32_t chunks, uint32_t count,
                                                         float* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= count) return;
    float s = 0.0f;
    for(uint32_t c = 0; c < chunks; ++c) s += partial[(size_t)c * count + i];
    out[i] = s;
}

// Bias gradients: column sums over the batch in row order.
__global__ __launch_bounds__(256) void k_bias_grads(const float* dh1, const float* dh2, const float* out_grad_src,
                                                    uint32_t B, float* g)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if(c < kMlpHidden)
    {
        float s1 = 0.0f, s2 = 0.0f;
        for(uint32_t b = 0; b < B; ++b)
        {
            s1 += dh1[(size_t)b * kMlpHidden + c];
            s2 += dh2[(size_t)b * kMlpHidden + c];
        }
        g[kMlpB1 + c] = s1;
        g[kMlpB2 + c] = s2;
    }
    else if(c < kMlpHidden + kMlpOut)
    {
        const int j = c - kMlpHidden;
        float s     = 0.0f;
        for(uint32_t b = 0; b < B; ++b) s += out_grad_src[(size_t)b * kMlpOut + j];
        g[kMlpB3 + j] = s;
    }
}

// fp32 dOut (for the b3 gradient) from the stored bf16-free recomputation:
// the loss kernel keeps mu and s; this recomputes the exact fp32 gradient.
__global__ __launch_bounds__(256) void k_dout_f32(const float* out, const float* Y, const uint32_t* idx, uint32_t B,
                                                  float* g)
{
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if(row >= B) return;
    const float invB = 1.0f / (float)B;
    const float* y   = Y + (size_t)idx[row] * 4;
    for(int j = 0; j < 4; ++j)
    {
        const float mu = out[(size_t)row * kMlpOut + j], s = out[(size_t)row * kMlpOut + 4 + j];
        const float sc = s < -4.0f ? -4.0f : s;
        const float iv = expf(-2.0f * sc);
        const float d  = y[j] - mu;
        g[(size_t)row * kMlpOut + j]     = -d * iv * invB;
        g[(size_t)row * kMlpOut + 4 + j] = s < -4.0f ? 0.0f : (1.0f - d * d * iv) * invB;
    }
}

// Adam (Kingma & Ba) on the fp32 parameters.
__global__ __launch_bounds__(256) void k_adam(float* w, const float* g, float* m, float* v, uint32_t n, float lr,
                                              float b1, float b2, float c1, float c2, float eps)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i]           = mi;
    v[i]           = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace

// ===========================================================================
// Host side
// ===========================================================================
namespace kpax {

FeatParams make_feat_params(const kpax_config& c)
{
    FeatParams F;
    for(int d = 0; d < 3; ++d)
    {
        F.mid[d]      = 0.5f * (c.ws_min[d] + c.ws_max[d]);
        F.half[d]     = 0.5f * (c.ws_max[d] - c.ws_min[d]);
        F.occ_mn[d]   = c.ws_min[d];
        F.occ_cell[d] = (c.ws_max[d] - c.ws_min[d]) / (float)kOccDim;
    }
    F.inv_pi     = 0.31830988618379067f;
    const float v = std::max(std::fabs(c.vel_min), std::fabs(c.vel_max));
    F.inv_vel    = v > 0.0f ? 1.0f / v : 1.0f;
    F.rate_scale = 0.25f;
    F.S          = state_dim(c.model);
    return F;
}

// Occupancy bits of environment `env` (after its obstacles are on the device).
int build_occupancy(kpax_handle* h, uint32_t env)
{
    Env& e = h->envs[env];
    if(!e.occ) HIPC(hipMalloc(&e.occ, kOccDim * kOccDim * kOccDim / 8));
    const FeatParams F = make_feat_params(h->cfg);
    HIPC(hipMemsetAsync(e.occ, 0, kOccDim * kOccDim * kOccDim / 8, h->stream));
    if(e.n_obs)
    {
        const uint32_t chunks = std::min<uint32_t>((e.n_obs + kOccChunk - 1) / kOccChunk, 1024);
        hipLaunchKernelGGL(k_occupancy, dim3(kOccDim * kOccDim * kOccDim / 256, chunks), dim3(256), 0, h->stream, F,
                           e.lo, e.hi, e.n_obs, e.occ);
    }
    HIPC(hipGetLastError());
    return KPAX_OK;
}

int launch_propose(kpax_handle* h, uint32_t n_init)
{
    ProposeArgs A;
    A.ctl        = h->ctl;
    A.qd         = h->qd;
    A.envs       = h->envs_dev;
    A.tree_st    = h->tree.st;
    A.prop       = h->prop;
    A.F          = make_feat_params(h->cfg);
    A.T          = h->T;
    A.Q          = h->Qmax;
    A.init_slots = h->d_init_slots;
    A.n_init     = n_init;
    const uint32_t blocks = n_init ? std::max<uint32_t>(1, (n_init + kMlpWaves - 1) / kMlpWaves)
                                   : 2 * (uint32_t)h->cu_count;
    hipLaunchKernelGGL(k_propose, dim3(blocks), dim3(256), 0, h->stream, (const MlpDev*)h->mlp, A);
    HIPC(hipGetLastError());
    return KPAX_OK;
}

}  // namespace kpax

struct kpax_trainer {
    int device         = 0;
    hipStream_t stream = nullptr;
    int cu_count       = 256;
    uint32_t maxB      = 0;
    float* w           = nullptr;  // fp32 parameters
    float* m           = nullptr;  // Adam moments
    float* v           = nullptr;
    float* g           = nullptr;  // gradients
    MlpDev* wd         = nullptr;  // bf16 layouts
    float* X           = nullptr;  // dataset
    float* Y           = nullptr;
    uint32_t n_data = 0, cap_data = 0;
    uint32_t* idx      = nullptr;
    TrainBufs bufs{};
    float* partial     = nullptr;
    float* doutf       = nullptr;
    float* loss_part   = nullptr;
    uint64_t steps     = 0;
    std::string err;
    std::vector<void*> allocs;
};

namespace {

int tfail(kpax_trainer* t, int code, const std::string& msg)
{
    if(t) t->err = msg;
    return code;
}

#define TCHK(expr)                                                                                   \
    do                                                                                               \
    {                                                                                                \
        hipError_t e_ = (expr);                                                                      \
        if(e_ != hipSuccess)                                                                         \
            return tfail(t, KPAX_DEVICE_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while(0)

template<class T>
int talloc(kpax_trainer* t, T** p, size_t count)
{
    void* q = nullptr;
    if(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
        return tfail(t, KPAX_NO_MEMORY, "hipMalloc failed");
    t->allocs.push_back(q);
    *p = (T*)q;
    return KPAX_OK;
}

int pack(kpax_trainer* t)
{
    hipLaunchKernelGGL(k_pack_weights, dim3(64), dim3(256), 0, t->stream, t->w, t->wd);
    TCHK(hipGetLastError());
    return KPAX_OK;
}

int ensure_data(kpax_trainer* t, uint32_t n)
{
    if(n <= t->cap_data) return KPAX_OK;
    if(t->X) (void)hipFree(t->X);
    if(t->Y) (void)hipFree(t->Y);
    t->X = t->Y = nullptr;
    t->cap_data = 0;
    TCHK(hipMalloc(&t->X, (size_t)n * kMlpIn * sizeof(float)));
    TCHK(hipMalloc(&t->Y, (size_t)n * 4 * sizeof(float)));
    t->cap_data = n;
    return KPAX_OK;
}

// Gradients of the batch idx[0..B) into t->g; loss parts into loss_part.
int batch_gradients(kpax_trainer* t, uint32_t B, float* loss_part)
{
    TrainBufs T  = t->bufs;
    T.X          = t->X;
    T.Y          = t->Y;
    T.idx        = t->idx;
    T.loss_part  = loss_part;
    const uint32_t tiles = (B + 31) / 32;
    const uint32_t blocks = std::min<uint32_t>((tiles + kMlpWaves - 1) / kMlpWaves, 2 * (uint32_t)t->cu_count);
    hipLaunchKernelGGL(k_train_forward, dim3(blocks), dim3(256), kMlpLdsBytes, t->stream, t->wd, T, B);
    hipLaunchKernelGGL(k_train_backward, dim3(blocks), dim3(256), 2 * kMlpWaves * kMlpTileBytes, t->stream, t->wd, T,
                       B);
    hipLaunchKernelGGL(k_dout_f32, dim3(blocks_for(B, 256)), dim3(256), 0, t->stream, T.out, T.Y, T.idx, B, t->doutf);
    const uint32_t chunks = (B + 255) / 256;
    auto wgrad = [&](const uint16_t* P, int ldp, int M, const uint16_t* Q, int ldq, int N, float* dst) {
        const uint32_t tasks = chunks * ((M + 31) / 32) * ((N + 31) / 32);
        hipLaunchKernelGGL(k_grad_tiles, dim3(std::max<uint32_t>(1, (tasks + kMlpWaves - 1) / kMlpWaves)), dim3(256),
                           0, t->stream, P, ldp, M, Q, ldq, N, B, t->partial);
        hipLaunchKernelGGL(k_reduce_partials, dim3(blocks_for((uint64_t)M * N, 256)), dim3(256), 0, t->stream,
                           t->partial, chunks, (uint32_t)(M * N), dst);
    };
    wgrad(T.xb, kMlpIn, kMlpIn, T.dh1, kMlpHidden, kMlpHidden, t->g + kMlpW1);
    wgrad(T.a1, kMlpHidden, kMlpHidden, T.dh2, kMlpHidden, kMlpHidden, t->g + kMlpW2);
    wgrad(T.a2, kMlpHidden, kMlpHidden, T.dout, 16, kMlpOut, t->g + kMlpW3);
    hipLaunchKernelGGL(k_bias_grads, dim3(1), dim3(256), 0, t->stream, T.dh1f, T.dh2f, t->doutf, B, t->g);
    TCHK(hipGetLastError());
    return KPAX_OK;
}

}  // namespace

extern "C" {

int kpax_set_proposal_min_budget(kpax_handle* h, uint32_t min_budget)
{
    if(!h) return KPAX_INVALID_ARG;
    h->prop_min_budget = min_budget;
    return KPAX_OK;
}

int kpax_set_proposal(kpax_handle* h, const float* params, float mix)
{
    if(!h) return KPAX_INVALID_ARG;
    if(!(mix >= 0.0f && mix <= 1.0f)) return fail(h, KPAX_INVALID_ARG, "mix must lie in [0, 1]");
    HIPC(hipSetDevice(h->cfg.device));
    HIPC(hipStreamSynchronize(h->stream));
    if(!params || mix == 0.0f)
        h->mix = 0.0f;
    else
    {
        if(h->C != 4) return fail(h, KPAX_INVALID_ARG, "the learned proposal is defined for the quadcopter");
        for(int i = 0; i < kMlpParams; ++i)
            if(!std::isfinite(params[i])) return fail(h, KPAX_INVALID_ARG, "non-finite network parameter");
        if(!h->mlp)
        {
            MlpDev* d = nullptr;
            int rc    = dalloc(h, &d, 1);
            if(rc != KPAX_OK) return rc;
            h->mlp = d;
            rc     = dalloc(h, &h->prop, (size_t)8 * h->T);
            if(rc != KPAX_OK) return rc;
        }
        Scratch sc;
        float* w = sc.get<float>(kMlpParams);
        if(!w) return fail(h, KPAX_NO_MEMORY, "scratch allocation failed");
        HIPC(hipMemcpy(w, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_pack_weights, dim3(64), dim3(256), 0, h->stream, w, (MlpDev*)h->mlp);
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(h->stream));
        h->mix = mix;
    }
    return refresh_launch_state(h);
}

int kpax_mlp_features(kpax_handle* h, uint32_t env, const float* states_soa, const float* goals_soa, uint32_t n,
                      float* out)
{
    if(!h || (n && (!states_soa || !goals_soa || !out))) return fail(h, KPAX_INVALID_ARG, "null argument");
    if(env >= (uint32_t)kMaxEnvs || !h->envs[env].occ) return fail(h, KPAX_INVALID_ARG, "unknown environment");
    if(n == 0) return KPAX_OK;
    HIPC(hipSetDevice(h->cfg.device));
    Scratch sc;
    float* st = sc.get<float>((size_t)h->S * n);
    float* g  = sc.get<float>((size_t)3 * n);
    float* f  = sc.get<float>((size_t)kMlpIn * n);
    if(!st || !g || !f) return fail(h, KPAX_NO_MEMORY, "scratch allocation failed");
    HIPC(hipMemcpy(st, states_soa, (size_t)h->S * n * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(g, goals_soa, (size_t)3 * n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_features, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, make_feat_params(h->cfg),
                       (const uint32_t*)h->envs[env].occ, st, g, n, f);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    HIPC(hipMemcpy(out, f, (size_t)kMlpIn * n * 4, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_mlp_infer(kpax_handle* h, const float* params, const float* features, uint32_t n, float* out)
{
    if(!h || !params || (n && (!features || !out))) return fail(h, KPAX_INVALID_ARG, "null argument");
    if(n == 0) return KPAX_OK;
    HIPC(hipSetDevice(h->cfg.device));
    Scratch sc;
    float* w   = sc.get<float>(kMlpParams);
    MlpDev* d  = sc.get<MlpDev>(1);
    float* x   = sc.get<float>((size_t)kMlpIn * n);
    float* o   = sc.get<float>((size_t)kMlpOut * n);
    if(!w || !d || !x || !o) return fail(h, KPAX_NO_MEMORY, "scratch allocation failed");
    HIPC(hipMemcpy(w, params, kMlpParams * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(x, features, (size_t)kMlpIn * n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_pack_weights, dim3(64), dim3(256), 0, h->stream, w, d);
    const uint32_t tiles = (n + 15) / 16;
    hipLaunchKernelGGL(k_mlp_infer, dim3(std::min<uint32_t>((tiles + kMlpWaves - 1) / kMlpWaves, 2 * h->cu_count)),
                       dim3(256), 0, h->stream, d, x, n, o);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    HIPC(hipMemcpy(out, o, (size_t)kMlpOut * n * 4, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_propose_sample(kpax_handle* h, const float* mean_sigma, const uint32_t* words, uint32_t n, float* controls)
{
    if(!h || (n && (!mean_sigma || !words || !controls))) return fail(h, KPAX_INVALID_ARG, "null argument");
    if(h->C != 4) return fail(h, KPAX_INVALID_ARG, "the learned proposal is defined for the quadcopter");
    if(n == 0) return KPAX_OK;
    HIPC(hipSetDevice(h->cfg.device));
    Scratch sc;
    float* ms   = sc.get<float>((size_t)8 * n);
    uint32_t* w = sc.get<uint32_t>((size_t)4 * n);
    float* u    = sc.get<float>((size_t)4 * n);
    if(!ms || !w || !u) return fail(h, KPAX_NO_MEMORY, "scratch allocation failed");
    HIPC(hipMemcpy(ms, mean_sigma, (size_t)8 * n * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(w, words, (size_t)4 * n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_propose_sample, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->P, ms, w, n, u);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    HIPC(hipMemcpy(controls, u, (size_t)4 * n * 4, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_mlp_init_params(uint64_t seed, float* params)
{
    if(!params) return KPAX_INVALID_ARG;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    std::fill(params, params + kMlpParams, 0.0f);
    const float a1 = std::sqrt(6.0f / kMlpFeatures), a2 = std::sqrt(6.0f / kMlpHidden),
                a3 = 0.1f * std::sqrt(6.0f / kMlpHidden);
    for(int k = 0; k < kMlpFeatures; ++k)
        for(int n = 0; n < kMlpHidden; ++n) params[kMlpW1 + k * kMlpHidden + n] = a1 * U(rng);
    for(int i = 0; i < kMlpHidden * kMlpHidden; ++i) params[kMlpW2 + i] = a2 * U(rng);
    for(int i = 0; i < kMlpHidden * kMlpOut; ++i) params[kMlpW3 + i] = a3 * U(rng);
    for(int j = 4; j < kMlpOut; ++j) params[kMlpB3 + j] = -0.6931472f;  // sigma = 1/2 at the start
    return KPAX_OK;
}

int kpax_trainer_create(int32_t device, uint32_t max_batch, kpax_trainer** out)
{
    if(!out || max_batch < 1 || max_batch > (1u << 20)) return KPAX_INVALID_ARG;
    *out = nullptr;
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return KPAX_NO_DEVICE;
    kpax_trainer* t = new(std::nothrow) kpax_trainer();
    if(!t) return KPAX_NO_MEMORY;
    t->device = device;
    t->maxB   = max_batch;
    int rc    = KPAX_OK;
#define TTRY(x)                      \
    do                               \
    {                                \
        rc = (x);                    \
        if(rc != KPAX_OK)            \
        {                            \
            kpax_trainer_destroy(t); \
            return rc;               \
        }                            \
    } while(0)
    if(hipSetDevice(device) != hipSuccess) TTRY(KPAX_DEVICE_ERROR);
    hipDeviceProp_t prop;
    if(hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        t->cu_count = prop.multiProcessorCount;
    if(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess) TTRY(KPAX_DEVICE_ERROR);
    const size_t B = max_batch;
    TTRY(talloc(t, &t->w, kMlpParams));
    TTRY(talloc(t, &t->m, kMlpParams));
    TTRY(talloc(t, &t->v, kMlpParams));
    TTRY(talloc(t, &t->g, kMlpParams));
    TTRY(talloc(t, &t->wd, 1));
    TTRY(talloc(t, &t->idx, B));
    TTRY(talloc(t, &t->bufs.xb, B * kMlpIn));
    TTRY(talloc(t, &t->bufs.h1, B * kMlpHidden));
    TTRY(talloc(t, &t->bufs.h2, B * kMlpHidden));
    TTRY(talloc(t, &t->bufs.a1, B * kMlpHidden));
    TTRY(talloc(t, &t->bufs.a2, B * kMlpHidden));
    TTRY(talloc(t, &t->bufs.out, B * kMlpOut));
    TTRY(talloc(t, &t->bufs.dout, B * 16));
    TTRY(talloc(t, &t->bufs.dh2, B * kMlpHidden));
    TTRY(talloc(t, &t->bufs.dh1, B * kMlpHidden));
    TTRY(talloc(t, &t->bufs.dh2f, B * kMlpHidden));
    TTRY(talloc(t, &t->bufs.dh1f, B * kMlpHidden));
    TTRY(talloc(t, &t->doutf, B * kMlpOut));
    TTRY(talloc(t, &t->partial, ((B + 255) / 256) * (size_t)kMlpHidden * kMlpHidden));
    TTRY(talloc(t, &t->loss_part, (B + 31) / 32 * 4096));
    std::vector<float> p(kMlpParams);
    kpax_mlp_init_params(1, p.data());
    TTRY(kpax_trainer_set_params(t, p.data()));
#undef TTRY
    *out = t;
    return KPAX_OK;
}

int kpax_trainer_destroy(kpax_trainer* t)
{
    if(!t) return KPAX_INVALID_ARG;
    if(t->stream) (void)hipStreamSynchronize(t->stream);
    for(void* p : t->allocs) (void)hipFree(p);
    if(t->X) (void)hipFree(t->X);
    if(t->Y) (void)hipFree(t->Y);
    if(t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
    return KPAX_OK;
}

const char* kpax_trainer_last_error(const kpax_trainer* t) { return t ? t->err.c_str() : "null trainer"; }

int kpax_trainer_set_params(kpax_trainer* t, const float* params)
{
    if(!t || !params) return KPAX_INVALID_ARG;
    TCHK(hipSetDevice(t->device));
    TCHK(hipMemcpy(t->w, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    TCHK(hipMemset(t->m, 0, kMlpParams * sizeof(float)));
    TCHK(hipMemset(t->v, 0, kMlpParams * sizeof(float)));
    t->steps = 0;
    const int rc = pack(t);
    if(rc != KPAX_OK) return rc;
    TCHK(hipStreamSynchronize(t->stream));
    return KPAX_OK;
}

int kpax_trainer_get_params(kpax_trainer* t, float* params)
{
    if(!t || !params) return KPAX_INVALID_ARG;
    TCHK(hipSetDevice(t->device));
    TCHK(hipStreamSynchronize(t->stream));
    TCHK(hipMemcpy(params, t->w, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_trainer_gradients(kpax_trainer* t, const float* X, const float* Y, uint32_t n, float* grads, float* loss)
{
    if(!t || !X || !Y || !grads || n == 0) return KPAX_INVALID_ARG;
    if(n > t->maxB) return tfail(t, KPAX_INVALID_ARG, "batch above the trainer's maximum");
    TCHK(hipSetDevice(t->device));
    int rc = ensure_data(t, n);
    if(rc != KPAX_OK) return rc;
    TCHK(hipMemcpy(t->X, X, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice));
    TCHK(hipMemcpy(t->Y, Y, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice));
    std::vector<uint32_t> id(n);
    for(uint32_t i = 0; i < n; ++i) id[i] = i;
    TCHK(hipMemcpy(t->idx, id.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    if((rc = batch_gradients(t, n, t->loss_part)) != KPAX_OK) return rc;
    TCHK(hipStreamSynchronize(t->stream));
    TCHK(hipMemcpy(grads, t->g, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    if(loss)
    {
        const uint32_t tiles = (n + 31) / 32;
        std::vector<float> lp(tiles);
        TCHK(hipMemcpy(lp.data(), t->loss_part, tiles * sizeof(float), hipMemcpyDeviceToHost));
        double s = 0.0;
        for(float x : lp) s += x;
        *loss = (float)(s / n);
    }
    return KPAX_OK;
}

int kpax_trainer_fit(kpax_trainer* t, const float* X, const float* Y, uint32_t n, uint32_t epochs, uint32_t batch,
                     float lr, uint64_t seed, float* epoch_loss)
{
    if(!t || !X || !Y || n == 0 || batch == 0 || epochs == 0) return KPAX_INVALID_ARG;
    if(batch > t->maxB) return tfail(t, KPAX_INVALID_ARG, "batch above the trainer's maximum");
    TCHK(hipSetDevice(t->device));
    int rc = ensure_data(t, n);
    if(rc != KPAX_OK) return rc;
    TCHK(hipMemcpy(t->X, X, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice));
    TCHK(hipMemcpy(t->Y, Y, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice));
    t->n_data = n;
    std::mt19937_64 rng(seed);
    std::vector<uint32_t> perm(n);
    for(uint32_t i = 0; i < n; ++i) perm[i] = i;
    const uint32_t nb = n / batch;  // full batches only
    if(nb == 0) return tfail(t, KPAX_INVALID_ARG, "fewer rows than one batch");
    const uint32_t tiles = (batch + 31) / 32;
    if((size_t)nb * tiles > (size_t)(t->maxB + 31) / 32 * 4096)
        return tfail(t, KPAX_INVALID_ARG, "too many batches per epoch");
    uint32_t* dperm = nullptr;
    TCHK(hipMalloc(&dperm, (size_t)n * sizeof(uint32_t)));
    std::vector<float> lp((size_t)nb * tiles);
    for(uint32_t ep = 0; ep < epochs; ++ep)
    {
        std::shuffle(perm.begin(), perm.end(), rng);
        if(hipMemcpyAsync(dperm, perm.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, t->stream) !=
           hipSuccess)
        {
            (void)hipFree(dperm);
            return tfail(t, KPAX_DEVICE_ERROR, "permutation upload failed");
        }
        for(uint32_t b = 0; b < nb; ++b)
        {
            t->idx = dperm + (size_t)b * batch;
            rc     = batch_gradients(t, batch, t->loss_part + (size_t)b * tiles);
            if(rc != KPAX_OK) break;
            ++t->steps;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.0f - std::pow(b1, (float)t->steps), c2 = 1.0f - std::pow(b2, (float)t->steps);
            hipLaunchKernelGGL(k_adam, dim3(blocks_for(kMlpParams, 256)), dim3(256), 0, t->stream, t->w, t->g, t->m,
                               t->v, (uint32_t)kMlpParams, lr, b1, b2, c1, c2, 1e-8f);
            if((rc = pack(t)) != KPAX_OK) break;
        }
        if(rc == KPAX_OK && epoch_loss)
        {
            if(hipMemcpyAsync(lp.data(), t->loss_part, lp.size() * sizeof(float), hipMemcpyDeviceToHost, t->stream) !=
                 hipSuccess ||
               hipStreamSynchronize(t->stream) != hipSuccess)
                rc = tfail(t, KPAX_DEVICE_ERROR, "loss read-back failed");
            double s = 0.0;
            for(float x : lp) s += x;
            epoch_loss[ep] = (float)(s / ((double)nb * batch));
        }
        if(rc != KPAX_OK) break;
    }
    (void)hipStreamSynchronize(t->stream);
    (void)hipFree(dperm);
    return rc;
}

}  // extern "C"
