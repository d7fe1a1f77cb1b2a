// kpax_device.hip — MI355X kernels of the Kino-PAX planner and the GPU half
// of the C ABI (include/kpax.h).
//
// One planner iteration is five launches on the handle's stream, with no host
// round trip between them (the reference makes at least eight blocking
// device->host transfers per iteration, KPAX.cu:303-543):
//   k_advance   1 block   publish the iteration record, clear scan state
//   k_expand    hot       one lane per child: Philox controls, RK4 with checks,
//                         region counters, admission (propagateFrontier_kernel1/2)
//   k_score     1 block   region scores and their normalisation (Graph::updateVertices)
//   k_append    scan      admitted children -> tree in slot order, goal test
//                         (updateFrontier_kernel, first half)
//   k_frontier  scan      re-admission + repeat counts -> next repeat list
//                         (updateFrontier_kernel second half + findInd/repeatInd)
// Tree and children are SoA in HBM: column d of node i at d * N + i.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "kpax_device.cuh"
#include "kpax_internal.h"

using namespace kpax;
using namespace kpax::dev;

namespace {

struct TreeView {
    float* st;
    float* ct;
    int32_t* steps;
    int32_t* parent;
    int32_t* reg;
};

struct ChildView {
    float* st;
    float* ct;
    int32_t* steps;
    int32_t* parent;
    int32_t* reg;
    int32_t* r2;
    uint8_t* adm;  // kValid | kCandidate | kScoreAdmit
};

struct RegionView {
    int32_t* valid;
    int32_t* total;
    int32_t* covered;
    float* score;
    uint64_t* r2_state;  // per R2 cell: first iteration << 32 | smallest slot, ~0 = never reached
};

constexpr int kTile       = 1024;  // items per scan tile (256 threads x 4)
constexpr int kLdsRegions = 4096;  // R1 counts up to this size are summed in LDS

// ---------------------------------------------------------------------------
// Fills (used instead of hipMemsetAsync so every per-plan launch can live in
// a captured graph).
// ---------------------------------------------------------------------------
__global__ void k_fill_i32(int32_t* p, int32_t v, uint64_t n)
{
    for(uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t v, uint64_t n)
{
    for(uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_f32(float* p, float v, uint64_t n)
{
    for(uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Root and first iteration record (KPAX.cu:60-88).
__global__ __launch_bounds__(256) void k_init(Ctl* ctl, TreeView tree, uint32_t* R, const float* start, int S, int C,
                                              RegionParams RP, PlanParams PP)
{
    const uint32_t N = PP.N;
    const int tid    = threadIdx.x;
    if(tid < S) tree.st[(size_t)tid * N] = start[tid];
    if(tid < C) tree.ct[(size_t)tid * N] = 0.0f;
    for(uint32_t i = tid; i < PP.root_repeat && i < N; i += blockDim.x) R[i] = 0u;
    if(tid == 0)
    {
        tree.steps[0]  = 0;
        tree.parent[0] = -1;
        int r1, r2;
        locate(RP, start, &r1, &r2);
        tree.reg[0] = r1;
        Ctl c;
        memset(&c, 0, sizeof(c));
        c.nxt        = make_iter(1u, 1u, PP.root_repeat, PP);
        c.cur        = c.nxt;
        c.cur.status = kNotStarted;  // nothing runs until k_advance publishes nxt
        c.goal       = INT_MAX;
        *ctl         = c;
    }
}

__global__ __launch_bounds__(256) void k_advance(Ctl* ctl, uint64_t* st_append, uint64_t* st_frontier, PlanParams PP)
{
    __shared__ uint32_t s_clear;
    if(threadIdx.x == 0)
    {
        const Iter n = ctl->nxt;
        ctl->cur     = n;
        if(n.status == kRunning)
        {
            ctl->expansions += n.M;
            ctl->Fn       = 0;
            ctl->f_accept = f_accept(n.t, n.ts, PP);
        }
        ctl->tile_expand   = 0;
        ctl->tile_append   = 0;
        ctl->tile_frontier = 0;
        // tiles the scans of this iteration can touch
        const uint64_t items = (uint64_t)n.ts + n.M;
        s_clear              = n.status == kRunning ? (uint32_t)((items + kTile - 1) / kTile) : 0u;
    }
    __syncthreads();
    for(uint32_t i = threadIdx.x; i < s_clear; i += blockDim.x)
    {
        st_append[i]   = 0;
        st_frontier[i] = 0;
    }
}

// ---------------------------------------------------------------------------
// Expansion: propagateFrontier_kernel1/2 (KPAX.cu:358-439) with counter-based
// draws.  Waves are persistent and advance their 64 children one integration
// step per trip through the loop; a lane whose child has finished (all its
// steps, or an invalid step) writes the child out and takes the next one
// from the wave's 64-child chunk, and the wave takes a new chunk from a
// global counter when its chunk runs out.  Children draw 5..20 steps and many
// stop early, so refilling keeps the lanes of a wave busy instead of waiting
// for the longest edge.  Each child's arithmetic is that of propagate().
// ---------------------------------------------------------------------------
// Everything k_expand reads that is fixed for a handle lives in one
// device-side block behind a single pointer: as by-value kernel arguments the
// structures exceed the scalar register budget and get spilled into vector
// lanes, costing a readlane at every use inside the step loop.
struct ExpandArgs {
    Ctl* ctl;
    TreeView tree;
    ChildView ch;
    const uint32_t* R;
    RegionView rv;
    ModelParams P;
    RegionParams RP;
    GridView G;
    PlanParams PP;
};

// Five resident waves per SIMD (96 VGPRs for the quadcopter) hide more of
// the grid lookups' latency than four.
template<int MODEL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_expand(
  const ExpandArgs* __restrict__ A, uint64_t seed, uint32_t query)
{
    Ctl* ctl              = A->ctl;
    const TreeView& tree  = A->tree;
    const ChildView& ch   = A->ch;
    const uint32_t* R     = A->R;
    const RegionView& rv  = A->rv;
    const ModelParams& P  = A->P;
    const RegionParams& RP = A->RP;
    const GridView& G     = A->G;
    const PlanParams& PP  = A->PP;
    constexpr int S = Dims<MODEL>::S, C = Dims<MODEL>::C;
    constexpr uint32_t kNone = 0xffffffffu;
    // Region counters are summed in LDS (total in [0, R1), valid in
    // [R1, 2 R1)) and added to HBM once per work-group: siblings land in the
    // same few regions, so per-child HBM atomics would serialise on them.
    HIP_DYNAMIC_SHARED(int32_t, s_hist)
    const Iter it = ctl->cur;
    if(it.status != kRunning) return;
    const uint32_t M = it.M, N = PP.N, t = it.t;
    const int R1      = RP.r1_count;
    const bool local  = R1 <= kLdsRegions;
    if(local)
    {
        for(int r = threadIdx.x; r < 2 * R1; r += blockDim.x) s_hist[r] = 0;
        __syncthreads();
    }
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const int lane      = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t c = kNone, parent = 0, admit_draw = 0;
    int i = 0, n = 0;
    float x[S], u[C];
    uint32_t q_next = 0, q_end = 0;  // the wave's current chunk [q_next, q_end)
    bool drained        = false;
    uint32_t steps_done = 0;
    for(;;)
    {
        if(!drained)
        {
            uint64_t need = __ballot(c == kNone);
            while(need)
            {
                if(q_next >= q_end)
                {
                    uint32_t base = 0;
                    if(lane == 0) base = atomicAdd(&ctl->tile_expand, 64u);
                    base = __shfl(base, 0, 64);
                    if(base >= M)
                    {
                        drained = true;
                        break;
                    }
                    q_next = base;
                    q_end  = base + 64u < M ? base + 64u : M;
                }
                const uint32_t avail = q_end - q_next;
                const uint32_t rank  = (uint32_t)__popcll(need & below);
                if(c == kNone && rank < avail)
                {
                    c      = q_next + rank;
                    parent = R[c / it.fan];
#pragma unroll
                    for(int d = 0; d < S; ++d) x[d] = tree.st[(size_t)d * N + parent];
                    const uint4 dc        = philox(c, t, query, PURPOSE_CONTROL, k0, k1);
                    const uint4 ds        = philox(c, t, query, PURPOSE_STEPS_ADMIT, k0, k1);
                    const uint32_t dcw[4] = {dc.x, dc.y, dc.z, dc.w};
#pragma unroll
                    for(int j = 0; j < C; ++j) u[j] = P.ctrl_min[j] + u01(dcw[j]) * P.ctrl_rng[j];
                    int kk = (int)(u01(ds.x) * (float)P.span);
                    if(kk > P.span - 1) kk = P.span - 1;
                    n          = P.steps_min + kk;
                    i          = 0;
                    admit_draw = ds.y;
                }
                const uint32_t wanted = (uint32_t)__popcll(need);
                q_next += wanted < avail ? wanted : avail;
                need = __ballot(c == kNone);
            }
        }
        if(__ballot(c != kNone) == 0) break;
        if(c != kNone)
        {
            const bool ok = step<MODEL>(P, G, x, u);
            ++i;
            if(!ok || i == n)
            {
                steps_done += (uint32_t)i;
                int r1, r2;
                locate(RP, x, &r1, &r2);
                if(local)
                {
                    atomicAdd(&s_hist[r1], 1);
                    if(ok) atomicAdd(&s_hist[R1 + r1], 1);
                }
                else
                {
                    atomicAdd(&rv.total[r1], 1);
                    if(ok) atomicAdd(&rv.valid[r1], 1);
                }
                uint32_t flags = 0;
                if(ok)
                {
                    flags = kValid | claim_cell(rv.r2_state, r2, t, c);
                    if(u01(admit_draw) < rv.score[r1]) flags |= kScoreAdmit;
                }
#pragma unroll
                for(int d = 0; d < S; ++d) ch.st[(size_t)d * N + c] = x[d];
#pragma unroll
                for(int d = 0; d < C; ++d) ch.ct[(size_t)d * N + c] = u[d];
                ch.steps[c]  = n;
                ch.parent[c] = (int32_t)parent;
                ch.reg[c]    = r1;
                ch.r2[c]     = r2;
                ch.adm[c]    = (uint8_t)flags;
                c            = kNone;
            }
        }
    }
    const uint32_t ws = wave_sum(steps_done);
    if(lane == 0 && ws) atomicAdd((unsigned long long*)&ctl->steps_done, (unsigned long long)ws);
    if(local)
    {
        __syncthreads();
        for(int r = threadIdx.x; r < R1; r += blockDim.x)
        {
            const int32_t a = s_hist[r], b = s_hist[R1 + r];
            if(a) atomicAdd(&rv.total[r], a);
            if(b) atomicAdd(&rv.valid[r], b);
        }
    }
}

// ---------------------------------------------------------------------------
// Region scores (Graph.cu:233-358, the evident intent): raw Syclop score per
// R1 region, their sum in a fixed order (1024 strided partial sums, then a
// pairwise tree), normalised eps + score / total, or 1 for regions with no
// valid child.  One block of 1024 threads.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_score(Ctl* ctl, RegionView rv, RegionParams RP, float eps,
                                                int32_t check_status)
{
    __shared__ float part[1024];
    if(check_status && ctl->cur.status != kRunning) return;
    const int t = threadIdx.x;
    float s     = 0.0f;
    for(int r = t; r < RP.r1_count; r += 1024)
    {
        float raw   = 0.0f;
        const int v = rv.valid[r];
        if(v > 0)
        {
            const int n0    = rv.total[r];
            const float cov = (float)rv.covered[r] / (float)RP.r2_per;
            const float ev  = eps + (float)v;
            const float fv  = ev / (ev + (float)(n0 - v)) * RP.vol;
            const float n   = (float)n0;
            raw             = (fv * fv) * (fv * fv) / ((1.0f + cov) * (1.0f + n * n));
        }
        rv.score[r] = raw;
        s += raw;
    }
    part[t] = s;
    __syncthreads();
    for(int off = 512; off >= 1; off >>= 1)
    {
        if(t < off) part[t] = part[t] + part[t + off];
        __syncthreads();
    }
    const float tot = part[0];
    for(int r = t; r < RP.r1_count; r += 1024) rv.score[r] = rv.valid[r] == 0 ? 1.0f : eps + rv.score[r] / tot;
    if(t == 0 && ctl) ctl->score_total = tot;
}

// ---------------------------------------------------------------------------
// Append: compaction of the admitted children into the tree in slot order
// and the goal test (updateFrontier_kernel, KPAX.cu:458-502).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_append(Ctl* ctl, ChildView ch, TreeView tree, const uint64_t* r2_state,
                                                int32_t* covered, uint64_t* status, PlanParams PP, int S, int C,
                                                float gx, float gy, float gz)
{
    __shared__ uint32_t lds[4];
    __shared__ uint32_t s_tile, s_pre;
    const Iter it = ctl->cur;
    if(it.status != kRunning) return;
    const uint32_t N = PP.N, M = it.M, ts = it.ts;
    const uint32_t ntiles = (M + kTile - 1) / kTile;
    for(;;)
    {
        const uint32_t tile = next_tile(&ctl->tile_append, &s_tile);
        if(tile >= ntiles) break;
        const uint32_t base = tile * kTile + threadIdx.x * 4;
        uint32_t f[4], cnt = 0;
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            const uint32_t c = base + j;
            const uint32_t a = c < M ? ch.adm[c] : 0u;
            // a candidate that holds its cell's key was the first to reach it:
            // it covers the cell and is admitted (KPAX.cu:394-395)
            bool first = false;
            if(a & kCandidate)
            {
                first = r2_state[ch.r2[c]] == cell_key(it.t, c);
                if(first) atomicAdd(&covered[ch.reg[c]], 1);
            }
            f[j] = (a & kValid) && ((a & kScoreAdmit) || first) ? 1u : 0u;
            cnt += f[j];
        }
        uint32_t agg;
        const uint32_t loc = block_excl_scan_256(cnt, &agg, lds);
        const uint32_t pre = tile_prefix(status, tile, agg, &s_pre);
        uint32_t dst       = ts + pre + loc;
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            if(!f[j]) continue;
            const uint32_t c = base + j;
            if(dst >= N)
            {
                ctl->overflow = 1;  // cannot happen while M <= N - ts; never write past the tree
                break;
            }
            for(int d = 0; d < S; ++d) tree.st[(size_t)d * N + dst] = ch.st[(size_t)d * N + c];
            for(int d = 0; d < C; ++d) tree.ct[(size_t)d * N + dst] = ch.ct[(size_t)d * N + c];
            tree.steps[dst]  = ch.steps[c];
            tree.parent[dst] = ch.parent[c];
            tree.reg[dst]    = ch.reg[c];
            const float dx = ch.st[c] - gx, dy = ch.st[(size_t)N + c] - gy, dz = ch.st[2 * (size_t)N + c] - gz;
            if(__builtin_sqrtf(dx * dx + dy * dy + dz * dz) < PP.goal_radius) atomicMin(&ctl->goal, (int32_t)dst);
            ++dst;
        }
        if(tile == ntiles - 1 && threadIdx.x == 0)
        {
            const uint32_t Fn = pre + agg;
            ctl->Fn           = Fn;
            ctl->f_accept     = f_accept(it.t, ts + Fn, PP);
        }
    }
}

// ---------------------------------------------------------------------------
// Next frontier: weights of all tree nodes (re-admission of old nodes with
// probability score + fAccept, KPAX.cu:505-517; repeat count 15 / 1 of new
// nodes by their region's valid count, KPAX.cu:474-479), scanned and
// run-length expanded into the repeat list (Planner.cu:52-80).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_frontier(Ctl* ctl, TreeView tree, uint32_t* R, RegionView rv,
                                                  uint64_t* status, PlanParams PP, uint64_t seed, uint32_t query)
{
    __shared__ uint32_t lds[4];
    __shared__ uint32_t s_tile, s_pre;
    const Iter it = ctl->cur;
    if(it.status != kRunning) return;
    const int32_t goal = ctl->goal;
    const uint32_t Fn  = ctl->Fn;
    if(goal != INT_MAX)
    {
        if(blockIdx.x == 0 && threadIdx.x == 0)
        {
            Iter n          = it;
            n.ts            = it.ts + Fn;
            n.status        = kDoneFound;
            ctl->nxt        = n;
            ctl->iterations = it.t;
        }
        return;
    }
    const float fa         = ctl->f_accept;
    const uint32_t N       = PP.N, ts_old = it.ts, tsn = it.ts + Fn;
    const uint32_t ntiles  = (tsn + kTile - 1) / kTile;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for(;;)
    {
        const uint32_t tile = next_tile(&ctl->tile_frontier, &s_tile);
        if(tile >= ntiles) break;
        const uint32_t base = tile * kTile + threadIdx.x * 4;
        uint32_t w[4], cnt = 0;
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            const uint32_t node = base + j;
            w[j]                = 0;
            if(node < tsn)
            {
                const int r = tree.reg[node];
                if(node < ts_old)
                {
                    const uint4 d = philox(node, it.t, query, PURPOSE_READMIT, k0, k1);
                    w[j]          = u01(d.x) <= rv.score[r] + fa ? 1u : 0u;
                }
                else
                    w[j] = rv.valid[r] < PP.sparse_valid_below ? PP.sparse_repeat : PP.dense_repeat;
            }
            cnt += w[j];
        }
        uint32_t agg;
        const uint32_t loc = block_excl_scan_256(cnt, &agg, lds);
        const uint32_t pre = tile_prefix(status, tile, agg, &s_pre);
        uint32_t pos       = pre + loc;
#pragma unroll
        for(int j = 0; j < 4; ++j)
            for(uint32_t k = 0; k < w[j]; ++k, ++pos)
                if(pos < N) R[pos] = base + j;
        if(tile == ntiles - 1 && threadIdx.x == 0)
        {
            ctl->nxt        = make_iter(it.t + 1, tsn, pre + agg, PP);
            ctl->iterations = it.t;
        }
    }
}

// Goal -> root walk by one thread (the reference walks it concurrently from
// every thread that reached the goal, KPAX.cu:482-501).  Rows go out in
// goal -> root order; the host reverses them.
__global__ void k_path(Ctl* ctl, TreeView tree, int S, int C, uint32_t N, float dt, float* rows, int32_t* steps,
                       int32_t* idx, uint32_t cap)
{
    if(threadIdx.x != 0 || blockIdx.x != 0) return;
    const int D  = S + C + 1;
    int32_t j    = ctl->goal;
    uint32_t len = 0;
    if(j != INT_MAX)
        while(j != -1 && len < cap)
        {
            float* row = rows + (size_t)len * D;
            for(int d = 0; d < S; ++d) row[d] = tree.st[(size_t)d * N + j];
            for(int d = 0; d < C; ++d) row[S + d] = tree.ct[(size_t)d * N + j];
            steps[len]  = tree.steps[j];
            row[S + C]  = dt * (float)tree.steps[j];
            idx[len]    = j;
            ++len;
            j = tree.parent[j];
        }
    ctl->path_len = (j == -1) ? len : 0u;
}

// ---------------------------------------------------------------------------
// Parity operators (one item per lane).
// ---------------------------------------------------------------------------
template<int MODEL>
__global__ __launch_bounds__(256) void k_propagate(ModelParams P, GridView G, const float* x0s, const float* us,
                                                   const int32_t* steps, uint32_t n, float* x1s, uint8_t* valid,
                                                   int32_t* fail)
{
    constexpr int S = Dims<MODEL>::S, C = Dims<MODEL>::C;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    float x0[S], u[C], x1[S];
#pragma unroll
    for(int d = 0; d < S; ++d) x0[d] = x0s[(size_t)d * n + i];
#pragma unroll
    for(int d = 0; d < C; ++d) u[d] = us[(size_t)d * n + i];
    const int f = propagate<MODEL>(P, G, x0, u, steps[i], x1);
#pragma unroll
    for(int d = 0; d < S; ++d) x1s[(size_t)d * n + i] = x1[d];
    valid[i] = f == 0 ? 1 : 0;
    fail[i]  = f;
}

__global__ __launch_bounds__(256) void k_collide(GridView G, const float* bmin, const float* bmax, uint32_t n,
                                                 uint8_t* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const float lo[3] = {bmin[i], bmin[(size_t)n + i], bmin[2 * (size_t)n + i]};
    const float hi[3] = {bmax[i], bmax[(size_t)n + i], bmax[2 * (size_t)n + i]};
    out[i]            = seg_free(G, lo, hi) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_regions(RegionParams RP, const float* st, uint32_t n, int32_t* r1,
                                                 int32_t* r2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const float p[3] = {st[i], st[(size_t)n + i], st[2 * (size_t)n + i]};
    locate(RP, p, &r1[i], &r2[i]);
}

__global__ __launch_bounds__(256) void k_histogram(RegionParams RP, const float* st, const uint8_t* valid, uint32_t n,
                                                   uint32_t t, RegionView rv)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const float p[3] = {st[i], st[(size_t)n + i], st[2 * (size_t)n + i]};
    int r1, r2;
    locate(RP, p, &r1, &r2);
    atomicAdd(&rv.total[r1], 1);
    if(valid[i])
    {
        atomicAdd(&rv.valid[r1], 1);
        claim_cell(rv.r2_state, r2, t, i);
    }
}

// Coverage after k_histogram: the winner of each newly reached cell counts it.
__global__ __launch_bounds__(256) void k_histogram_cover(RegionParams RP, const float* st, const uint8_t* valid,
                                                         uint32_t n, uint32_t t, RegionView rv)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n || !valid[i]) return;
    const float p[3] = {st[i], st[(size_t)n + i], st[2 * (size_t)n + i]};
    int r1, r2;
    locate(RP, p, &r1, &r2);
    if(rv.r2_state[r2] == cell_key(t, i)) atomicAdd(&rv.covered[r1], 1);
}

// Run-length expansion of weights (repeatInd after findInd).
__global__ __launch_bounds__(256) void k_repeat_expand(const uint32_t* w, uint32_t n, uint32_t* out, uint32_t cap,
                                                       uint64_t* status, uint32_t* ticket, uint32_t* total)
{
    __shared__ uint32_t lds[4];
    __shared__ uint32_t s_tile, s_pre;
    const uint32_t ntiles = (n + kTile - 1) / kTile;
    for(;;)
    {
        const uint32_t tile = next_tile(ticket, &s_tile);
        if(tile >= ntiles) break;
        const uint32_t base = tile * kTile + threadIdx.x * 4;
        uint32_t v[4], cnt = 0;
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            v[j] = base + j < n ? w[base + j] : 0u;
            cnt += v[j];
        }
        uint32_t agg;
        const uint32_t loc = block_excl_scan_256(cnt, &agg, lds);
        const uint32_t pre = tile_prefix(status, tile, agg, &s_pre);
        uint32_t pos       = pre + loc;
#pragma unroll
        for(int j = 0; j < 4; ++j)
            for(uint32_t k = 0; k < v[j]; ++k, ++pos)
                if(pos < cap) out[pos] = base + j;
        if(tile == ntiles - 1 && threadIdx.x == 0) *total = pre + agg;
    }
}

// Exclusive scan of counts into CSR offsets out[0..n] (out[n] = total).
__global__ __launch_bounds__(256) void k_scan_offsets(const int32_t* in, uint32_t n, int32_t* out, uint64_t* status,
                                                      uint32_t* ticket)
{
    __shared__ uint32_t lds[4];
    __shared__ uint32_t s_tile, s_pre;
    const uint32_t ntiles = (n + kTile - 1) / kTile;
    for(;;)
    {
        const uint32_t tile = next_tile(ticket, &s_tile);
        if(tile >= ntiles) break;
        const uint32_t base = tile * kTile + threadIdx.x * 4;
        uint32_t v[4], cnt = 0;
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            v[j] = base + j < n ? (uint32_t)in[base + j] : 0u;
            cnt += v[j];
        }
        uint32_t agg;
        const uint32_t loc = block_excl_scan_256(cnt, &agg, lds);
        const uint32_t pre = tile_prefix(status, tile, agg, &s_pre);
        uint32_t pos       = pre + loc;
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            if(base + j < n) out[base + j] = (int32_t)pos;
            pos += v[j];
        }
        if(tile == ntiles - 1 && threadIdx.x == 0) out[n] = (int32_t)(pre + agg);
    }
}

// Grid build (spatialHash.cu:9-54 without the 16-slot cap): count, scan,
// fill, then sort each cell's ids so the layout is deterministic.
__device__ __forceinline__ void obstacle_cells(const GridParams& gp, float4 a, float4 b, int lo[3], int hi[3])
{
    const float av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
#pragma unroll
    for(int d = 0; d < 3; ++d)
    {
        lo[d] = cell_of(av[d], gp.mn[d], gp.cell[d], gp.D[d]);
        hi[d] = cell_of(bv[d], gp.mn[d], gp.cell[d], gp.D[d]);
    }
}

__global__ void k_grid_count(GridParams gp, const float4* lo, const float4* hi, uint32_t n, int32_t* counts)
{
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if(o >= n) return;
    int a[3], b[3];
    obstacle_cells(gp, lo[o], hi[o], a, b);
    for(int z = a[2]; z <= b[2]; ++z)
        for(int y = a[1]; y <= b[1]; ++y)
            for(int x = a[0]; x <= b[0]; ++x) atomicAdd(&counts[x + gp.D[0] * (y + gp.D[1] * z)], 1);
}

__global__ void k_grid_fill(GridParams gp, const float4* lo, const float4* hi, uint32_t n, const int32_t* start,
                            int32_t* cursor, int32_t* items)
{
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if(o >= n) return;
    int a[3], b[3];
    obstacle_cells(gp, lo[o], hi[o], a, b);
    for(int z = a[2]; z <= b[2]; ++z)
        for(int y = a[1]; y <= b[1]; ++y)
            for(int x = a[0]; x <= b[0]; ++x)
            {
                const int c = x + gp.D[0] * (y + gp.D[1] * z);
                items[start[c] + atomicAdd(&cursor[c], 1)] = (int32_t)o;
            }
}

__global__ void k_grid_sort(const int32_t* start, int32_t* items, uint32_t G)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= G) return;
    const int b = start[c], e = start[c + 1];
    for(int i = b + 1; i < e; ++i)
    {
        const int32_t v = items[i];
        int j           = i - 1;
        while(j >= b && items[j] > v)
        {
            items[j + 1] = items[j];
            --j;
        }
        items[j + 1] = v;
    }
}

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)std::max<uint64_t>(1, (n + bs - 1) / bs); }

}  // namespace

// ===========================================================================
// Handle
// ===========================================================================
struct kpax_handle {
    kpax_config cfg{};
    int S = 0, C = 0;
    ModelParams P{};
    RegionParams RP{};
    GridParams GPm{};
    PlanParams PP{};
    std::string err;
    int cu_count          = 256;
    uint32_t expand_blocks = 1;
    void* expand_args      = nullptr;  // device copy of ExpandArgs
    hipStream_t stream    = nullptr;
    // tree and children
    TreeView tree{};
    ChildView ch{};
    uint32_t* R           = nullptr;
    RegionView rv{};
    Ctl* ctl              = nullptr;
    Ctl* h_ctl            = nullptr;  // pinned
    Iter* h_iter          = nullptr;  // pinned, 2 slots
    hipEvent_t it_ev[2]   = {nullptr, nullptr};
    uint64_t* st_append   = nullptr;
    uint64_t* st_frontier = nullptr;
    uint32_t n_status     = 0;
    float* d_start        = nullptr;
    float* path_rows      = nullptr;
    int32_t* path_steps   = nullptr;
    int32_t* path_idx     = nullptr;
    uint32_t path_cap     = 0;
    // environment
    std::vector<float> obs_host;
    uint32_t n_obs        = 0;
    float4* obs_lo        = nullptr;
    float4* obs_hi        = nullptr;
    int32_t* grid_start   = nullptr;
    int32_t* grid_items   = nullptr;
    uint32_t grid_cells   = 0;
    uint32_t grid_k       = 0;
    // last plan
    uint32_t last_tree    = 0;
    bool planned          = false;
    // profiling; KPAX_TRACE=1 prints every iteration record to stderr
    bool profiling        = false;
    bool trace            = false;
    std::vector<hipEvent_t> ev;
    kpax_stats stats{};
    std::vector<void*> allocs;
};

namespace {

int fail(kpax_handle* h, int code, const std::string& msg)
{
    if(h) h->err = msg;
    return code;
}

#define HIPC(expr)                                                                                 \
    do                                                                                             \
    {                                                                                              \
        hipError_t e_ = (expr);                                                                    \
        if(e_ != hipSuccess)                                                                       \
            return fail(h, KPAX_DEVICE_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while(0)

template<class T>
int dalloc(kpax_handle* h, T** p, size_t count)
{
    void* q       = nullptr;
    hipError_t e  = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if(e != hipSuccess) return fail(h, KPAX_NO_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    h->allocs.push_back(q);
    *p = (T*)q;
    return KPAX_OK;
}

GridView grid_view(const kpax_handle* h)
{
    GridView g;
    g.gp    = h->GPm;
    g.start = h->grid_start;
    g.items = h->grid_items;
    g.lo    = h->obs_lo;
    g.hi    = h->obs_hi;
    return g;
}

// Scratch device buffers for the parity operators (allocated per call).
struct Scratch {
    std::vector<void*> ptrs;
    ~Scratch()
    {
        for(void* p : ptrs) (void)hipFree(p);
    }
    template<class T> T* get(size_t count)
    {
        void* q = nullptr;
        if(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(q);
        return (T*)q;
    }
};

size_t expand_lds_bytes(const kpax_handle* h)
{
    return h->RP.r1_count <= kLdsRegions ? 2 * (size_t)h->RP.r1_count * sizeof(int32_t) : 0;
}

// Work-groups of k_expand that fit on the device at once: its waves are
// persistent, so more would only wait for the first ones to drain the queue.
template<int MODEL>
uint32_t expand_blocks_for(const kpax_handle* h)
{
    int per_cu = 0;
    if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_expand<MODEL>, 256, expand_lds_bytes(h)) != hipSuccess ||
       per_cu < 1)
        per_cu = 1;
    return (uint32_t)per_cu * (uint32_t)h->cu_count;
}

uint32_t expand_blocks(const kpax_handle* h)
{
    switch(h->cfg.model)
    {
        case KPAX_MODEL_DOUBLE_INTEGRATOR: return expand_blocks_for<KPAX_MODEL_DOUBLE_INTEGRATOR>(h);
        case KPAX_MODEL_DUBINS_AIRPLANE: return expand_blocks_for<KPAX_MODEL_DUBINS_AIRPLANE>(h);
        default: return expand_blocks_for<KPAX_MODEL_QUADCOPTER>(h);
    }
}

// Writes the device-side argument block of k_expand (after creation and
// whenever the collision grid moves).
int upload_expand_args(kpax_handle* h)
{
    ExpandArgs a;
    a.ctl  = h->ctl;
    a.tree = h->tree;
    a.ch   = h->ch;
    a.R    = h->R;
    a.rv   = h->rv;
    a.P    = h->P;
    a.RP   = h->RP;
    a.G    = grid_view(h);
    a.PP   = h->PP;
    HIPC(hipMemcpy(h->expand_args, &a, sizeof(a), hipMemcpyHostToDevice));
    return KPAX_OK;
}

int launch_expand(kpax_handle* h, uint64_t seed, uint32_t query, uint32_t blocks)
{
    const size_t lds = expand_lds_bytes(h);
    const ExpandArgs* A = (const ExpandArgs*)h->expand_args;
    switch(h->cfg.model)
    {
        case KPAX_MODEL_DOUBLE_INTEGRATOR:
            hipLaunchKernelGGL(k_expand<KPAX_MODEL_DOUBLE_INTEGRATOR>, dim3(blocks), dim3(256), lds, h->stream, A, seed,
                               query);
            break;
        case KPAX_MODEL_DUBINS_AIRPLANE:
            hipLaunchKernelGGL(k_expand<KPAX_MODEL_DUBINS_AIRPLANE>, dim3(blocks), dim3(256), lds, h->stream, A, seed,
                               query);
            break;
        default:
            hipLaunchKernelGGL(k_expand<KPAX_MODEL_QUADCOPTER>, dim3(blocks), dim3(256), lds, h->stream, A, seed,
                               query);
            break;
    }
    return KPAX_OK;
}

void ensure_events(kpax_handle* h)
{
    const size_t need = 2 * 5 * ((size_t)h->cfg.max_iter + 3);
    while(h->ev.size() < need)
    {
        hipEvent_t e;
        if(hipEventCreate(&e) != hipSuccess) break;
        h->ev.push_back(e);
    }
}

inline void rec(kpax_handle* h, size_t slot)
{
    if(h->profiling && slot < h->ev.size()) (void)hipEventRecord(h->ev[slot], h->stream);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int kpax_create(const kpax_config* cfg, kpax_handle** out)
{
    if(!cfg || !out) return KPAX_INVALID_ARG;
    *out = nullptr;
    kpax_handle* h = new(std::nothrow) kpax_handle();
    if(!h) return KPAX_NO_MEMORY;
    const std::string why = validate_config(*cfg);
    if(!why.empty())
    {
        delete h;
        return KPAX_INVALID_ARG;
    }
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0)
    {
        delete h;
        return KPAX_NO_DEVICE;
    }
    h->cfg   = *cfg;
    h->trace = std::getenv("KPAX_TRACE") != nullptr;
    h->S     = state_dim(cfg->model);
    h->C   = control_dim(cfg->model);
    h->P   = make_model_params(*cfg);
    h->RP  = make_region_params(*cfg);
    h->GPm = make_grid_params(*cfg);
    h->PP  = make_plan_params(*cfg);
    int rc = KPAX_OK;
#define TRY(x)                   \
    do                           \
    {                            \
        rc = (x);                \
        if(rc != KPAX_OK)        \
        {                        \
            kpax_destroy(h);     \
            return rc;           \
        }                        \
    } while(0)
    if(hipSetDevice(cfg->device) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    hipDeviceProp_t prop;
    if(hipGetDeviceProperties(&prop, cfg->device) == hipSuccess && prop.multiProcessorCount > 0)
        h->cu_count = prop.multiProcessorCount;
    if(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    h->expand_blocks = expand_blocks(h);
    const size_t N = cfg->max_tree;
    TRY(dalloc(h, &h->tree.st, N * h->S));
    TRY(dalloc(h, &h->tree.ct, N * h->C));
    TRY(dalloc(h, &h->tree.steps, N));
    TRY(dalloc(h, &h->tree.parent, N));
    TRY(dalloc(h, &h->tree.reg, N));
    TRY(dalloc(h, &h->ch.st, N * h->S));
    TRY(dalloc(h, &h->ch.ct, N * h->C));
    TRY(dalloc(h, &h->ch.steps, N));
    TRY(dalloc(h, &h->ch.parent, N));
    TRY(dalloc(h, &h->ch.reg, N));
    TRY(dalloc(h, &h->ch.r2, N));
    TRY(dalloc(h, &h->ch.adm, N));
    TRY(dalloc(h, &h->R, N));
    const size_t r1 = (size_t)h->RP.r1_count, r2 = r1 * h->RP.r2_per;
    TRY(dalloc(h, &h->rv.valid, r1));
    TRY(dalloc(h, &h->rv.total, r1));
    TRY(dalloc(h, &h->rv.covered, r1));
    TRY(dalloc(h, &h->rv.score, r1));
    TRY(dalloc(h, &h->rv.r2_state, r2));
    TRY(dalloc(h, &h->ctl, 1));
    h->n_status = (uint32_t)((2 * N + kTile - 1) / kTile + 1);
    TRY(dalloc(h, &h->st_append, h->n_status));
    TRY(dalloc(h, &h->st_frontier, h->n_status));
    TRY(dalloc(h, &h->d_start, 16));
    h->path_cap = cfg->max_iter + 2;
    TRY(dalloc(h, &h->path_rows, (size_t)h->path_cap * (h->S + h->C + 1)));
    TRY(dalloc(h, &h->path_steps, h->path_cap));
    TRY(dalloc(h, &h->path_idx, h->path_cap));
    if(hipHostMalloc((void**)&h->h_ctl, sizeof(Ctl), hipHostMallocDefault) != hipSuccess) TRY(KPAX_NO_MEMORY);
    if(hipHostMalloc((void**)&h->h_iter, 2 * sizeof(Iter), hipHostMallocDefault) != hipSuccess) TRY(KPAX_NO_MEMORY);
    for(int i = 0; i < 2; ++i)
        if(hipEventCreateWithFlags(&h->it_ev[i], hipEventDisableTiming) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    // an empty environment until kpax_set_obstacles
    h->grid_cells = (uint32_t)h->GPm.D[0] * h->GPm.D[1] * h->GPm.D[2];
    TRY(dalloc(h, &h->grid_start, (size_t)h->grid_cells + 1));
    if(hipMemset(h->grid_start, 0, ((size_t)h->grid_cells + 1) * sizeof(int32_t)) != hipSuccess)
        TRY(KPAX_DEVICE_ERROR);
    if(hipMemset(h->st_append, 0, h->n_status * sizeof(uint64_t)) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    if(hipMemset(h->st_frontier, 0, h->n_status * sizeof(uint64_t)) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    TRY(dalloc(h, (char**)&h->expand_args, sizeof(ExpandArgs)));
    TRY(upload_expand_args(h));
#undef TRY
    *out = h;
    return KPAX_OK;
}

int kpax_destroy(kpax_handle* h)
{
    if(!h) return KPAX_INVALID_ARG;
    if(h->stream) (void)hipStreamSynchronize(h->stream);
    for(void* p : h->allocs) (void)hipFree(p);
    if(h->obs_lo) (void)hipFree(h->obs_lo);
    if(h->obs_hi) (void)hipFree(h->obs_hi);
    if(h->grid_items) (void)hipFree(h->grid_items);
    if(h->h_ctl) (void)hipHostFree(h->h_ctl);
    if(h->h_iter) (void)hipHostFree(h->h_iter);
    for(hipEvent_t e : h->ev) (void)hipEventDestroy(e);
    for(hipEvent_t e : h->it_ev)
        if(e) (void)hipEventDestroy(e);
    if(h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return KPAX_OK;
}

const char* kpax_last_error(const kpax_handle* h)
{
    if(!h) return "null handle";
    return h->err.c_str();
}

int kpax_set_obstacles(kpax_handle* h, const float* aabb6, uint32_t n)
{
    if(!h) return KPAX_INVALID_ARG;
    if(n && !aabb6) return fail(h, KPAX_INVALID_ARG, "obstacles pointer is NULL");
    for(uint32_t i = 0; i < n; ++i)
        for(int d = 0; d < 3; ++d)
        {
            const float a = aabb6[6 * (size_t)i + d], b = aabb6[6 * (size_t)i + 3 + d];
            if(!std::isfinite(a) || !std::isfinite(b) || a > b)
                return fail(h, KPAX_INVALID_ARG, "obstacle " + std::to_string(i) + " is not a finite min<=max box");
        }
    HIPC(hipSetDevice(h->cfg.device));
    HIPC(hipStreamSynchronize(h->stream));
    if(h->obs_lo) (void)hipFree(h->obs_lo);
    if(h->obs_hi) (void)hipFree(h->obs_hi);
    if(h->grid_items) (void)hipFree(h->grid_items);
    h->obs_lo = h->obs_hi = nullptr;
    h->grid_items         = nullptr;
    h->obs_host.assign(aabb6, aabb6 + 6 * (size_t)n);
    h->n_obs = n;
    std::vector<float4> lo(std::max<uint32_t>(n, 1)), hi(std::max<uint32_t>(n, 1));
    for(uint32_t i = 0; i < n; ++i)
    {
        const float* o = aabb6 + 6 * (size_t)i;
        lo[i]          = make_float4(o[0], o[1], o[2], 0.0f);
        hi[i]          = make_float4(o[3], o[4], o[5], 0.0f);
    }
    HIPC(hipMalloc(&h->obs_lo, lo.size() * sizeof(float4)));
    HIPC(hipMalloc(&h->obs_hi, hi.size() * sizeof(float4)));
    HIPC(hipMemcpy(h->obs_lo, lo.data(), lo.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(h->obs_hi, hi.data(), hi.size() * sizeof(float4), hipMemcpyHostToDevice));
    const uint32_t G = h->grid_cells;
    Scratch sc;
    int32_t* counts  = sc.get<int32_t>(G);
    int32_t* cursor  = sc.get<int32_t>(G);
    uint64_t* status = sc.get<uint64_t>(G / kTile + 2);
    uint32_t* ticket = sc.get<uint32_t>(1);
    if(!counts || !cursor || !status || !ticket) return fail(h, KPAX_NO_MEMORY, "grid scratch allocation failed");
    HIPC(hipMemsetAsync(counts, 0, (size_t)G * 4, h->stream));
    HIPC(hipMemsetAsync(cursor, 0, (size_t)G * 4, h->stream));
    HIPC(hipMemsetAsync(status, 0, (G / kTile + 2) * 8, h->stream));
    HIPC(hipMemsetAsync(ticket, 0, 4, h->stream));
    if(n) hipLaunchKernelGGL(k_grid_count, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->GPm, h->obs_lo, h->obs_hi, n, counts);
    hipLaunchKernelGGL(k_scan_offsets, dim3(std::min<uint32_t>(blocks_for(G, kTile), 4 * h->cu_count)), dim3(256), 0, h->stream, 
      counts, G, h->grid_start, status, ticket);
    HIPC(hipGetLastError());
    int32_t K = 0;
    HIPC(hipMemcpyAsync(&K, h->grid_start + G, 4, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->grid_k = (uint32_t)K;
    HIPC(hipMalloc(&h->grid_items, std::max<size_t>(K, 1) * sizeof(int32_t)));
    if(n) hipLaunchKernelGGL(k_grid_fill, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->GPm, h->obs_lo, h->obs_hi, n, h->grid_start,
                                                                  cursor, h->grid_items);
    hipLaunchKernelGGL(k_grid_sort, dim3(blocks_for(G, 256)), dim3(256), 0, h->stream, h->grid_start, h->grid_items, G);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    return upload_expand_args(h);
}

int kpax_plan(kpax_handle* h, const float* start, const float* goal, uint64_t seed, kpax_result* out)
{
    if(!h || !start || !goal || !out) return fail(h, KPAX_INVALID_ARG, "null argument");
    memset(out, 0, sizeof(*out));
    out->goal_index = -1;
    out->verified   = -1;
    for(int d = 0; d < h->S; ++d)
        if(!std::isfinite(start[d])) return fail(h, KPAX_INVALID_ARG, "start state is not finite");
    const auto t0 = std::chrono::steady_clock::now();
    HIPC(hipSetDevice(h->cfg.device));
    const uint32_t N = h->PP.N;
    if(h->profiling) ensure_events(h);
    float sg[16] = {};
    for(int d = 0; d < h->S; ++d) sg[d] = start[d];
    HIPC(hipMemcpyAsync(h->d_start, sg, sizeof(sg), hipMemcpyHostToDevice, h->stream));
    const uint32_t fb = std::min<uint32_t>(2048, 8 * h->cu_count);
    rec(h, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(fb), dim3(256), 0, h->stream, h->rv.valid, 0, h->RP.r1_count);
    hipLaunchKernelGGL(k_fill_i32, dim3(fb), dim3(256), 0, h->stream, h->rv.total, 0, h->RP.r1_count);
    hipLaunchKernelGGL(k_fill_i32, dim3(fb), dim3(256), 0, h->stream, h->rv.covered, 0, h->RP.r1_count);
    hipLaunchKernelGGL(k_fill_f32, dim3(fb), dim3(256), 0, h->stream, h->rv.score, 0.0f, h->RP.r1_count);
    hipLaunchKernelGGL(k_fill_u64, dim3(fb), dim3(256), 0, h->stream, h->rv.r2_state, ~0ull, (uint64_t)h->RP.r1_count * h->RP.r2_per);
    hipLaunchKernelGGL(k_init, dim3(1), dim3(256), 0, h->stream, h->ctl, h->tree, h->R, h->d_start, h->S, h->C, h->RP, h->PP);
    rec(h, 1);
    HIPC(hipGetLastError());
    const uint32_t eb = std::min<uint32_t>(blocks_for(N, 256), h->expand_blocks);
    const uint32_t sb = std::min<uint32_t>(blocks_for(N, kTile), 4 * h->cu_count);
    uint32_t launched = 0;
    for(uint32_t it = 1; it <= h->cfg.max_iter; ++it)
    {
        const size_t e = 10 * (size_t)it;
        hipLaunchKernelGGL(k_advance, dim3(1), dim3(256), 0, h->stream, h->ctl, h->st_append, h->st_frontier, h->PP);
        rec(h, e + 0);
        launch_expand(h, seed, 0, eb);
        rec(h, e + 1);
        hipLaunchKernelGGL(k_append, dim3(sb), dim3(256), 0, h->stream, h->ctl, h->ch, h->tree, h->rv.r2_state,
                           h->rv.covered, h->st_append, h->PP, h->S, h->C, goal[0], goal[1], goal[2]);
        rec(h, e + 2);
        hipLaunchKernelGGL(k_score, dim3(1), dim3(1024), 0, h->stream, h->ctl, h->rv, h->RP, h->PP.eps, 1);
        rec(h, e + 3);
        hipLaunchKernelGGL(k_frontier, dim3(sb), dim3(256), 0, h->stream, h->ctl, h->tree, h->R, h->rv, h->st_frontier, h->PP, seed, 0);
        rec(h, e + 4);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(&h->h_iter[it & 1], &h->ctl->nxt, sizeof(Iter), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipEventRecord(h->it_ev[it & 1], h->stream));
        launched = it;
        // One iteration stays queued behind the one whose status is read, so
        // the GPU never waits for the host; an iteration queued after the
        // plan ended finds status != 0 and returns at once.
        if(it >= 2)
        {
            HIPC(hipEventSynchronize(h->it_ev[(it - 1) & 1]));
            if(h->trace)
            {
                const Iter& n = h->h_iter[(it - 1) & 1];
                std::fprintf(stderr, "kpax: after iteration %u: next t=%u ts=%u W=%u M=%u fan=%u status=%d\n",
                             it - 1, n.t, n.ts, n.W, n.M, n.fan, n.status);
            }
            if(h->h_iter[(it - 1) & 1].status != kRunning) break;
        }
    }
    HIPC(hipStreamSynchronize(h->stream));
    HIPC(hipMemcpy(h->h_ctl, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    const Ctl c          = *h->h_ctl;
    if(c.overflow) return fail(h, KPAX_DEVICE_ERROR, "tree index past the configured capacity");
    int status;
    switch(c.nxt.status)
    {
        case kDoneFound: status = KPAX_OK; break;
        case kDoneBudget: status = KPAX_BUDGET_EXHAUSTED; break;
        default: status = KPAX_NOT_FOUND; break;  // max_iter reached
    }
    out->status          = status;
    out->iterations      = c.iterations;
    out->tree_size       = c.nxt.ts;
    out->expansions      = c.expansions;
    out->integrated_steps = c.steps_done;
    out->state_dim       = h->S;
    out->control_dim     = h->C;
    h->last_tree         = c.nxt.ts;
    h->planned           = true;
    if(status == KPAX_OK && c.goal != INT_MAX)
    {
        out->goal_index = c.goal;
        hipLaunchKernelGGL(k_path, dim3(1), dim3(64), 0, h->stream, h->ctl, h->tree, h->S, h->C, N, h->cfg.dt, h->path_rows, h->path_steps,
                                         h->path_idx, h->path_cap);
        HIPC(hipGetLastError());
        uint32_t len = 0;
        HIPC(hipMemcpyAsync(&len, &h->ctl->path_len, 4, hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        const int D = h->S + h->C + 1;
        std::vector<float> rows((size_t)len * D);
        std::vector<int32_t> st(len), ix(len);
        if(len)
        {
            HIPC(hipMemcpy(rows.data(), h->path_rows, rows.size() * 4, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(st.data(), h->path_steps, len * 4, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(ix.data(), h->path_idx, len * 4, hipMemcpyDeviceToHost));
        }
        out->n_rows     = len;
        out->rows       = (float*)malloc(std::max<size_t>(rows.size(), 1) * 4);
        out->steps      = (int32_t*)malloc(std::max<size_t>(len, 1) * 4);
        out->node_index = (int32_t*)malloc(std::max<size_t>(len, 1) * 4);
        for(uint32_t r = 0; r < len; ++r)
        {
            const uint32_t s = len - 1 - r;  // root first
            memcpy(out->rows + (size_t)r * D, rows.data() + (size_t)s * D, D * 4);
            out->steps[r]      = st[s];
            out->node_index[r] = ix[s];
        }
    }
    const auto t1 = std::chrono::steady_clock::now();
    out->plan_ms  = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if(h->profiling)
    {
        float ms = 0.0f;
        if(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]) == hipSuccess)
        {
            h->stats.phase_ms[4] += ms;
            h->stats.phase_launches[4] += 1;
        }
        for(uint32_t it = 1; it <= std::min(launched, c.iterations); ++it)
        {
            const size_t e = 10 * (size_t)it;
            for(int p = 0; p < 4; ++p)
                if(hipEventElapsedTime(&ms, h->ev[e + p], h->ev[e + p + 1]) == hipSuccess)
                {
                    h->stats.phase_ms[p] += ms;
                    h->stats.phase_launches[p] += 1;
                }
        }
        h->stats.expansions += c.expansions;
        h->stats.integrated_steps += c.steps_done;
        h->stats.plans += 1;
    }
    if(h->cfg.verify && out->n_rows > 0)
    {
        float dev = 0.0f;
        const bool ok = verify_path_fp64(h->cfg, h->obs_host.data(), h->n_obs, out->rows, out->steps, out->n_rows,
                                         goal, &dev);
        out->verified       = ok ? 1 : 0;
        out->verify_max_dev = dev;
    }
    return KPAX_OK;
}

void kpax_result_free(kpax_result* r)
{
    if(!r) return;
    free(r->rows);
    free(r->steps);
    free(r->node_index);
    r->rows       = nullptr;
    r->steps      = nullptr;
    r->node_index = nullptr;
    r->n_rows     = 0;
}

int kpax_get_tree(kpax_handle* h, uint32_t cap, float* states, float* controls, int32_t* steps, int32_t* parents,
                  int32_t* regions, uint32_t* tree_size)
{
    if(!h) return KPAX_INVALID_ARG;
    if(!h->planned) return fail(h, KPAX_INVALID_ARG, "no plan has run on this handle");
    const uint32_t ts = h->last_tree, N = h->PP.N;
    if(tree_size) *tree_size = ts;
    if(cap < ts && (states || controls)) return fail(h, KPAX_INVALID_ARG, "cap is smaller than the tree");
    HIPC(hipSetDevice(h->cfg.device));
    for(int d = 0; states && d < h->S; ++d)
        HIPC(hipMemcpy(states + (size_t)d * cap, h->tree.st + (size_t)d * N, (size_t)ts * 4, hipMemcpyDeviceToHost));
    for(int d = 0; controls && d < h->C; ++d)
        HIPC(hipMemcpy(controls + (size_t)d * cap, h->tree.ct + (size_t)d * N, (size_t)ts * 4, hipMemcpyDeviceToHost));
    if(steps) HIPC(hipMemcpy(steps, h->tree.steps, (size_t)ts * 4, hipMemcpyDeviceToHost));
    if(parents) HIPC(hipMemcpy(parents, h->tree.parent, (size_t)ts * 4, hipMemcpyDeviceToHost));
    if(regions) HIPC(hipMemcpy(regions, h->tree.reg, (size_t)ts * 4, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_get_regions(kpax_handle* h, int32_t* valid, int32_t* total, int32_t* covered, float* scores)
{
    if(!h) return KPAX_INVALID_ARG;
    HIPC(hipSetDevice(h->cfg.device));
    const size_t n = (size_t)h->RP.r1_count * 4;
    if(valid) HIPC(hipMemcpy(valid, h->rv.valid, n, hipMemcpyDeviceToHost));
    if(total) HIPC(hipMemcpy(total, h->rv.total, n, hipMemcpyDeviceToHost));
    if(covered) HIPC(hipMemcpy(covered, h->rv.covered, n, hipMemcpyDeviceToHost));
    if(scores) HIPC(hipMemcpy(scores, h->rv.score, n, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_get_grid(kpax_handle* h, int32_t* cell_start, int32_t* items, uint32_t items_cap, uint32_t* n_cells,
                  uint32_t* n_items)
{
    if(!h) return KPAX_INVALID_ARG;
    if(n_cells) *n_cells = h->grid_cells;
    if(n_items) *n_items = h->grid_k;
    HIPC(hipSetDevice(h->cfg.device));
    if(cell_start)
        HIPC(hipMemcpy(cell_start, h->grid_start, ((size_t)h->grid_cells + 1) * 4, hipMemcpyDeviceToHost));
    if(items && h->grid_k)
    {
        if(items_cap < h->grid_k) return fail(h, KPAX_INVALID_ARG, "items_cap is smaller than the grid");
        HIPC(hipMemcpy(items, h->grid_items, (size_t)h->grid_k * 4, hipMemcpyDeviceToHost));
    }
    return KPAX_OK;
}

int kpax_set_profiling(kpax_handle* h, int32_t on)
{
    if(!h) return KPAX_INVALID_ARG;
    h->profiling = on != 0;
    if(h->profiling) ensure_events(h);
    return KPAX_OK;
}

int kpax_get_stats(kpax_handle* h, kpax_stats* out)
{
    if(!h || !out) return KPAX_INVALID_ARG;
    *out = h->stats;
    return KPAX_OK;
}

int kpax_reset_stats(kpax_handle* h)
{
    if(!h) return KPAX_INVALID_ARG;
    memset(&h->stats, 0, sizeof(h->stats));
    return KPAX_OK;
}

// ---- parity operators ------------------------------------------------------

int kpax_propagate_batch(kpax_handle* h, const float* x0_soa, const float* u_soa, const int32_t* steps, uint32_t n,
                         float* x1_soa, uint8_t* valid, int32_t* fail_step)
{
    if(!h || (n && (!x0_soa || !u_soa || !steps || !x1_soa || !valid || !fail_step)))
        return fail(h, KPAX_INVALID_ARG, "null argument");
    if(n == 0) return KPAX_OK;
    for(uint32_t i = 0; i < n; ++i)
        if(steps[i] < 0 || steps[i] > 1000000) return fail(h, KPAX_INVALID_ARG, "step count out of range");
    HIPC(hipSetDevice(h->cfg.device));
    Scratch sc;
    const int S = h->S, C = h->C;
    float* dx0 = sc.get<float>((size_t)S * n);
    float* du  = sc.get<float>((size_t)C * n);
    int32_t* dst = sc.get<int32_t>(n);
    float* dx1 = sc.get<float>((size_t)S * n);
    uint8_t* dv = sc.get<uint8_t>(n);
    int32_t* df = sc.get<int32_t>(n);
    if(!dx0 || !du || !dst || !dx1 || !dv || !df) return fail(h, KPAX_NO_MEMORY, "scratch allocation failed");
    HIPC(hipMemcpy(dx0, x0_soa, (size_t)S * n * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(du, u_soa, (size_t)C * n * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(dst, steps, (size_t)n * 4, hipMemcpyHostToDevice));
    const GridView G = grid_view(h);
    const uint32_t b = blocks_for(n, 256);
    switch(h->cfg.model)
    {
        case KPAX_MODEL_DOUBLE_INTEGRATOR:
            hipLaunchKernelGGL(k_propagate<KPAX_MODEL_DOUBLE_INTEGRATOR>, dim3(b), dim3(256), 0, h->stream, h->P, G, dx0, du, dst, n, dx1, dv, df);
            break;
        case KPAX_MODEL_DUBINS_AIRPLANE:
            hipLaunchKernelGGL(k_propagate<KPAX_MODEL_DUBINS_AIRPLANE>, dim3(b), dim3(256), 0, h->stream, h->P, G, dx0, du, dst, n, dx1, dv, df);
            break;
        default:
            hipLaunchKernelGGL(k_propagate<KPAX_MODEL_QUADCOPTER>, dim3(b), dim3(256), 0, h->stream, h->P, G, dx0, du, dst, n, dx1, dv, df);
            break;
    }
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    HIPC(hipMemcpy(x1_soa, dx1, (size_t)S * n * 4, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(valid, dv, n, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(fail_step, df, (size_t)n * 4, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_collide_batch(kpax_handle* h, const float* bbmin_soa, const float* bbmax_soa, uint32_t n, uint8_t* free_out)
{
    if(!h || (n && (!bbmin_soa || !bbmax_soa || !free_out))) return fail(h, KPAX_INVALID_ARG, "null argument");
    if(n == 0) return KPAX_OK;
    HIPC(hipSetDevice(h->cfg.device));
    Scratch sc;
    float* lo = sc.get<float>(3 * (size_t)n);
    float* hi = sc.get<float>(3 * (size_t)n);
    uint8_t* o = sc.get<uint8_t>(n);
    if(!lo || !hi || !o) return fail(h, KPAX_NO_MEMORY, "scratch allocation failed");
    HIPC(hipMemcpy(lo, bbmin_soa, 12 * (size_t)n, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(hi, bbmax_soa, 12 * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_collide, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, grid_view(h), lo, hi, n, o);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    HIPC(hipMemcpy(free_out, o, n, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_region_batch(kpax_handle* h, const float* states_soa, uint32_t n, int32_t* r1, int32_t* r2)
{
    if(!h || (n && (!states_soa || !r1 || !r2))) return fail(h, KPAX_INVALID_ARG, "null argument");
    if(n == 0) return KPAX_OK;
    HIPC(hipSetDevice(h->cfg.device));
    Scratch sc;
    float* st  = sc.get<float>(3 * (size_t)n);
    int32_t* a = sc.get<int32_t>(n);
    int32_t* b = sc.get<int32_t>(n);
    if(!st || !a || !b) return fail(h, KPAX_NO_MEMORY, "scratch allocation failed");
    HIPC(hipMemcpy(st, states_soa, 12 * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_regions, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->RP, st, n, a, b);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    HIPC(hipMemcpy(r1, a, 4 * (size_t)n, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(r2, b, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_region_histogram(kpax_handle* h, const float* states_soa, const uint8_t* valid, uint32_t n, int32_t* total,
                          int32_t* valid_count, int32_t* covered)
{
    if(!h || !total || !valid_count || !covered || (n && (!states_soa || !valid)))
        return fail(h, KPAX_INVALID_ARG, "null argument");
    HIPC(hipSetDevice(h->cfg.device));
    const size_t r1 = (size_t)h->RP.r1_count, r2 = r1 * h->RP.r2_per;
    Scratch sc;
    float* st  = sc.get<float>(3 * (size_t)n);
    uint8_t* v = sc.get<uint8_t>(n);
    RegionView rv{};
    rv.total    = sc.get<int32_t>(r1);
    rv.valid    = sc.get<int32_t>(r1);
    rv.covered  = sc.get<int32_t>(r1);
    rv.r2_state = sc.get<uint64_t>(r2);
    if(!st || !v || !rv.total || !rv.valid || !rv.covered || !rv.r2_state)
        return fail(h, KPAX_NO_MEMORY, "scratch allocation failed");
    if(n)
    {
        HIPC(hipMemcpy(st, states_soa, 12 * (size_t)n, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(v, valid, n, hipMemcpyHostToDevice));
    }
    HIPC(hipMemcpy(rv.total, total, r1 * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(rv.valid, valid_count, r1 * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(rv.covered, covered, r1 * 4, hipMemcpyHostToDevice));
    // the batch is one iteration on an empty coverage map
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(r2, 256)), dim3(256), 0, h->stream, rv.r2_state, ~0ull, r2);
    if(n)
    {
        hipLaunchKernelGGL(k_histogram, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->RP, st, v, n, 1u, rv);
        hipLaunchKernelGGL(k_histogram_cover, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->RP, st, v, n, 1u,
                           rv);
    }
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    HIPC(hipMemcpy(total, rv.total, r1 * 4, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(valid_count, rv.valid, r1 * 4, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(covered, rv.covered, r1 * 4, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_scores(kpax_handle* h, const int32_t* valid_count, const int32_t* total, const int32_t* covered, float* scores)
{
    if(!h || !valid_count || !total || !covered || !scores) return fail(h, KPAX_INVALID_ARG, "null argument");
    HIPC(hipSetDevice(h->cfg.device));
    const size_t r1 = (size_t)h->RP.r1_count;
    Scratch sc;
    RegionView rv{};
    rv.valid   = sc.get<int32_t>(r1);
    rv.total   = sc.get<int32_t>(r1);
    rv.covered = sc.get<int32_t>(r1);
    rv.score   = sc.get<float>(r1);
    if(!rv.valid || !rv.total || !rv.covered || !rv.score) return fail(h, KPAX_NO_MEMORY, "scratch allocation failed");
    HIPC(hipMemcpy(rv.valid, valid_count, r1 * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(rv.total, total, r1 * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(rv.covered, covered, r1 * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_score, dim3(1), dim3(1024), 0, h->stream, nullptr, rv, h->RP, h->PP.eps, 0);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    HIPC(hipMemcpy(scores, rv.score, r1 * 4, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_repeat_expand(kpax_handle* h, const uint32_t* weights, uint32_t n, uint32_t* out, uint32_t out_cap,
                       uint32_t* total)
{
    if(!h || !total || (n && !weights) || (out_cap && !out)) return fail(h, KPAX_INVALID_ARG, "null argument");
    uint64_t sum = 0;
    for(uint32_t i = 0; i < n; ++i) sum += weights[i];
    if(sum > 0xffffffffull) return fail(h, KPAX_INVALID_ARG, "total weight exceeds 2^32-1");
    *total = 0;
    if(n == 0) return KPAX_OK;
    HIPC(hipSetDevice(h->cfg.device));
    Scratch sc;
    const uint32_t ntiles = (n + kTile - 1) / kTile;
    uint32_t* dw     = sc.get<uint32_t>(n);
    uint32_t* dout   = sc.get<uint32_t>(std::max<uint32_t>(out_cap, 1));
    uint64_t* status = sc.get<uint64_t>(ntiles + 1);
    uint32_t* ticket = sc.get<uint32_t>(2);
    if(!dw || !dout || !status || !ticket) return fail(h, KPAX_NO_MEMORY, "scratch allocation failed");
    HIPC(hipMemcpy(dw, weights, (size_t)n * 4, hipMemcpyHostToDevice));
    HIPC(hipMemsetAsync(status, 0, (ntiles + 1) * 8, h->stream));
    HIPC(hipMemsetAsync(ticket, 0, 8, h->stream));
    hipLaunchKernelGGL(k_repeat_expand, dim3(std::min<uint32_t>(ntiles, 4 * h->cu_count)), dim3(256), 0, h->stream, dw, n, dout, out_cap, status,
                                                                                         ticket, ticket + 1);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    HIPC(hipMemcpy(total, ticket + 1, 4, hipMemcpyDeviceToHost));
    const uint32_t k = std::min<uint32_t>(*total, out_cap);
    if(k) HIPC(hipMemcpy(out, dout, (size_t)k * 4, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

}  // extern "C"
