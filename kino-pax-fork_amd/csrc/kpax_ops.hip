// kpax_ops.hip — the hot path's operators one at a time (parity entry points
// of include/kpax.h) and the collision-grid build of every environment.
//
// The operators run the same device functions as the planner kernels
// (kpax_device.cuh), one item per lane, on host buffers copied in and out.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "kpax_handle.h"

using namespace kpax;
using namespace kpax::dev;

namespace {

// The operators stage the grid's LDS image like the expansion does when the
// host passes its bytes as dynamic LDS (G.lg.bytes = 0: the CSR grid).
template<int MODEL>
__global__ __launch_bounds__(256) void k_propagate(ModelParams P, GridView G, const float* x0s, const float* us,
                                                   const int32_t* steps, uint32_t n, float* x1s, uint8_t* valid,
                                                   int32_t* fail)
{
    HIP_DYNAMIC_SHARED(uint4, s_grid)
    constexpr int S = Dims<MODEL>::S, C = Dims<MODEL>::C;
    stage_lds_grid(G, s_grid);
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
    HIP_DYNAMIC_SHARED(uint4, s_grid)
    stage_lds_grid(G, s_grid);
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
        const uint32_t pre = tile_prefix(status, tile, agg, 0u, 0u, &s_pre);
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
        const uint32_t pre = tile_prefix(status, tile, agg, 0u, 0u, &s_pre);
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
        lo[d] = cell_of(av[d], gp.mn[d], gp.inv[d], gp.D[d]);
        hi[d] = cell_of(bv[d], gp.mn[d], gp.inv[d], gp.D[d]);
    }
}

// One wave per obstacle: its lanes walk the obstacle's cells (a column of the
// forest spans a few hundred), so no lane runs a long serial loop of atomics.
__device__ __forceinline__ uint32_t obstacle_span(const int a[3], const int b[3], uint32_t* nx, uint32_t* nxy)
{
    *nx  = (uint32_t)(b[0] - a[0] + 1);
    *nxy = *nx * (uint32_t)(b[1] - a[1] + 1);
    return *nxy * (uint32_t)(b[2] - a[2] + 1);
}

__device__ __forceinline__ int span_cell(const GridParams& gp, const int a[3], uint32_t k, uint32_t nx, uint32_t nxy)
{
    const int x = a[0] + (int)(k % nx), y = a[1] + (int)((k % nxy) / nx), z = a[2] + (int)(k / nxy);
    return x + gp.D[0] * (y + gp.D[1] * z);
}

__global__ __launch_bounds__(64) void k_grid_count(GridParams gp, const float4* lo, const float4* hi, int32_t* counts)
{
    const uint32_t o = blockIdx.x;
    int a[3], b[3];
    obstacle_cells(gp, lo[o], hi[o], a, b);
    uint32_t nx, nxy;
    const uint32_t span = obstacle_span(a, b, &nx, &nxy);
    for(uint32_t k = threadIdx.x; k < span; k += 64) atomicAdd(&counts[span_cell(gp, a, k, nx, nxy)], 1);
}

__global__ __launch_bounds__(64) void k_grid_fill(GridParams gp, const float4* lo, const float4* hi,
                                                  const int32_t* start, int32_t* cursor, int32_t* items)
{
    const uint32_t o = blockIdx.x;
    int a[3], b[3];
    obstacle_cells(gp, lo[o], hi[o], a, b);
    uint32_t nx, nxy;
    const uint32_t span = obstacle_span(a, b, &nx, &nxy);
    for(uint32_t k = threadIdx.x; k < span; k += 64)
    {
        const int c = span_cell(gp, a, k, nx, nxy);
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

// The boxes of the sorted CSR items, in item order.
__global__ void k_grid_boxes(const int32_t* items, const float4* lo, const float4* hi, uint32_t K, float4* box)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= K) return;
    const int32_t o = items[k];
    box[2 * (size_t)k]     = lo[o];
    box[2 * (size_t)k + 1] = hi[o];
}

// ---- allocation-free repeat-expansion (kpax_repeat_expand_noalloc) --------
// The look-back status words and the tile ticket of the expansion, for up to
// KPAX_NOALLOC_MAX_ITEMS weights: one array per GPU, held until the process
// ends and shared by every call on that GPU.  In host code `scratch` only
// names the array; its device address comes from hipGetSymbolAddress.
constexpr uint32_t kNoAllocWords = KPAX_NOALLOC_MAX_ITEMS / kTile + 4;
static __device__ __align__(256) uint64_t scratch[kNoAllocWords];

// Zeroes the status words and the ticket a call uses, and its total.
__global__ __launch_bounds__(256) void k_noalloc_clear(uint64_t* words, uint32_t n, uint32_t* total)
{
    for(uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) words[i] = 0;
    if(blockIdx.x == 0 && threadIdx.x == 0) *total = 0;
}

// Environment 0's grid for the operators: with its LDS image unless the
// launch option is off or it has none.
GridView operator_grid(const kpax_handle* h)
{
    GridView G = grid_view(h, 0);
    if(!h->use_lds || !h->envs[0].has_lds) G.lg.bytes = 0;
    return G;
}

}  // namespace

namespace kpax {

// Device address of `scratch` on the current device, looked up once per
// device (the device must be current).
int noalloc_scratch(kpax_handle* h, uint64_t** out)
{
    static std::mutex mu;
    static std::map<int, uint64_t*> addr;
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if(it == addr.end())
    {
        void* p            = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if(e != hipSuccess || !p)
            return fail(h, KPAX_DEVICE_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        it = addr.emplace(dev, (uint64_t*)p).first;
    }
    *out = it->second;
    return KPAX_OK;
}

std::vector<uint4> build_lds_image(const kpax_config& cfg, const float* aabb6, uint32_t n, uint32_t budget,
                                   double max_mean, dev::LdsGrid* lg)
{
    *lg = dev::LdsGrid{};
    if(n > 65535u) return {};
    // candidate dimensions per axis: powers of two up to the CSR grid's
    std::vector<int> cand[3];
    for(int d = 0; d < 3; ++d)
    {
        for(int v = 1; v < cfg.grid_dims[d]; v *= 2) cand[d].push_back(v);
        cand[d].push_back(cfg.grid_dims[d]);
    }
    auto align16 = [](size_t b) { return (b + 15) & ~(size_t)15; };
    float mn[3], ext[3];
    for(int d = 0; d < 3; ++d)
    {
        mn[d]  = cfg.ws_min[d];
        ext[d] = cfg.ws_max[d] - cfg.ws_min[d];
    }
    const size_t box_bytes = align16(24 * (size_t)n);
    // cells each obstacle spans per axis, for every candidate dimension
    std::vector<uint32_t> span[3];
    for(int d = 0; d < 3; ++d)
    {
        span[d].resize(cand[d].size() * (size_t)n);
        for(size_t j = 0; j < cand[d].size(); ++j)
        {
            const float inv = (float)cand[d][j] / ext[d];
            for(uint32_t i = 0; i < n; ++i)
                span[d][j * n + i] = (uint32_t)(dev::cell_of(aabb6[6 * (size_t)i + 3 + d], mn[d], inv, cand[d][j]) -
                                                dev::cell_of(aabb6[6 * (size_t)i + d], mn[d], inv, cand[d][j]) + 1);
        }
    }
    int best[3]       = {0, 0, 0};
    double best_cost  = 0.0;
    size_t best_bytes = 0;
    for(size_t ja = 0; ja < cand[0].size(); ++ja)
        for(size_t jb = 0; jb < cand[1].size(); ++jb)
            for(size_t jc = 0; jc < cand[2].size(); ++jc)
            {
                const int a = cand[0][ja], b = cand[1][jb], c = cand[2][jc];
                const size_t G  = (size_t)a * b * c;
                const size_t lo = box_bytes + align16(2 * (G + 1));
                if(lo > budget) continue;
                const uint32_t* sa = &span[0][ja * n];
                const uint32_t* sb = &span[1][jb * n];
                const uint32_t* sc = &span[2][jc * n];
                size_t items       = 0;
                for(uint32_t i = 0; i < n; ++i) items += (size_t)sa[i] * sb[i] * sc[i];
                const size_t bytes = lo + align16(2 * items);
                if(items > 65535u || bytes > budget) continue;
                // mean obstacles per cell (tests per query), then the smaller image
                const double cost = (double)items / (double)G;
                if(!best[0] || cost < best_cost || (cost == best_cost && bytes < best_bytes))
                {
                    best[0] = a, best[1] = b, best[2] = c;
                    best_cost  = cost;
                    best_bytes = bytes;
                }
            }
    if(!best[0] || best_cost > max_mean) return {};
    dev::LdsGrid L{};
    for(int d = 0; d < 3; ++d)
    {
        L.D[d]   = best[d];
        L.mn[d]  = mn[d];
        L.inv[d] = (float)best[d] / ext[d];
    }
    const size_t G = (size_t)best[0] * best[1] * best[2];
    std::vector<uint32_t> count(G + 1, 0);
    auto cells = [&](uint32_t i, int lo[3], int hi[3]) {
        for(int d = 0; d < 3; ++d)
        {
            lo[d] = dev::cell_of(aabb6[6 * (size_t)i + d], L.mn[d], L.inv[d], L.D[d]);
            hi[d] = dev::cell_of(aabb6[6 * (size_t)i + 3 + d], L.mn[d], L.inv[d], L.D[d]);
        }
    };
    for(uint32_t i = 0; i < n; ++i)
    {
        int lo[3], hi[3];
        cells(i, lo, hi);
        for(int z = lo[2]; z <= hi[2]; ++z)
            for(int y = lo[1]; y <= hi[1]; ++y)
                for(int x = lo[0]; x <= hi[0]; ++x) count[x + (size_t)L.D[0] * (y + (size_t)L.D[1] * z) + 1]++;
    }
    for(size_t g = 0; g < G; ++g) count[g + 1] += count[g];
    const size_t items = count[G];
    L.off_start = (uint32_t)box_bytes;
    L.off_items = (uint32_t)(box_bytes + align16(2 * (G + 1)));
    L.bytes     = (uint32_t)(L.off_items + align16(2 * items));
    std::vector<uint4> img(L.bytes / 16);
    uint8_t* base = (uint8_t*)img.data();
    std::memset(base, 0, L.bytes);
    float* bx = (float*)base;
    for(uint32_t i = 0; i < n; ++i)
        for(int j = 0; j < 6; ++j) bx[6 * (size_t)i + j] = aabb6[6 * (size_t)i + j];
    uint16_t* st = (uint16_t*)(base + L.off_start);
    uint16_t* it = (uint16_t*)(base + L.off_items);
    for(size_t g = 0; g <= G; ++g) st[g] = (uint16_t)count[g];
    std::vector<uint32_t> fill(count.begin(), count.end() - 1);
    for(uint32_t i = 0; i < n; ++i)  // ascending ids per cell
    {
        int lo[3], hi[3];
        cells(i, lo, hi);
        for(int z = lo[2]; z <= hi[2]; ++z)
            for(int y = lo[1]; y <= hi[1]; ++y)
                for(int x = lo[0]; x <= hi[0]; ++x) it[fill[x + (size_t)L.D[0] * (y + (size_t)L.D[1] * z)]++] = (uint16_t)i;
    }
    *lg = L;
    return img;
}

void free_env(Env& e)
{
    if(e.start) (void)hipFree(e.start);
    if(e.items) (void)hipFree(e.items);
    if(e.box) (void)hipFree(e.box);
    if(e.lo) (void)hipFree(e.lo);  // e.hi points into the same buffer
    if(e.occ) (void)hipFree(e.occ);
    if(e.lds_img) (void)hipFree(e.lds_img);
    e = Env{};
}

// A device buffer of at least n items, reallocated only when it must grow.
template<class T>
int ensure(kpax_handle* h, T** p, size_t* cap, size_t n)
{
    n = std::max<size_t>(n, 1);
    if(*p && *cap >= n) return KPAX_OK;
    if(*p) (void)hipFree(*p);
    *p   = nullptr;
    *cap = 0;
    HIPC(hipMalloc((void**)p, n * sizeof(T)));
    *cap = n;
    return KPAX_OK;
}

// Replaces environment `env` by n boxes and builds its CSR grid; the planner
// reads the grid through envs_dev[env].  Runs on the handle's stream after
// everything queued before it.  The environment's buffers and the build's
// scratch are kept from one build to the next and grown when needed; the
// item count is computed on the host with the same cell map, so the build is
// queued without waiting for the device in between.
int build_grid(kpax_handle* h, uint32_t env, const float* aabb6, uint32_t n)
{
    HIPC(hipStreamSynchronize(h->stream));
    Env& e = h->envs[env];
    e.host.assign(aabb6, aabb6 + 6 * (size_t)n);
    e.n_obs = n;
    std::vector<float4> lohi(2 * (size_t)std::max<uint32_t>(n, 1));
    size_t K = 0;
    for(uint32_t i = 0; i < n; ++i)
    {
        const float* o = aabb6 + 6 * (size_t)i;
        lohi[i]        = make_float4(o[0], o[1], o[2], 0.0f);
        lohi[n + i]    = make_float4(o[3], o[4], o[5], 0.0f);
        size_t span    = 1;  // cells of the obstacle, as k_grid_count counts them
        for(int d = 0; d < 3; ++d)
            span *= (size_t)(cell_of(o[3 + d], h->GPm.mn[d], h->GPm.inv[d], h->GPm.D[d]) -
                             cell_of(o[d], h->GPm.mn[d], h->GPm.inv[d], h->GPm.D[d]) + 1);
        K += span;
    }
    if(K > (size_t)INT32_MAX) return fail(h, KPAX_INVALID_ARG, "the obstacles fill more than 2^31 grid cells");
    const uint32_t G = h->G;
    int rc;
    if((rc = ensure(h, &e.lo, &e.lo_cap, 2 * (size_t)std::max<uint32_t>(n, 1))) != KPAX_OK) return rc;
    e.hi = e.lo + std::max<uint32_t>(n, 1);
    if((rc = ensure(h, &e.start, &e.start_cap, (size_t)G + 1)) != KPAX_OK) return rc;
    if((rc = ensure(h, &e.items, &e.items_cap, K)) != KPAX_OK) return rc;
    if((rc = ensure(h, &e.box, &e.box_cap, 2 * K)) != KPAX_OK) return rc;
    HIPC(hipMemcpyAsync(e.lo, lohi.data(), lohi.size() * sizeof(float4), hipMemcpyHostToDevice, h->stream));
    size_t cap = 0;
    if(!h->grid_scratch)
    {
        // counts, cursor (G each), look-back status words, ticket
        const size_t words = 2 * (size_t)G + 2 * ((size_t)G / kTile + 2) + 2;
        if((rc = ensure(h, &h->grid_scratch, &cap, words)) != KPAX_OK) return rc;
    }
    int32_t* counts  = h->grid_scratch;
    int32_t* cursor  = counts + G;
    uint64_t* status = (uint64_t*)(cursor + G);  // 8 G bytes from the base: aligned
    uint32_t* ticket = (uint32_t*)(status + G / kTile + 2);
    HIPC(hipMemsetAsync(counts, 0, (size_t)G * 4, h->stream));
    HIPC(hipMemsetAsync(cursor, 0, (size_t)G * 4, h->stream));
    HIPC(hipMemsetAsync(status, 0, (G / kTile + 2) * 8, h->stream));
    HIPC(hipMemsetAsync(ticket, 0, 4, h->stream));
    if(n)
        hipLaunchKernelGGL(k_grid_count, dim3(n), dim3(64), 0, h->stream, h->GPm, e.lo, e.hi, counts);
    hipLaunchKernelGGL(k_scan_offsets, dim3(std::min<uint32_t>(blocks_for(G, kTile), 4 * h->cu_count)), dim3(256), 0,
                       h->stream, counts, G, e.start, status, ticket);
    e.k = (uint32_t)K;
    if(n)
        hipLaunchKernelGGL(k_grid_fill, dim3(n), dim3(64), 0, h->stream, h->GPm, e.lo, e.hi, e.start, cursor,
                           e.items);
    hipLaunchKernelGGL(k_grid_sort, dim3(blocks_for(G, 256)), dim3(256), 0, h->stream, e.start, e.items, G);
    if(K)
        hipLaunchKernelGGL(k_grid_boxes, dim3(blocks_for(K, 256)), dim3(256), 0, h->stream, e.items, e.lo, e.hi,
                           (uint32_t)K, e.box);
    HIPC(hipGetLastError());
    if((rc = build_occupancy(h, env)) != KPAX_OK) return rc;
    // an image only where its cells list about as few obstacles as the CSR
    // grid's: a coarser grid tests more boxes per query
    const double fine_mean       = (double)K / (double)std::max<uint32_t>(G, 1);
    const std::vector<uint4> img = build_lds_image(h->cfg, aabb6, n, h->lds_budget, 1.5 * fine_mean + 0.05, &e.lg);
    if(!img.empty())
    {
        if((rc = ensure(h, &e.lds_img, &e.lds_cap, img.size())) != KPAX_OK) return rc;
        HIPC(hipMemcpyAsync(e.lds_img, img.data(), img.size() * sizeof(uint4), hipMemcpyHostToDevice, h->stream));
        e.lg.img = e.lds_img;
    }
    e.has_lds = !img.empty();
    if((rc = refresh_lds_launch(h)) != KPAX_OK) return rc;
    const GridView g = grid_view(h, env);
    HIPC(hipMemcpyAsync(h->envs_dev + env, &g, sizeof(g), hipMemcpyHostToDevice, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    return KPAX_OK;
}

}  // namespace kpax

extern "C" {

int kpax_get_grid(kpax_handle* h, int32_t* cell_start, int32_t* items, uint32_t items_cap, uint32_t* n_cells,
                  uint32_t* n_items)
{
    if(!h) return KPAX_INVALID_ARG;
    const Env& e = h->envs[0];
    if(n_cells) *n_cells = h->G;
    if(n_items) *n_items = e.k;
    HIPC(hipSetDevice(h->cfg.device));
    if(cell_start) HIPC(hipMemcpy(cell_start, e.start, ((size_t)h->G + 1) * 4, hipMemcpyDeviceToHost));
    if(items && e.k)
    {
        if(items_cap < e.k) return fail(h, KPAX_INVALID_ARG, "items_cap is smaller than the grid");
        HIPC(hipMemcpy(items, e.items, (size_t)e.k * 4, hipMemcpyDeviceToHost));
    }
    return KPAX_OK;
}

int kpax_get_lds_grid(kpax_handle* h, uint32_t env, int32_t* dims, uint32_t* bytes)
{
    if(!h) return KPAX_INVALID_ARG;
    if(env >= (uint32_t)kMaxEnvs) return fail(h, KPAX_INVALID_ARG, "environment index out of range");
    const Env& e = h->envs[env];
    for(int d = 0; dims && d < 3; ++d) dims[d] = e.has_lds ? e.lg.D[d] : 0;
    if(bytes) *bytes = e.has_lds ? e.lg.bytes : 0;
    return KPAX_OK;
}

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
    float* dx0   = sc.get<float>((size_t)S * n);
    float* du    = sc.get<float>((size_t)C * n);
    int32_t* dst = sc.get<int32_t>(n);
    float* dx1   = sc.get<float>((size_t)S * n);
    uint8_t* dv  = sc.get<uint8_t>(n);
    int32_t* df  = sc.get<int32_t>(n);
    if(!dx0 || !du || !dst || !dx1 || !dv || !df) return fail(h, KPAX_NO_MEMORY, "scratch allocation failed");
    HIPC(hipMemcpy(dx0, x0_soa, (size_t)S * n * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(du, u_soa, (size_t)C * n * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(dst, steps, (size_t)n * 4, hipMemcpyHostToDevice));
    const GridView G  = operator_grid(h);
    const uint32_t b  = blocks_for(n, 256);
    const size_t lds  = G.lg.bytes;
    switch(h->cfg.model)
    {
        case KPAX_MODEL_DOUBLE_INTEGRATOR:
            hipLaunchKernelGGL(k_propagate<KPAX_MODEL_DOUBLE_INTEGRATOR>, dim3(b), dim3(256), lds, h->stream, h->P, G,
                               dx0, du, dst, n, dx1, dv, df);
            break;
        case KPAX_MODEL_DUBINS_AIRPLANE:
            hipLaunchKernelGGL(k_propagate<KPAX_MODEL_DUBINS_AIRPLANE>, dim3(b), dim3(256), lds, h->stream, h->P, G,
                               dx0, du, dst, n, dx1, dv, df);
            break;
        default:
            hipLaunchKernelGGL(k_propagate<KPAX_MODEL_QUADCOPTER>, dim3(b), dim3(256), lds, h->stream, h->P, G, dx0,
                               du, dst, n, dx1, dv, df);
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
    float* lo  = sc.get<float>(3 * (size_t)n);
    float* hi  = sc.get<float>(3 * (size_t)n);
    uint8_t* o = sc.get<uint8_t>(n);
    if(!lo || !hi || !o) return fail(h, KPAX_NO_MEMORY, "scratch allocation failed");
    HIPC(hipMemcpy(lo, bbmin_soa, 12 * (size_t)n, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(hi, bbmax_soa, 12 * (size_t)n, hipMemcpyHostToDevice));
    const GridView G = operator_grid(h);
    hipLaunchKernelGGL(k_collide, dim3(blocks_for(n, 256)), dim3(256), (size_t)G.lg.bytes, h->stream, G, lo, hi, n,
                       o);
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
    HIPC(hipMemsetAsync(rv.r2_state, 0xff, r2 * sizeof(uint64_t), h->stream));
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
    uint32_t* dw          = sc.get<uint32_t>(n);
    uint32_t* dout        = sc.get<uint32_t>(std::max<uint32_t>(out_cap, 1));
    uint64_t* status      = sc.get<uint64_t>(ntiles + 1);
    uint32_t* ticket      = sc.get<uint32_t>(2);
    if(!dw || !dout || !status || !ticket) return fail(h, KPAX_NO_MEMORY, "scratch allocation failed");
    HIPC(hipMemcpy(dw, weights, (size_t)n * 4, hipMemcpyHostToDevice));
    HIPC(hipMemsetAsync(status, 0, (ntiles + 1) * 8, h->stream));
    HIPC(hipMemsetAsync(ticket, 0, 8, h->stream));
    hipLaunchKernelGGL(k_repeat_expand, dim3(std::min<uint32_t>(ntiles, 4 * h->cu_count)), dim3(256), 0, h->stream,
                       dw, n, dout, out_cap, status, ticket, ticket + 1);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    HIPC(hipMemcpy(total, ticket + 1, 4, hipMemcpyDeviceToHost));
    const uint32_t k = std::min<uint32_t>(*total, out_cap);
    if(k) HIPC(hipMemcpy(out, dout, (size_t)k * 4, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_repeat_expand_noalloc(kpax_handle* h, const uint32_t* d_weights, uint32_t n, uint32_t* d_out,
                               uint32_t out_cap, uint32_t* d_total)
{
    if(!h || !d_total || (n && !d_weights) || (out_cap && !d_out)) return fail(h, KPAX_INVALID_ARG, "null argument");
    if(n > KPAX_NOALLOC_MAX_ITEMS)
        return fail(h, KPAX_INVALID_ARG, "n exceeds KPAX_NOALLOC_MAX_ITEMS (" +
                                             std::to_string(KPAX_NOALLOC_MAX_ITEMS) + ") weights");
    HIPC(hipSetDevice(h->cfg.device));
    uint64_t* ptr = h->noalloc;
    if(!ptr)
    {
        const int rc = noalloc_scratch(h, &ptr);
        if(rc != KPAX_OK) return rc;
        h->noalloc = ptr;
    }
    const uint32_t ntiles = (n + kTile - 1) / kTile;
    const uint32_t words  = ntiles + 2;  // status words [0, ntiles], the ticket in word ntiles + 1
    hipLaunchKernelGGL(k_noalloc_clear, dim3(std::min<uint32_t>(blocks_for(words, 256), 64)), dim3(256), 0,
                       h->stream, ptr, words, d_total);
    if(n)
        hipLaunchKernelGGL(k_repeat_expand, dim3(std::min<uint32_t>(ntiles, 4 * h->cu_count)), dim3(256), 0,
                           h->stream, d_weights, n, d_out, out_cap, ptr, (uint32_t*)(ptr + ntiles + 1), d_total);
    HIPC(hipGetLastError());
    return KPAX_OK;
}

int kpax_scan_compact(kpax_handle* h, const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t out_cap,
                      uint32_t* count)
{
    if(!h || !count || (n && !flags) || (out_cap && !out)) return fail(h, KPAX_INVALID_ARG, "null argument");
    std::vector<uint32_t> w(n);
    for(uint32_t i = 0; i < n; ++i) w[i] = flags[i] ? 1u : 0u;
    return kpax_repeat_expand(h, w.data(), n, out, out_cap, count);
}

}  // extern "C"
