// kpax_planner.hip — MI355X kernels of the Kino-PAX planner and the planner
// half of the C ABI (include/kpax.h): single plans and batches of queries.
//
// A handle holds Qmax query slots over one node pool of T nodes.  Each
// admitted query owns a slot and a contiguous node range [base, base + cap)
// of the tree, children and repeat-list arrays (SoA, column d of node i at
// d * T + i), plus its own region counters.  One planner iteration advances
// every running query with a handful of launches on the handle's stream and
// no host round trip between them (the reference makes at least eight
// blocking device->host transfers per iteration of its single query,
// KPAX.cu:303-543):
//   k_prepare   coalesced block 0 first runs the advance: publish each slot's
//                         iteration record, scan the slots' child / tile
//                         counts into offsets; then each child's parent,
//                         Philox control and step draws,
//                         written at its processing position (longest drawn
//                         edge first within each work-group's segment: on by
//                         default for a single-query handle)
//   k_expand    hot       one lane per child: RK4 with checks, region and
//                         sub-region of the result, admission
//                         (propagateFrontier_kernel1/2)
//   k_append    scan      admitted children -> tree in slot order, goal test
//                         (updateFrontier_kernel, first half)
//   k_propose   (learned proposal only) the network on the appended nodes
//   k_score     1 block per slot: region scores (Graph::updateVertices)
//   k_frontier  scan      re-admission + repeat counts -> next repeat list
//                         (updateFrontier_kernel second half + findInd/repeatInd)
// Work of all slots is flattened: children, append tiles and frontier tiles
// are numbered across slots by the advance kernel's offsets, and each item
// finds its slot by binary search.  The advance and the frontier pass write
// the slots' statuses straight into a pinned host buffer (one per iteration
// parity), so an iteration is its launches alone, on the stream directly or
// (launch option) as one captured HIP graph per status parity.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <map>
#include <sstream>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "kpax_handle.h"
#include "kpax_mlp.cuh"

using namespace kpax;
using namespace kpax::dev;

namespace {

// ---------------------------------------------------------------------------
// Slot set-up
// ---------------------------------------------------------------------------

// Clears the region state of the listed slots (fill kernel rather than
// hipMemsetAsync, so every per-plan launch is capturable).
__global__ __launch_bounds__(256) void k_reset_regions(RegionView rv, const uint32_t* slots, uint32_t n, uint32_t R1,
                                                       uint32_t R2)
{
    const uint64_t total = (uint64_t)n * R2;
    for(uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x)
    {
        const uint32_t j = (uint32_t)(i / R2), e = (uint32_t)(i - (uint64_t)j * R2);
        const uint64_t q = slots[j];
        rv.r2_state[q * R2 + e] = ~0ull;
        if(e < R1)
        {
            const uint64_t o = q * R1 + e;
            rv.valid[o]      = 0;
            rv.total[o]      = 0;
            rv.covered[o]    = 0;
            rv.score[o]      = 0.0f;
        }
    }
}

// Root and first iteration record of each listed slot (KPAX.cu:60-88).  One
// block per slot; starts holds 16 floats per listed slot.  The slot's first
// status goes into both parities of the host's status buffer (Qn = Qmax + 1
// words each).
__global__ __launch_bounds__(256) void k_init(Ctl* ctl, const QueryDev* qd, TreeView tree, uint32_t* R,
                                              int32_t* slot_status, uint32_t Qn, const uint32_t* slots,
                                              const float* starts, int S, int C, uint32_t T, RegionParams RP,
                                              PlanParams PP)
{
    const uint32_t q   = slots[blockIdx.x];
    const QueryDev d   = qd[q];
    const float* start = starts + 16 * (size_t)blockIdx.x;
    const size_t b     = d.base;
    const int tid      = threadIdx.x;
    if(tid < S) tree.st[(size_t)tid * T + b] = start[tid];
    if(tid < C) tree.ct[(size_t)tid * T + b] = 0.0f;
    for(uint32_t i = tid; i < PP.root_repeat && i < d.cap; i += blockDim.x) R[b + i] = 0u;
    if(tid == 0)
    {
        tree.steps[b]  = 0;
        tree.parent[b] = -1;
        float p[3]     = {start[0], start[1], start[2]};
        int r1, r2;
        locate(RP, p, &r1, &r2);
        tree.reg[b] = r1;
        Ctl c;
        memset(&c, 0, sizeof(c));
        c.nxt          = make_iter(1u, 1u, PP.root_repeat, d.cap, PP);
        c.cur          = c.nxt;
        c.cur.status   = kNotStarted;  // nothing runs until the advance publishes nxt
        c.goal         = INT_MAX;
        ctl[q]              = c;
        slot_status[q]      = c.nxt.status;
        slot_status[Qn + q] = c.nxt.status;
    }
}

// Everything k_expand reads that is fixed for a launch lives in one
// device-side block behind a single __restrict__ pointer: as by-value kernel
// arguments the structures exceed the scalar register budget and get spilled
// into vector lanes, costing a readlane at every use inside the step loop,
// while values in the block can be re-read with scalar loads whenever the
// compiler needs the registers.  `s0` is slot 0's record, query and grid,
// copied by the advance each iteration for the single-slot form.
struct SlotUniform {
    Iter cur;
    QueryDev qd;
    GridView G;
};

// What the batched expansion's refill needs of a running slot, gathered by
// the advance into one 32-byte record: the iteration, the node range's base
// and the environment's CSR grid, one load instead of the chain slot ->
// query -> environment.
struct SlotRun {
    uint32_t t, base;
    const int32_t* start;
    const float4* box;
    uint32_t pad[2];
};
static_assert(sizeof(SlotRun) == 32, "a slot's run record is two 16-byte words");

struct ExpandArgs {
    Ctl* ctl;
    const QueryDev* qd;
    BatchCtl* bc;
    const uint32_t* child_off;
    const SlotRun* run;  // per slot, written by the advance
    const GridView* envs;
    TreeView tree;
    ChildView ch;
    const uint32_t* R;
    RegionView rv;
    ModelParams P;
    RegionParams RP;
    PlanParams PP;
    GridParams GP;
    uint32_t T, Q, R2;
    uint32_t chunk;  // children per work ticket (multiple of 16)
    uint4* prep;            // per processing position: {child, parent, steps | slot << 17, admit draw}, {controls}
    uint32_t* tickets;      // kTicketGroups chunk counters, kTicketStride words apart
    uint32_t refill;        // idle lanes that trigger a wave's service pass (1..64)
    int32_t prio_q;         // issue-priority step quantum (a quarter of the longest edge); 0: priorities off
    uint32_t ordered;       // children in longest-first order per preparation segment (else child order)
    uint32_t order_blocks;  // work-groups of the preparation pass (its segments)
    const float* prop;
    SlotUniform s0;
};

// ---------------------------------------------------------------------------
// Advance: one thread per slot publishes nxt -> cur; the block scans the
// running slots' children, append tiles and frontier tiles (an upper bound:
// the tree grows by at most M) into offsets, resets the work tickets and
// moves the scan epoch on.
// ---------------------------------------------------------------------------
struct AdvanceArgs {
    Ctl* ctl;
    const QueryDev* qd;
    BatchCtl* bc;
    uint32_t *child_off, *atile_off, *ftile_off;
    int32_t* slot_status;
    uint32_t Q;
    uint64_t *st_append, *st_frontier;
    uint32_t n_status;
    PlanParams PP;
    const GridView* envs;
    SlotUniform* s0;
    uint32_t tile;
    uint32_t* tickets;
    SlotRun* run;
};

// The advance, by block 0 (1024 threads) of the preparation pass.
__device__ __forceinline__ void advance_block(const AdvanceArgs& V)
{
    Ctl* ctl = V.ctl;
    const QueryDev* qd = V.qd;
    BatchCtl* bc = V.bc;
    uint32_t *child_off = V.child_off, *atile_off = V.atile_off, *ftile_off = V.ftile_off;
    int32_t* slot_status = V.slot_status;
    const uint32_t Q = V.Q, n_status = V.n_status, tile = V.tile;
    uint64_t *st_append = V.st_append, *st_frontier = V.st_frontier;
    const PlanParams& PP = V.PP;
    const GridView* envs = V.envs;
    SlotUniform* s0 = V.s0;
    uint32_t* tickets = V.tickets;
    if(threadIdx.x < kTicketGroups) tickets[threadIdx.x * kTicketStride] = 0;
    __shared__ uint32_t lds[16];
    __shared__ uint32_t s_epoch;
    const uint32_t q = threadIdx.x;
    uint32_t m = 0, at = 0, ft = 0, run = 0;
    if(q < Q)
    {
        Ctl* c       = &ctl[q];
        const Iter n = c->nxt;
        c->cur       = n;
        if(q == 0)
        {
            s0->cur = n;
            s0->qd  = qd[0];
            s0->G   = envs[qd[0].env];
        }
        if(n.status == kRunning)
        {
            if(V.run)
            {
                const QueryDev& d = qd[q];
                SlotRun r{};
                r.t     = n.t;
                r.base  = d.base;
                r.start = envs[d.env].start;
                r.box   = envs[d.env].box;
                V.run[q] = r;
            }
            c->expansions += n.M;
            c->Fn       = 0;
            c->f_accept = f_accept(n.t, n.ts, qd[q].cap, PP.eps);  // kept when no child is admitted
            m           = n.M;
            at          = (n.M + tile - 1) / tile;
            ft          = (uint32_t)(((uint64_t)n.ts + n.M + tile - 1) / tile);
            run         = 1;
        }
    }
    uint32_t tm, ta, tf, tr;
    const uint32_t em = block_excl_scan<16>(m, &tm, lds);
    const uint32_t ea = block_excl_scan<16>(at, &ta, lds);
    const uint32_t ef = block_excl_scan<16>(ft, &tf, lds);
    (void)block_excl_scan<16>(run, &tr, lds);
    if(q < Q)
    {
        child_off[q] = em;
        atile_off[q] = ea;
        ftile_off[q] = ef;
    }
    if(q == 0)
    {
        child_off[Q]     = tm;
        atile_off[Q]     = ta;
        ftile_off[Q]     = tf;
        bc->children     = tm;
        bc->atiles       = ta;
        bc->ftiles       = tf;
        bc->running      = tr;
        bc->tile_append  = 0;
        bc->tile_frontier = 0;
        uint32_t e       = bc->epoch + 1;
        if(e > kEpochMask) e = 1;
        bc->epoch      = e;
        s_epoch        = e;
        slot_status[Q] = (int32_t)tr;
    }
    __syncthreads();
    // epoch 1 starts a new cycle of epochs: words left from the last cycle
    // could carry the same epoch, so the status arrays are cleared once
    if(s_epoch == 1)
        for(uint32_t i = threadIdx.x; i < n_status; i += blockDim.x)
        {
            st_append[i]   = 0;
            st_frontier[i] = 0;
        }
}


// ---------------------------------------------------------------------------
// Processing order of the children, longest drawn edge first, so the last
// children a launch starts are short ones and the launch does not wait on a
// long edge begun at its end.  Results do not depend on the order (every
// child is a function of its index and of the tree before the iteration).
// The preparation pass's work-group b takes segment b of the child index
// space, [b per, (b + 1) per) with per = ceil(children / work-groups), and
// sorts it in place by drawn step count (a counting sort in LDS); the
// expansion's chunk k is the (k / segments)-th chunk of segment
// k % segments, so the launch starts every segment's longest children first,
// then their next longest, ...: the global longest-first order to within a
// chunk, with no pass that counts the buckets of all segments first.  In
// child order there is one segment and the position is the child index.
// ---------------------------------------------------------------------------
// 256 work-groups of 1024 threads: four waves per SIMD hide the Philox chains
// and the repeat-list loads (one wave per SIMD left both exposed).
constexpr int kOrderBlocks = 256, kOrderThreads = 1024, kOrderBuckets = 32;

__device__ __forceinline__ uint32_t order_bucket(const ModelParams& P, int n)
{
    const uint32_t back = (uint32_t)(P.steps_min + P.span - 1 - n);  // 0 for the longest edge
    return P.span <= kOrderBuckets ? back : (uint32_t)((uint64_t)back * kOrderBuckets / (uint32_t)P.span);
}

// Children per segment: the children over the preparation pass's `nb`
// work-groups for the longest-first order, but at least kMinSegment (a small
// iteration's segments would otherwise hold a child or two each, and a wave
// would take a ticket per child); one segment in child order.
constexpr uint32_t kMinSegment = 1024;
__device__ __forceinline__ uint32_t segment_len(uint32_t total, uint32_t nb, bool ordered)
{
    if(!ordered) return total;
    const uint32_t per = (total + nb - 1) / nb;
    return per > kMinSegment ? per : kMinSegment;
}

// Positions [*b, *e) of work chunk k (empty when k lands past the end of the
// last, shorter segment).  per >= 1, nseg = ceil(total / per).  Row j holds
// the j-th chunk of every segment, rotated by j: a ticket group takes every
// 64th chunk index, and without the rotation it would keep to the same few
// segments, whose children (consecutive parents, often in one part of the
// workspace) can run long or short together.
__device__ __forceinline__ void chunk_range(uint64_t k, uint32_t total, uint32_t per, uint32_t nseg, uint32_t chunk,
                                            uint32_t* b, uint32_t* e)
{
    const uint64_t j = k / nseg, s = (k % nseg + j) % nseg;
    const uint64_t s0 = s * per, lo = s0 + j * chunk;
    uint64_t hi = s0 + (j + 1) * chunk;
    if(hi > s0 + per) hi = s0 + per;
    if(hi > total) hi = total;
    *b = (uint32_t)(lo < total ? lo : total);
    *e = (uint32_t)(hi > *b ? hi : *b);
}

// Issue priority of a wave from the steps its longest live edge has left, in
// quarters q of the longest drawn edge (0..3).  A launch lasts as long as its
// longest chain of steps; the waves holding the long edges issue first and
// the waves with short edges fill their stalls, instead of every wave
// sharing the SIMD evenly until the long edges are all that is left.
__device__ __forceinline__ void set_issue_priority(int left, int q)
{
    if(__ballot(left > 3 * q))
        __builtin_amdgcn_s_setprio(3);
    else if(__ballot(left > 2 * q))
        __builtin_amdgcn_s_setprio(2);
    else if(__ballot(left > q))
        __builtin_amdgcn_s_setprio(1);
    else
        __builtin_amdgcn_s_setprio(0);
}

// ---------------------------------------------------------------------------
// Expansion: propagateFrontier_kernel1/2 (KPAX.cu:358-439) with counter-based
// draws.  Waves are persistent and advance their 64 children one integration
// step per trip through the loop; a lane whose child has finished (all its
// steps, or an invalid step) writes the child out and takes the next one
// from the wave's chunk of children, and the wave takes a new chunk from a
// global ticket when its chunk runs out.  Children draw 5..20 steps and many
// stop early, so refilling keeps the lanes of a wave busy instead of waiting
// for the longest edge.  Each child's arithmetic is that of propagate().
//
// MULTI = false: one slot (slot 0); its record, seed and grid are uniform.
// MULTI = true: each lane looks up the slot of its child (children of all
// slots are numbered consecutively) and carries that slot's iteration, node
// range and grid pointers.  Every child's R1 region and integrated step
// count go out with it; k_append sums them per query, so the expansion makes
// no counter atomics.
// ---------------------------------------------------------------------------
// Five resident waves per SIMD (96 VGPRs for the quadcopter) hide more of
// the grid lookups' latency than four; the batched form carries a slot's
// state per lane and keeps four.  The single-slot form stages its
// environment's grid image in LDS (work-groups of 640 threads: two per CU
// hold the five waves per SIMD and two images); the batched form's lanes
// query the grids of several environments in global memory (256 threads).
// GQ: the grid query path (kGridLds: the single-slot form with its
// environment's image staged in LDS; kGridCsr: the CSR grid in HBM).
template<int MODEL, bool MULTI, int GQ>
__global__ __launch_bounds__(MULTI ? 256 : 640) __attribute__((amdgpu_waves_per_eu(MULTI ? 4 : 5))) void k_expand(
  const ExpandArgs* __restrict__ A)
{
    HIP_DYNAMIC_SHARED(uint4, s_grid)
    const TreeView& tree   = A->tree;
    const ChildView& ch    = A->ch;
    const RegionView& rv   = A->rv;
    const ModelParams& P   = A->P;
    const RegionParams& RP = A->RP;
    const uint32_t T       = A->T;
    constexpr int S = Dims<MODEL>::S, C = Dims<MODEL>::C;
    constexpr uint32_t kNone = 0xffffffffu;
    const uint32_t total = MULTI ? A->bc->children : (A->s0.cur.status == kRunning ? A->s0.cur.M : 0u);
    if(total == 0) return;
    // Chunks of children are handed out by kTicketGroups counters, each on
    // its own cache line (one device-scope counter for all waves serialised
    // the launch): the blocks of group g (block index mod the group count, so
    // one XCD of the round-robin dispatch feeds each counter) take chunks
    // g, g + groups, g + 2 groups, ...  A wave whose first chunk would
    // already lie past the end leaves without taking a ticket: in a small
    // iteration thousands of persistent waves would otherwise queue on the
    // counters before the kernel could end.  The waves that stay still take
    // every chunk of their group, and every group has blocks.
    const uint32_t groups  = gridDim.x < (uint32_t)kTicketGroups ? gridDim.x : (uint32_t)kTicketGroups;
    const uint32_t group   = blockIdx.x % groups;
    const uint32_t rank0   = (blockIdx.x / groups) * (blockDim.x >> 6);  // the work-group's first wave
    const uint32_t rank    = rank0 + (threadIdx.x >> 6);
    // work chunks: segments of the preparation pass, interleaved chunk by chunk
    const uint32_t per    = segment_len(total, A->order_blocks, A->ordered != 0);
    const uint32_t nseg   = (total + per - 1) / per;
    const uint64_t chunks = (uint64_t)nseg * ((per + A->chunk - 1) / A->chunk);
    // a work-group none of whose waves has a first chunk leaves at once
    // (before the LDS staging, whose barrier needs every thread)
    if((uint64_t)rank0 * groups + group >= chunks) return;
    const int R1 = RP.r1_count;
    // slot-uniform in the single-slot form, per lane in the batched form
    uint32_t t = 0, slot = 0, base = 0;
    GridView G;
    if constexpr(!MULTI)
    {
        t    = A->s0.cur.t;
        base = A->s0.qd.base;
        G    = A->s0.G;
        if constexpr(GQ == kGridLds)
            stage_lds_grid(G, s_grid);
        else
            G.lds = nullptr;
    }
    else
    {
        G.gp  = A->GP;  // grid parameters are shared; pointers are set per child
        G.lds = nullptr;
    }
    if((uint64_t)rank * groups + group >= chunks) return;
    uint32_t* const ticket = A->tickets + group * kTicketStride;
    const int lane       = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t c = kNone, parent = 0, admit_draw = 0;
    int i = 0, n = 0;
    float x[S], u[C];
    uint32_t q_next = 0, q_end = 0;  // the wave's current chunk [q_next, q_end)
    bool drained = false;
    bool live = false, pend = false, valid = false;  // integrating / finished, not yet written / its verdict
    int cells[3] = {0, 0, 0};  // grid cells of the current step's start point
    const uint32_t refill = A->refill;
    const int prio_q      = A->prio_q;
    for(;;)
    {
        // Service pass, once `refill` lanes are idle (or all are): write out the
        // finished children, then hand the idle lanes new ones.  Deferring it
        // lets the wave run the write-out and the draws for several lanes at
        // once instead of on nearly every trip.  Once the chunks are drained
        // no lane can be refilled, so the finished children wait for one
        // write-out when the last lane is done: the drain's trips then carry
        // no service passes, which sit on the launch's critical path.
        const uint64_t idle = __ballot(!live);
        if(idle == ~0ull || (!drained && (uint32_t)__popcll(idle) >= refill))
        {
            if(pend)
            {
                int r1, r2;
                locate(RP, x, &r1, &r2);
                const size_t ro = MULTI ? (size_t)slot * R1 : 0;
                uint32_t flags  = 0;
                if(valid)
                {
                    flags = kValid | claim_cell(rv.r2_state + (MULTI ? (size_t)slot * A->R2 : 0), r2, t, c);
                    if(u01(admit_draw) < rv.score[ro + r1]) flags |= kScoreAdmit;
                }
                const uint32_t o = base + c;
                uint32_t w[kChildWords];
#pragma unroll
                for(int d = 0; d < 12; ++d) w[d] = d < S ? __builtin_bit_cast(uint32_t, x[d < S ? d : 0]) : 0u;
#pragma unroll
                for(int d = 0; d < 4; ++d) w[12 + d] = d < C ? __builtin_bit_cast(uint32_t, u[d < C ? d : 0]) : 0u;
                uint4* rec = ch.rec + (size_t)o * (kChildWords / 4);
#pragma unroll
                for(int v = 0; v < kChildWords / 4; ++v)
                    rec[v] = make_uint4(w[4 * v], w[4 * v + 1], w[4 * v + 2], w[4 * v + 3]);
                // steps integrated: n for a valid child
                ch.meta[o] = make_uint4((uint32_t)i | (flags << kMetaFlagShift), parent, (uint32_t)r1, (uint32_t)r2);
                c    = kNone;
                pend = false;
            }
            if(!drained)
            {
                uint64_t need = __ballot(c == kNone);
                while(need)
                {
                    if(q_next >= q_end)
                    {
                        // the next non-empty chunk of the group (chunks past the
                        // end of the last, shorter segment are empty)
                        uint32_t b = total, e = total;
                        if(lane == 0)
                            for(;;)
                            {
                                const uint64_t k = (uint64_t)atomicAdd(ticket, 1u) * groups + group;
                                if(k >= chunks)
                                {
                                    b = e = total;
                                    break;
                                }
                                chunk_range(k, total, per, nseg, A->chunk, &b, &e);
                                if(b < e) break;
                            }
                        b = __shfl(b, 0, 64);
                        e = __shfl(e, 0, 64);
                        if(b >= total)
                        {
                            drained = true;
                            break;
                        }
                        q_next = b;
                        q_end  = e;
                    }
                    const uint32_t avail = q_end - q_next;
                    const uint32_t rank  = (uint32_t)__popcll(need & below);
                    if(c == kNone && rank < avail)
                    {
                        // the child's draws and parent (its node index in the
                        // pool: the state loads need no slot record), prepared
                        // by k_prepare
                        const uint4* pr = A->prep + 2 * (size_t)(q_next + rank);
                        const uint4 r0 = pr[0], r1 = pr[1];
                        const uint32_t pn = r0.y;
#pragma unroll
                        for(int d = 0; d < S; ++d) x[d] = tree.st[(size_t)d * T + pn];
                        if constexpr(MULTI)
                        {
                            slot            = r0.z >> kPrepSlotShift;
                            const SlotRun r = A->run[slot];
                            t               = r.t;
                            base            = r.base;
                            G.start         = r.start;
                            G.box           = r.box;
                        }
                        c                 = r0.x;  // index within the slot's children
                        parent            = pn - base;
                        const uint32_t uw[4] = {r1.x, r1.y, r1.z, r1.w};
#pragma unroll
                        for(int j = 0; j < C; ++j) u[j] = __builtin_bit_cast(float, uw[j]);
                        n          = (int)(r0.z & ((1u << kPrepSlotShift) - 1u));
                        i          = 0;
                        admit_draw = r0.w;
                        live       = true;
                        if constexpr(GQ == kGridLds)
                            lds_cells(G, x, cells);
                        else
                            csr_cells(G, x, cells);
                    }
                    const uint32_t wanted = (uint32_t)__popcll(need);
                    q_next += wanted < avail ? wanted : avail;
                    need = __ballot(c == kNone);
                }
            }
            if(__ballot(live) == 0) break;  // drained, and every child written out
        }
        if(prio_q) set_issue_priority(live ? n - i : 0, prio_q);
        if(live)
        {
            valid = step_carry<MODEL, GQ>(P, G, x, u, cells);
            ++i;
            if(!valid || i == n)
            {
                live = false;
                pend = true;
            }
        }
    }
}

// Preparation of the iteration's children, in a coalesced pass before the
// expansion: each child's parent (repeat list), control draw (uniform, or the
// parent's Gaussian with the query's proposal weight), step count and
// admission draw, stored as a 32-byte record at its processing position, so
// the expansion's refill is two vector loads and the parent state.
//
// advance = 1: block 0 first runs the iteration's advance (no launch of its
// own), and every block takes the iteration records from the slots' `nxt`,
// which this launch does not change, instead of `cur` and the argument
// block's copies, which block 0 rewrites; in the batched form each block also
// scans the slots' child counts itself, since the offsets block 0 publishes
// are only for the passes after this one.
template<int MODEL, bool MULTI>
__global__ __launch_bounds__(kOrderThreads) void k_prepare(const ExpandArgs* __restrict__ A, int ordered,
                                                           AdvanceArgs V, int advance)
{
    constexpr int C = Dims<MODEL>::C;
    __shared__ uint32_t pos[kOrderBuckets];
    // the batched form's child offsets per slot (with the advance: this
    // block's own scan of the slots' next iterations)
    __shared__ uint32_t s_off[MULTI ? kMaxSlots + 1 : 1];
    __shared__ uint32_t s_scan[kOrderThreads / 64];
    if(advance && blockIdx.x == 0) advance_block(V);
    uint32_t total;
    Iter it0{};
    QueryDev q0{};
    const uint32_t* child_off = A->child_off;
    if constexpr(MULTI)
    {
        if(advance)
        {
            // the advance publishes the same offsets, but only for the passes
            // after this one
            const uint32_t q = threadIdx.x;
            uint32_t m       = 0;
            if(q < A->Q)
            {
                const Iter n = A->ctl[q].nxt;
                if(n.status == kRunning) m = n.M;
            }
            const uint32_t em = block_excl_scan<kOrderThreads / 64>(m, &total, s_scan);
            if(q < A->Q) s_off[q] = em;
            if(q == 0) s_off[A->Q] = total;
            __syncthreads();
            child_off = s_off;
        }
        else
            total = A->bc->children;
    }
    else
    {
        it0   = advance ? A->ctl[0].nxt : A->s0.cur;
        q0    = advance ? A->qd[0] : A->s0.qd;
        total = it0.status == kRunning ? it0.M : 0u;
    }
    const ModelParams& P = A->P;
    // ordered: this work-group's segment; child order: an even share
    const uint32_t per   = ordered ? segment_len(total, gridDim.x, true) : (total + gridDim.x - 1) / gridDim.x;
    const uint32_t b0 = blockIdx.x * per, b1 = b0 + per < total ? b0 + per : total;
    // the draws of child c (index c - child_off[slot] within its slot)
    auto draws = [&](uint32_t k, uint32_t* c, uint32_t* t, uint32_t* slot, uint64_t* seed, uint32_t* query) {
        if constexpr(MULTI)
        {
            *slot  = segment_of(child_off, A->Q, k);
            *c     = k - child_off[*slot];
            *t     = advance ? A->ctl[*slot].nxt.t : A->ctl[*slot].cur.t;
            *seed  = A->qd[*slot].seed;
            *query = A->qd[*slot].query;
        }
        else
        {
            *slot  = 0;
            *c     = k;
            *t     = it0.t;
            *seed  = q0.seed;
            *query = q0.query;
        }
    };
    if(ordered)
    {
        // counting sort of the segment: bucket counts, then their exclusive
        // scan (longest bucket first) from the segment's start
        if(threadIdx.x < kOrderBuckets) pos[threadIdx.x] = 0;
        __syncthreads();
        for(uint32_t k = b0 + threadIdx.x; k < b1; k += blockDim.x)
        {
            uint32_t c, t, slot, query;
            uint64_t seed;
            draws(k, &c, &t, &slot, &seed, &query);
            const uint4 ds = philox(c, t, query, PURPOSE_STEPS_ADMIT, (uint32_t)seed, (uint32_t)(seed >> 32));
            atomicAdd(&pos[order_bucket(P, draw_steps(P, ds.x))], 1u);
        }
        __syncthreads();
        if(threadIdx.x == 0)
        {
            uint32_t run = b0;
            for(int b = 0; b < kOrderBuckets; ++b)
            {
                const uint32_t n = pos[b];
                pos[b]           = run;
                run += n;
            }
        }
        __syncthreads();
    }
    for(uint32_t k = b0 + threadIdx.x; k < b1; k += blockDim.x)
    {
        uint32_t c, t, slot, query, fan, base;
        uint64_t seed;
        float mix;
        draws(k, &c, &t, &slot, &seed, &query);
        if constexpr(MULTI)
        {
            fan  = advance ? A->ctl[slot].nxt.fan : A->ctl[slot].cur.fan;
            base = A->qd[slot].base;
            mix  = A->qd[slot].mix;
        }
        else
        {
            fan  = it0.fan;
            base = q0.base;
            mix  = q0.mix;
        }
        const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
        const uint32_t parent = A->R[base + c / fan];
        const uint4 dc        = philox(c, t, query, PURPOSE_CONTROL, k0, k1);
        const uint4 ds        = philox(c, t, query, PURPOSE_STEPS_ADMIT, k0, k1);
        const uint32_t dcw[4] = {dc.x, dc.y, dc.z, dc.w};
        float u[C];
#pragma unroll
        for(int j = 0; j < C; ++j) u[j] = P.ctrl_min[j] + u01(dcw[j]) * P.ctrl_rng[j];
        if(mix > 0.0f && u01(ds.z) < mix)
        {
            // the parent's Gaussian instead (learned proposal)
            const size_t pn = base + parent;
            const uint4 dp  = philox(c, t, query, PURPOSE_PROPOSAL, k0, k1);
            float mean[C], sg[C];
#pragma unroll
            for(int j = 0; j < C; ++j)
            {
                mean[j] = A->prop[(size_t)j * A->T + pn];
                sg[j]   = A->prop[(size_t)(4 + j) * A->T + pn];
            }
            propose_control<C>(mean, sg, dp, P, u);
        }
        const int n      = draw_steps(P, ds.x);
        const uint32_t p = ordered ? atomicAdd(&pos[order_bucket(P, n)], 1u) : k;
        uint32_t uw[4]   = {0u, 0u, 0u, 0u};
#pragma unroll
        for(int j = 0; j < C; ++j) uw[j] = __builtin_bit_cast(uint32_t, u[j]);
        uint4* pr = A->prep + 2 * (size_t)p;
        pr[0]     = make_uint4(c, base + parent, (uint32_t)n | (slot << kPrepSlotShift), ds.y);
        pr[1]     = make_uint4(uw[0], uw[1], uw[2], uw[3]);
    }
}

// ---------------------------------------------------------------------------
// Region scores (Graph.cu:233-358, the evident intent): raw Syclop score per
// R1 region, their sum in a fixed order (1024 strided partial sums, then a
// pairwise tree), normalised eps + score / total, or 1 for regions with no
// valid child.  One block of 1024 threads per slot; with ctl == nullptr one
// block scores the arrays as given (parity operator).
// ---------------------------------------------------------------------------
// The scores of one slot's regions by one block of NT threads (NT divides
// 1024): raw scores into sc (LDS or global), the 1024 strided partial sums
// and their pairwise tree in part (LDS, 1024 floats), then the normalised
// scores in place.  Returns the total.  Contains barriers.
template<int NT>
__device__ __forceinline__ float block_scores(const int32_t* va, const int32_t* to, const int32_t* co, float* sc,
                                              float* part, const RegionParams& RP, float eps)
{
    const int t = threadIdx.x;
    const int R1 = RP.r1_count;
    for(int pi = t; pi < 1024; pi += NT)
    {
        float s = 0.0f;
        // four regions per round, their counters loaded together (clamped
        // indices: the loads are unconditional, so they issue as one round trip)
        for(int r0 = pi; r0 < R1; r0 += 4 * 1024)
        {
            int v[4], n0[4], cv[4];
#pragma unroll
            for(int k = 0; k < 4; ++k)
            {
                const int r  = r0 + 1024 * k;
                const int rr = r < R1 ? r : R1 - 1;
                v[k]         = va[rr];
                n0[k]        = to[rr];
                cv[k]        = co[rr];
            }
#pragma unroll
            for(int k = 0; k < 4; ++k)
            {
                const int r = r0 + 1024 * k;
                if(r >= R1) break;
                float raw = 0.0f;
                if(v[k] > 0)
                {
                    const float cov = (float)cv[k] / (float)RP.r2_per;
                    const float ev  = eps + (float)v[k];
                    const float fv  = ev / (ev + (float)(n0[k] - v[k])) * RP.vol;
                    const float n   = (float)n0[k];
                    raw             = (fv * fv) * (fv * fv) / ((1.0f + cov) * (1.0f + n * n));
                }
                sc[r] = raw;
                s += raw;
            }
        }
        part[pi] = s;
    }
    __syncthreads();
    // the pairwise tree part[i] += part[i + off], off = 512 .. 1: the levels
    // down to 64 partials across the block, the last six within wave 0 by
    // shuffles (the same additions in the same order, without barriers)
    for(int off = 512; off >= 64; off >>= 1)
    {
        for(int pi = t; pi < off; pi += NT) part[pi] = part[pi] + part[pi + off];
        __syncthreads();
    }
    if(t < 64)
    {
        float x = part[t];
#pragma unroll
        for(int off = 32; off >= 1; off >>= 1)
        {
            const float y = __shfl_down(x, off, 64);
            if(t < off) x = x + y;
        }
        if(t == 0) part[0] = x;
    }
    __syncthreads();
    const float tot = part[0];
    for(int r = t; r < R1; r += NT) sc[r] = va[r] == 0 ? 1.0f : eps + sc[r] / tot;
    __syncthreads();
    return tot;
}

__global__ __launch_bounds__(1024) void k_score(Ctl* ctl, RegionView rv, RegionParams RP, float eps)
{
    __shared__ float part[1024];
    const uint32_t q = blockIdx.x;
    if(ctl && ctl[q].cur.status != kRunning) return;
    const size_t o  = (size_t)q * RP.r1_count;
    const float tot = block_scores<1024>(rv.valid + o, rv.total + o, rv.covered + o, rv.score + o, part, RP, eps);
    if(threadIdx.x == 0 && ctl) ctl[q].score_total = tot;
}

// ---------------------------------------------------------------------------
// Append: the region counters of the iteration's children (total for every
// child, at its failing state if invalid; valid for valid ones:
// KPAX.cu:386-390) and their integrated steps, compaction of the admitted children into the tree in slot
// order and the goal test (updateFrontier_kernel, KPAX.cu:458-502).  Tiles
// of all slots are numbered consecutively; each slot's tiles form one
// look-back segment.  A tile's children belong to one slot, so its counters
// are summed in LDS (total in [0, R1), valid in [R1, 2 R1)) and only the
// regions it touched reach HBM: siblings land in the same few regions, and
// per-child device-scope atomics would serialise on them.
// ---------------------------------------------------------------------------
// The tile's critical path is a chain of dependent memory round trips, so
// the work that does not depend on the scan runs while it waits: the tile's
// ticket is taken while the LDS counters are cleared; the first admitted
// child's record is loaded before the scan; the aggregate is published
// before the counters are flushed, and the flush (fire-and-forget atomics)
// runs before the wait on the predecessor tiles.
template<int NT>
__global__ __launch_bounds__(NT) void k_append(Ctl* ctl, const QueryDev* qd, BatchCtl* bc, const uint32_t* atile_off,
                                               uint32_t Q, ChildView ch, TreeView tree, RegionView rv,
                                               uint64_t* status, PlanParams PP, int S, int C, uint32_t T,
                                               uint32_t R1, uint32_t R2)
{
    constexpr uint32_t kT = 4 * NT;  // children per tile
    __shared__ uint32_t lds[2 * (NT / 64)];
    __shared__ uint32_t s_tile, s_pre;
    HIP_DYNAMIC_SHARED(int32_t, s_hist)
    const uint32_t ntiles = bc->atiles, epoch = bc->epoch;
    if(blockIdx.x >= ntiles) return;  // no ticket to take: leave the counter to the blocks with work
    const bool local      = R1 <= (uint32_t)kLdsRegions;
    for(;;)
    {
        // the ticket's round trip overlaps the clearing of the counters (the
        // previous tile's flush ended before its look-back's barriers)
        if(threadIdx.x == 0) s_tile = atomicAdd(&bc->tile_append, 1u);
        if(local)
            for(uint32_t r = threadIdx.x; r < 2 * R1; r += blockDim.x) s_hist[r] = 0;
        __syncthreads();
        const uint32_t tile = s_tile;
        if(tile >= ntiles) break;
        const uint32_t q      = segment_of(atile_off, Q, tile);
        const uint32_t first  = atile_off[q], lt = tile - first, nt = atile_off[q + 1] - first;
        const Iter it         = ctl[q].cur;
        const QueryDev d      = qd[q];
        const uint32_t M      = it.M, ts = it.ts, cap = d.cap;
        const size_t base     = d.base;
        const uint64_t* r2s   = rv.r2_state + (size_t)q * R2;
        int32_t* cov          = rv.covered + (size_t)q * R1;
        int32_t* tot          = rv.total + (size_t)q * R1;
        int32_t* val          = rv.valid + (size_t)q * R1;
        const uint32_t cbase  = lt * kT + threadIdx.x * 4;
        uint32_t f[4], cnt = 0, steps = 0;
        // The four meta words, then the cell words of the candidates, each as
        // one round trip: the loads are unconditional (past the last child,
        // the last child's word; for a non-candidate, cell 0's word) so they
        // issue together instead of one wait per child.
        uint4 mj[4];  // the children's meta words: steps | flags, parent, region, R2 cell
        uint64_t key[4];
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            const uint32_t c = cbase + j;
            mj[j]            = ch.meta[base + (c < M ? c : M - 1)];
        }
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            const bool cand = cbase + j < M && ((mj[j].x >> kMetaFlagShift) & kCandidate);
            key[j]          = r2s[cand ? mj[j].w : 0u];
        }
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            const uint32_t c    = cbase + j;
            if(c >= M) mj[j] = make_uint4(0u, 0u, 0u, 0u);
            const uint4 meta    = mj[j];
            const uint32_t a    = meta.x >> kMetaFlagShift;
            if(c < M)
            {
                steps += meta.x & ((1u << kMetaFlagShift) - 1u);
                const int r = (int)meta.z;
                if(local)
                {
                    atomicAdd(&s_hist[r], 1);
                    if(a & kValid) atomicAdd(&s_hist[R1 + r], 1);
                }
                else
                {
                    atomicAdd(&tot[r], 1);
                    if(a & kValid) atomicAdd(&val[r], 1);
                }
            }
            // a candidate that holds its cell's key was the first to reach it:
            // it covers the cell and is admitted (KPAX.cu:394-395)
            bool first_in_cell = false;
            if(a & kCandidate)
            {
                first_in_cell = key[j] == cell_key(it.t, c);
                if(first_in_cell) atomicAdd(&cov[meta.z], 1);
            }
            f[j] = (a & kValid) && ((a & kScoreAdmit) || first_in_cell) ? 1u : 0u;
            if(f[j] && PP.prune)
            {
                // goal-progress pruning (PruneKPAX): decided before the scan, so
                // the kept children are appended without holes
                const float4 cxyz = __builtin_bit_cast(float4, ch.rec[(base + c) * (kChildWords / 4)]);
                const size_t pn   = base + meta.y;
                const float child[3]  = {cxyz.x, cxyz.y, cxyz.z};
                const float parent[3] = {tree.st[pn], tree.st[(size_t)T + pn], tree.st[2 * (size_t)T + pn]};
                const uint4 dp        = philox(c, it.t, d.query, PURPOSE_PRUNE, (uint32_t)d.seed,
                                               (uint32_t)(d.seed >> 32));
                f[j] = prune_keep(PP, child, parent, d.goal, dp.x) ? 1u : 0u;
            }
            cnt += f[j];
        }
        // the first admitted child's record, loaded while the tile scans
        int jf = -1;
#pragma unroll
        for(int j = 3; j >= 0; --j)
            if(f[j]) jf = j;
        uint4 rf[kChildWords / 4];
        if(jf >= 0)
        {
            const uint4* rec = ch.rec + (base + cbase + (uint32_t)jf) * (kChildWords / 4);
#pragma unroll
            for(int v = 0; v < kChildWords / 4; ++v) rf[v] = rec[v];
        }
        uint32_t agg, steps_tile;
        const uint32_t loc = block_excl_scan_sum<NT / 64>(cnt, steps, &agg, &steps_tile, lds);
        if(threadIdx.x == 0)
        {
            tile_publish(status, tile, agg, first, epoch);
            if(steps_tile) atomicAdd((unsigned long long*)&ctl[q].steps_done, (unsigned long long)steps_tile);
        }
        if(local)
            for(uint32_t r = threadIdx.x; r < R1; r += blockDim.x)
            {
                const int32_t a = s_hist[r], b = s_hist[R1 + r];
                if(a) atomicAdd(&tot[r], a);
                if(b) atomicAdd(&val[r], b);
            }
        const uint32_t pre = tile_prefix_published(status, tile, agg, first, epoch, &s_pre);
        uint32_t dst       = ts + pre + loc;
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            if(!f[j]) continue;
            const size_t c = base + cbase + j;
            if(dst >= cap)
            {
                ctl[q].overflow = 1;  // cannot happen while M <= cap - ts; never write past the range
                break;
            }
            const size_t o = base + dst;
            uint4 r4[kChildWords / 4];
            if(j == jf)
            {
#pragma unroll
                for(int v = 0; v < kChildWords / 4; ++v) r4[v] = rf[v];
            }
            else
            {
                const uint4* rec = ch.rec + c * (kChildWords / 4);
#pragma unroll
                for(int v = 0; v < kChildWords / 4; ++v) r4[v] = rec[v];
            }
            const uint4 m4 = mj[j];
            uint32_t w[kChildWords];
#pragma unroll
            for(int v = 0; v < kChildWords / 4; ++v)
            {
                w[4 * v]     = r4[v].x;
                w[4 * v + 1] = r4[v].y;
                w[4 * v + 2] = r4[v].z;
                w[4 * v + 3] = r4[v].w;
            }
            for(int dd = 0; dd < S; ++dd) tree.st[(size_t)dd * T + o] = __builtin_bit_cast(float, w[dd]);
            for(int dd = 0; dd < C; ++dd) tree.ct[(size_t)dd * T + o] = __builtin_bit_cast(float, w[12 + dd]);
            tree.steps[o]  = (int32_t)(m4.x & ((1u << kMetaFlagShift) - 1u));
            tree.parent[o] = (int32_t)m4.y;
            tree.reg[o]    = (int32_t)m4.z;
            const float dx = __builtin_bit_cast(float, w[0]) - d.goal[0],
                        dy = __builtin_bit_cast(float, w[1]) - d.goal[1],
                        dz = __builtin_bit_cast(float, w[2]) - d.goal[2];
            if(__builtin_sqrtf(dx * dx + dy * dy + dz * dz) < PP.goal_radius) atomicMin(&ctl[q].goal, (int32_t)dst);
            ++dst;
        }
        if(lt == nt - 1 && threadIdx.x == 0)
        {
            const uint32_t Fn = pre + agg;
            ctl[q].Fn         = Fn;
            ctl[q].f_accept   = f_accept(it.t, ts + Fn, cap, PP.eps);
        }
    }
}

// ---------------------------------------------------------------------------
// Next frontier: weights of all tree nodes (re-admission of old nodes with
// probability score + fAccept, KPAX.cu:505-517; repeat count 15 / 1 of new
// nodes by their region's valid count, KPAX.cu:474-479), scanned and
// run-length expanded into the repeat list (Planner.cu:52-80).  A slot whose
// goal was reached ends here instead.
// ---------------------------------------------------------------------------
// SCORE (the single-slot form, R1 <= kLdsRegions): the scores of this
// iteration (Graph::updateVertices) are computed here, by every working block
// into LDS for its own re-admission draws, and written out by block 0 for the
// next expansion, instead of by k_score (one launch fewer per iteration).
template<int NT, bool SCORE>
__global__ __launch_bounds__(NT) void k_frontier(Ctl* ctl, const QueryDev* qd, BatchCtl* bc,
                                                 const uint32_t* ftile_off, int32_t* slot_status, uint32_t Q,
                                                 TreeView tree, uint32_t* R, RegionView rv, uint64_t* status,
                                                 PlanParams PP, uint32_t R1, RegionParams RP)
{
    constexpr uint32_t kT = 4 * NT;  // nodes per tile
    __shared__ uint32_t lds[NT / 64];
    __shared__ uint32_t s_tile, s_pre;
    HIP_DYNAMIC_SHARED(float, s_dyn)
    const uint32_t ntiles = bc->ftiles, epoch = bc->epoch;
    if(blockIdx.x >= ntiles) return;  // no ticket to take: leave the counter to the blocks with work
    // the first ticket's round trip overlaps the scores
    if(threadIdx.x == 0) s_tile = atomicAdd(&bc->tile_frontier, 1u);
    const float* lscore = nullptr;
    if constexpr(SCORE)
    {
        // s_dyn: R1 scores then 1024 partial sums
        const float tot = block_scores<NT>(rv.valid, rv.total, rv.covered, s_dyn, s_dyn + R1, RP, PP.eps);
        if(blockIdx.x == 0)
        {
            for(uint32_t r = threadIdx.x; r < R1; r += NT) rv.score[r] = s_dyn[r];
            if(threadIdx.x == 0) ctl[0].score_total = tot;
        }
        lscore = s_dyn;
    }
    __syncthreads();
    uint32_t tile = s_tile;
    __syncthreads();  // every thread has the first ticket before the next is taken
    for(;; tile = next_tile(&bc->tile_frontier, &s_tile))
    {
        if(tile >= ntiles) break;
        const uint32_t q     = segment_of(ftile_off, Q, tile);
        const uint32_t first = ftile_off[q], lt = tile - first;
        const Iter it        = ctl[q].cur;
        const int32_t goal   = ctl[q].goal;
        const uint32_t Fn    = ctl[q].Fn;
        if(goal != INT_MAX)
        {
            if(lt == 0 && threadIdx.x == 0)
            {
                Iter n            = it;
                n.ts              = it.ts + Fn;
                n.status          = kDoneFound;
                ctl[q].nxt        = n;
                ctl[q].iterations = it.t;
                slot_status[q]    = kDoneFound;
            }
            continue;
        }
        const uint32_t tsn = it.ts + Fn, nt = (tsn + kT - 1) / kT;
        if(lt >= nt) continue;  // the slot's tile budget assumed every child admitted
        const QueryDev d       = qd[q];
        const float fa         = ctl[q].f_accept;
        const size_t base      = d.base;
        const uint32_t k0      = (uint32_t)d.seed, k1 = (uint32_t)(d.seed >> 32);
        const float* score     = SCORE ? lscore : rv.score + (size_t)q * R1;
        const int32_t* valid   = rv.valid + (size_t)q * R1;
        const uint32_t nbase   = lt * kT + threadIdx.x * 4;
        // the nodes' regions, then the new nodes' valid counts, each as one
        // round trip (unconditional loads at clamped indices; tsn >= 1)
        int rg[4], vc[4];
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            const uint32_t node = nbase + j;
            rg[j]               = tree.reg[base + (node < tsn ? node : tsn - 1)];
        }
#pragma unroll
        for(int j = 0; j < 4; ++j) vc[j] = valid[rg[j]];
        uint32_t w[4], cnt = 0;
#pragma unroll
        for(int j = 0; j < 4; ++j)
        {
            const uint32_t node = nbase + j;
            w[j]                = 0;
            if(node < tsn)
            {
                const int r = rg[j];
                if(node < it.ts)
                {
                    const uint4 dr = philox(node, it.t, d.query, PURPOSE_READMIT, k0, k1);
                    w[j]           = u01(dr.x) <= score[r] + fa ? 1u : 0u;
                }
                else
                    w[j] = vc[j] < PP.sparse_valid_below ? PP.sparse_repeat : PP.dense_repeat;
            }
            cnt += w[j];
        }
        uint32_t agg;
        const uint32_t loc = block_excl_scan<NT / 64>(cnt, &agg, lds);
        const uint32_t pre = tile_prefix(status, tile, agg, first, epoch, &s_pre);
        uint32_t pos       = pre + loc;
#pragma unroll
        for(int j = 0; j < 4; ++j)
            for(uint32_t k = 0; k < w[j]; ++k, ++pos)
                if(pos < d.cap) R[base + pos] = nbase + j;
        if(lt == nt - 1 && threadIdx.x == 0)
        {
            const Iter n      = make_iter(it.t + 1, tsn, pre + agg, d.cap, PP);
            ctl[q].nxt        = n;
            ctl[q].iterations = it.t;
            slot_status[q]    = n.status;
        }
    }
}

// Goal -> root walk of each listed slot (the reference walks it concurrently
// from every thread that reached the goal, KPAX.cu:482-501).  Lane 0 follows
// the parent chain; then the block gathers the rows.  Rows go out in
// goal -> root order; the host reverses them.
// Row control column k holds control ctl_of_col.{x,y,z,w}[k] (the reference's
// storage order, row_control).
__global__ __launch_bounds__(64) void k_path(Ctl* ctl, const QueryDev* qd, const uint32_t* slots, TreeView tree, int S,
                                             int C, uint32_t T, float dt, float* rows, int32_t* steps, int32_t* idx,
                                             uint32_t cap, int4 ctl_of_col)
{
    __shared__ uint32_t s_len;
    const uint32_t q   = slots[blockIdx.x];
    const size_t base  = qd[q].base;
    const int D        = S + C + 1;
    int32_t* my_idx    = idx + (size_t)q * cap;
    if(threadIdx.x == 0)
    {
        int32_t j    = ctl[q].goal;
        uint32_t len = 0;
        if(j != INT_MAX)
            while(j != -1 && len < cap)
            {
                my_idx[len++] = j;
                j             = tree.parent[base + j];
            }
        s_len          = (j == -1) ? len : 0u;
        ctl[q].path_len = s_len;
    }
    __syncthreads();
    const uint32_t len = s_len;
    float* my_rows     = rows + (size_t)q * cap * D;
    for(uint32_t e = threadIdx.x; e < len * (uint32_t)D; e += blockDim.x)
    {
        const uint32_t r = e / D, col = e - r * D;
        const size_t node = base + (uint32_t)my_idx[r];
        float v;
        if(col < (uint32_t)S)
            v = tree.st[(size_t)col * T + node];
        else if(col < (uint32_t)(S + C))
        {
            const uint32_t k = col - S;
            const int j      = k == 0 ? ctl_of_col.x : k == 1 ? ctl_of_col.y : k == 2 ? ctl_of_col.z : ctl_of_col.w;
            v                = tree.ct[(size_t)j * T + node];
        }
        else
            v = dt * (float)tree.steps[node];
        my_rows[e] = v;
    }
    for(uint32_t r = threadIdx.x; r < len; r += blockDim.x) steps[(size_t)q * cap + r] = tree.steps[base + my_idx[r]];
}

// ===========================================================================
// Host side
// ===========================================================================

size_t append_lds_bytes(const kpax_handle* h)
{
    return h->RP.r1_count <= kLdsRegions ? 2 * (size_t)h->RP.r1_count * sizeof(int32_t) : 0;
}

// Work-groups of k_expand that fit on the device at once: its waves are
// persistent, so more would only wait for the first ones to drain the queue.
// The single-slot expansion stages its environment's grid image in LDS when
// the handle has one in use (the image's bytes are its dynamic LDS).
bool lds_expand(const kpax_handle* h) { return h->Qmax == 1 && h->use_lds && h->lds_bytes > 0; }

template<int MODEL, bool MULTI, int GQ>
uint32_t expand_blocks_for(const kpax_handle* h)
{
    int per_cu = 0;
    if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_expand<MODEL, MULTI, GQ>, (int)h->expand_threads,
                                                    GQ == kGridLds ? h->lds_bytes : 0) != hipSuccess ||
       per_cu < 1)
        per_cu = 1;
    return (uint32_t)per_cu * (uint32_t)h->cu_count;
}

template<int MODEL>
uint32_t expand_blocks_model(const kpax_handle* h)
{
    if(h->Qmax != 1) return expand_blocks_for<MODEL, true, kGridCsr>(h);
    return lds_expand(h) ? expand_blocks_for<MODEL, false, kGridLds>(h) : expand_blocks_for<MODEL, false, kGridCsr>(h);
}

uint32_t expand_blocks(const kpax_handle* h)
{
    switch(h->cfg.model)
    {
        case KPAX_MODEL_DOUBLE_INTEGRATOR: return expand_blocks_model<KPAX_MODEL_DOUBLE_INTEGRATOR>(h);
        case KPAX_MODEL_DUBINS_AIRPLANE: return expand_blocks_model<KPAX_MODEL_DUBINS_AIRPLANE>(h);
        default: return expand_blocks_model<KPAX_MODEL_QUADCOPTER>(h);
    }
}

// Work-groups of the preparation pass (its segments of the longest-first
// order): 256 (one per CU on MI355X), fewer on a device with fewer CUs.
uint32_t order_blocks(const kpax_handle* h)
{
    return std::min<uint32_t>(kOrderBlocks, std::max<uint32_t>(1u, (uint32_t)h->cu_count));
}

int upload_expand_args(kpax_handle* h)
{
    ExpandArgs a;
    a.ctl       = h->ctl;
    a.qd        = h->qd;
    a.bc        = h->bc;
    a.child_off = h->child_off;
    a.run       = (const SlotRun*)h->slot_run;
    a.envs      = h->envs_dev;
    a.tree      = h->tree;
    a.ch        = h->ch;
    a.R         = h->R;
    a.rv        = h->rv;
    a.P         = h->P;
    a.RP        = h->RP;
    a.PP        = h->PP;
    a.GP        = h->GPm;
    a.T         = h->T;
    a.Q         = h->Qmax;
    a.R2        = h->R2;
    a.chunk     = h->expand_chunk;
    a.prep      = h->prep;
    a.tickets   = h->tickets;
    a.refill    = h->refill_lanes;
    a.prio_q    = h->wave_priority ? std::max(1, (h->P.steps_min + h->P.span - 1 + 3) / 4) : 0;
    a.ordered   = h->order_children ? 1u : 0u;
    a.order_blocks = order_blocks(h);
    a.prop      = h->prop;
    memset(&a.s0, 0, sizeof(a.s0));
    a.s0.cur.status = kNotStarted;
    HIPC(hipMemcpy(h->expand_args, &a, sizeof(a), hipMemcpyHostToDevice));
    return KPAX_OK;
}

template<int MODEL>
void launch_expand_model(kpax_handle* h)
{
    const ExpandArgs* A = (const ExpandArgs*)h->expand_args;
    if(h->Qmax != 1)
        hipLaunchKernelGGL((k_expand<MODEL, true, kGridCsr>), dim3(h->expand_blocks), dim3(h->expand_threads), 0,
                           h->stream, A);
    else if(h->expand_lds)
        hipLaunchKernelGGL((k_expand<MODEL, false, kGridLds>), dim3(h->expand_blocks), dim3(h->expand_threads),
                           h->lds_bytes, h->stream, A);
    else
        hipLaunchKernelGGL((k_expand<MODEL, false, kGridCsr>), dim3(h->expand_blocks), dim3(h->expand_threads), 0,
                           h->stream, A);
}


// The slot statuses of the iteration being enqueued: the pinned host buffer
// of its parity (Qmax statuses and the running count), written by the advance
// and the frontier pass; the host reads it once the iteration's event is
// done, with no copy on the stream.
int32_t* status_of_parity(const kpax_handle* h) { return h->slot_status + (size_t)h->status_parity * (h->Qmax + 1); }

AdvanceArgs advance_args(const kpax_handle* h)
{
    AdvanceArgs V;
    V.ctl         = h->ctl;
    V.qd          = h->qd;
    V.bc          = h->bc;
    V.child_off   = h->child_off;
    V.atile_off   = h->atile_off;
    V.ftile_off   = h->ftile_off;
    V.slot_status = status_of_parity(h);
    V.Q           = h->Qmax;
    V.st_append   = h->st_append;
    V.st_frontier = h->st_frontier;
    V.n_status    = h->n_status;
    V.PP          = h->PP;
    V.envs        = h->envs_dev;
    V.s0          = &((ExpandArgs*)h->expand_args)->s0;
    V.tile        = h->scan_tile;
    V.tickets     = h->tickets;
    V.run         = h->Qmax > 1 ? (SlotRun*)h->slot_run : nullptr;  // only the batched expansion reads it
    return V;
}

// The children's records for this iteration (k_prepare), each segment in
// longest-first order or all in child order.
template<int MODEL>
void launch_prepare(kpax_handle* h)
{
    const ExpandArgs* A = (const ExpandArgs*)h->expand_args;
    const dim3 ob(order_blocks(h));
    const AdvanceArgs V = advance_args(h);
    const int ordered   = h->order_children ? 1 : 0;
    if(h->Qmax == 1)
        hipLaunchKernelGGL((k_prepare<MODEL, false>), ob, dim3(kOrderThreads), 0, h->stream, A, ordered, V, 1);
    else
        hipLaunchKernelGGL((k_prepare<MODEL, true>), ob, dim3(kOrderThreads), 0, h->stream, A, ordered, V, 1);
}

// The preparation passes, then the expansion; `between` (profiling) is
// recorded after the former.
template<int MODEL>
int launch_prepare_and_expand(kpax_handle* h, hipEvent_t between)
{
    launch_prepare<MODEL>(h);
    if(between) HIPC(hipEventRecord(between, h->stream));
    launch_expand_model<MODEL>(h);
    return KPAX_OK;
}

int launch_expand(kpax_handle* h, hipEvent_t between)
{
    switch(h->cfg.model)
    {
        case KPAX_MODEL_DOUBLE_INTEGRATOR:
            return launch_prepare_and_expand<KPAX_MODEL_DOUBLE_INTEGRATOR>(h, between);
        case KPAX_MODEL_DUBINS_AIRPLANE: return launch_prepare_and_expand<KPAX_MODEL_DUBINS_AIRPLANE>(h, between);
        default: return launch_prepare_and_expand<KPAX_MODEL_QUADCOPTER>(h, between);
    }
}

// The append and frontier scans in blocks of NT threads (tiles of 4 NT items).
template<int NT>
void launch_append(kpax_handle* h)
{
    hipLaunchKernelGGL(k_append<NT>, dim3(h->scan_blocks), dim3(NT), append_lds_bytes(h), h->stream, h->ctl, h->qd,
                       h->bc, h->atile_off, h->Qmax, h->ch, h->tree, h->rv, h->st_append, h->PP, h->S, h->C, h->T,
                       (uint32_t)h->RP.r1_count, h->R2);
}

// Scores computed inside the frontier pass (single-slot handles whose R1
// regions fit the LDS) instead of by k_score.
bool fused_scores(const kpax_handle* h) { return h->Qmax == 1 && h->RP.r1_count <= kLdsRegions; }

template<int NT>
void launch_frontier(kpax_handle* h)
{
    if(fused_scores(h))
        hipLaunchKernelGGL((k_frontier<NT, true>), dim3(h->scan_blocks), dim3(NT),
                           ((size_t)h->RP.r1_count + 1024) * sizeof(float), h->stream, h->ctl, h->qd, h->bc,
                           h->ftile_off, status_of_parity(h), h->Qmax, h->tree, h->R, h->rv, h->st_frontier, h->PP,
                           (uint32_t)h->RP.r1_count, h->RP);
    else
        hipLaunchKernelGGL((k_frontier<NT, false>), dim3(h->scan_blocks), dim3(NT), 0, h->stream, h->ctl, h->qd,
                           h->bc, h->ftile_off, status_of_parity(h), h->Qmax, h->tree, h->R, h->rv, h->st_frontier,
                           h->PP, (uint32_t)h->RP.r1_count, h->RP);
}

// The launches of one iteration, whose kernels write the slot statuses into
// the pinned buffer of parity p.  With `ev` set, phase boundaries are
// recorded between the launches.
int enqueue_iteration(kpax_handle* h, int p, bool ev)
{
    const uint32_t Q     = h->Qmax;
    h->status_parity     = p;
    if(ev) HIPC(hipEventRecord(h->ph_ev[p][0], h->stream));
    int rc = launch_expand(h, ev ? h->ph_ev[p][1] : nullptr);
    if(rc != KPAX_OK) return rc;
    if(ev) HIPC(hipEventRecord(h->ph_ev[p][2], h->stream));
    switch(h->scan_tile)
    {
        case 1024: launch_append<256>(h); break;
        case 2048: launch_append<512>(h); break;
        default: launch_append<1024>(h); break;
    }
    if(ev) HIPC(hipEventRecord(h->ph_ev[p][3], h->stream));
    // the proposal and score boundaries only around launches that run (an
    // event pair with nothing between them would time the events themselves)
    h->ph_ran[p][0] = h->mix > 0.0f && h->prop_slots > 0;
    h->ph_ran[p][1] = !fused_scores(h);
    if(h->ph_ran[p][0])
    {
        if((rc = launch_propose(h, 0)) != KPAX_OK) return rc;
        if(ev) HIPC(hipEventRecord(h->ph_ev[p][4], h->stream));
    }
    if(h->ph_ran[p][1])
    {
        hipLaunchKernelGGL(k_score, dim3(Q), dim3(1024), 0, h->stream, h->ctl, h->rv, h->RP, h->PP.eps);
        if(ev) HIPC(hipEventRecord(h->ph_ev[p][5], h->stream));
    }
    switch(h->scan_tile)
    {
        case 1024: launch_frontier<256>(h); break;
        case 2048: launch_frontier<512>(h); break;
        default: launch_frontier<1024>(h); break;
    }
    if(ev) HIPC(hipEventRecord(h->ph_ev[p][6], h->stream));
    HIPC(hipGetLastError());
    return KPAX_OK;
}

// Captures one iteration per status parity as a HIP graph, for each
// expansion a single-slot handle may launch (with the LDS grid image, and
// without it for environments that have none).  A capture the runtime
// refuses leaves the handle on direct launches of the same kernels.
void capture_graphs(kpax_handle* h)
{
    h->graph_tried = true;
    const bool keep = h->expand_lds;
    for(int v = 0; v < (lds_expand(h) ? 2 : 1); ++v)
        for(int p = 0; p < 2; ++p)
        {
            h->expand_lds = v == 1;
            hipGraph_t g  = nullptr;
            if(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed) != hipSuccess)
            {
                (void)hipGetLastError();
                h->use_graph  = false;
                h->expand_lds = keep;
                return;
            }
            h->prop_slots      = 1;  // a captured iteration holds the proposal pass whatever runs
            const int rc       = enqueue_iteration(h, p, false);
            const hipError_t e = hipStreamEndCapture(h->stream, &g);
            if(rc != KPAX_OK || e != hipSuccess || !g ||
               hipGraphInstantiate(&h->graph_exec[v][p], g, nullptr, nullptr, 0) != hipSuccess)
            {
                (void)hipGetLastError();
                if(g) (void)hipGraphDestroy(g);
                h->use_graph  = false;
                h->expand_lds = keep;
                return;
            }
            (void)hipGraphDestroy(g);
        }
    h->expand_lds = keep;
}

// First-fit allocator of node ranges in [0, T).
struct NodeRanges {
    std::map<uint32_t, uint32_t> free_;  // start -> end
    explicit NodeRanges(uint32_t T) { free_[0] = T; }
    bool alloc(uint32_t n, uint32_t* base)
    {
        for(auto it = free_.begin(); it != free_.end(); ++it)
            if(it->second - it->first >= n)
            {
                *base               = it->first;
                const uint32_t end  = it->second;
                free_.erase(it);
                if(*base + n < end) free_[*base + n] = end;
                return true;
            }
        return false;
    }
    void release(uint32_t base, uint32_t n)
    {
        uint32_t lo = base, hi = base + n;
        auto nx = free_.lower_bound(lo);
        if(nx != free_.end() && nx->first == hi)
        {
            hi = nx->second;
            nx = free_.erase(nx);
        }
        if(nx != free_.begin())
        {
            auto pv = std::prev(nx);
            if(pv->second == lo)
            {
                lo = pv->first;
                free_.erase(pv);
            }
        }
        free_[lo] = hi;
    }
};

struct QueryIn {
    const float* start;
    const float* goal;
    uint64_t seed;
    uint32_t cap;
    uint32_t env;
};

void account_iteration(kpax_handle* h, int p)
{
    if(h->h_status[(size_t)p * (h->Qmax + 1) + h->Qmax] <= 0) return;  // no slot was running
    // event pairs -> phases prepare, expand, append, propose (when it ran),
    // score (when k_score ran), frontier; each phase ends at its own event
    // and starts at the last one recorded before it
    constexpr int kPhase[6] = {6, 0, 1, 5, 2, 3};
    float ms   = 0.0f;
    int from   = 0;
    for(int k = 0; k < 6; ++k)
    {
        if((k == 3 && !h->ph_ran[p][0]) || (k == 4 && !h->ph_ran[p][1])) continue;
        if(hipEventElapsedTime(&ms, h->ph_ev[p][from], h->ph_ev[p][k + 1]) == hipSuccess)
        {
            h->stats.phase_ms[kPhase[k]] += ms;
            h->stats.phase_launches[kPhase[k]] += 1;
        }
        from = k + 1;
    }
}

// Runs queries through the handle's slots until every one has finished.
// Results land in out[i] exactly as a plan of query i alone would return.
int run_queries_impl(kpax_handle* h, const QueryIn* qs, uint32_t n, kpax_result* out)
{
    const auto t0   = std::chrono::steady_clock::now();
    const uint32_t Q = h->Qmax;
    const int S = h->S, C = h->C, D = S + C + 1;
    for(uint32_t i = 0; i < n; ++i)
    {
        memset(&out[i], 0, sizeof(out[i]));
        out[i].goal_index  = -1;
        out[i].verified    = -1;
        out[i].state_dim   = S;
        out[i].control_dim = C;
    }
    std::vector<uint32_t> waiting;
    for(uint32_t i = 0; i < n; ++i)
    {
        const QueryIn& q = qs[i];
        bool ok          = q.start && q.goal && q.cap >= 1 && q.cap <= h->T && q.env < (uint32_t)kMaxEnvs &&
                  h->envs[q.env].start != nullptr;
        for(int d = 0; ok && d < S; ++d) ok = std::isfinite(q.start[d]);
        for(int d = 0; ok && d < 3; ++d) ok = std::isfinite(q.goal[d]);
        if(ok)
            waiting.push_back(i);
        else
            out[i].status = KPAX_INVALID_ARG;
    }
    HIPC(hipSetDevice(h->cfg.device));
    if(h->use_graph && !h->graph_tried) capture_graphs(h);
    const bool prof  = h->profiling;
    const bool graph = h->use_graph && !prof && !h->trace && !h->iter_hook;
    NodeRanges ranges(h->T);
    std::vector<int64_t> slot_query(Q, -1);
    std::vector<uint32_t> slot_base(Q, 0), slot_cap(Q, 0);
    std::vector<uint64_t> admit_iter(Q, 0);
    std::vector<uint32_t> free_slots;
    for(uint32_t s = Q; s-- > 0;) free_slots.push_back(s);  // lowest slot first
    uint32_t active = 0;
    uint64_t it     = 0;  // iterations launched by this call
    h->prop_slots   = 0;  // direct launches skip the proposal pass while no admitted slot samples from it

    auto admit = [&](bool first) -> int {
        uint32_t n_init = 0, n_prop = 0;
        for(size_t k = 0; k < waiting.size() && !free_slots.empty();)
        {
            const uint32_t qi = waiting[k];
            uint32_t base     = 0;
            if(!ranges.alloc(qs[qi].cap, &base))
            {
                ++k;
                continue;
            }
            const uint32_t s = free_slots.back();
            free_slots.pop_back();
            QueryDev& d = h->h_qd[s];
            memset(&d, 0, sizeof(d));
            d.base  = base;
            d.cap   = qs[qi].cap;
            d.env   = qs[qi].env;
            d.query = 0;  // the same draws as a plan of this query alone
            d.seed  = qs[qi].seed;
            d.mix   = d.cap >= h->prop_min_budget ? h->mix : 0.0f;  // proposal policy by tree budget
            if(d.mix > 0.0f) ++h->prop_slots, ++n_prop;
            // a single-slot handle expands with the LDS grid when this
            // query's environment has an image (else the CSR grid)
            if(Q == 1) h->expand_lds = lds_expand(h) && h->envs[d.env].has_lds;
            for(int j = 0; j < 3; ++j) d.goal[j] = qs[qi].goal[j];
            h->h_init_slots[n_init] = s;
            for(int j = 0; j < 16; ++j) h->h_init_starts[16 * (size_t)n_init + j] = j < S ? qs[qi].start[j] : 0.0f;
            ++n_init;
            slot_query[s] = qi;
            slot_base[s]  = base;
            slot_cap[s]   = qs[qi].cap;
            admit_iter[s] = it + 1;
            ++active;
            waiting.erase(waiting.begin() + (long)k);
        }
        if(!n_init) return KPAX_OK;
        if(prof && first) HIPC(hipEventRecord(h->init_ev[0], h->stream));
        HIPC(hipMemcpyAsync(h->qd, h->h_qd, Q * sizeof(QueryDev), hipMemcpyHostToDevice, h->stream));
        HIPC(hipMemcpyAsync(h->d_init_slots, h->h_init_slots, n_init * sizeof(uint32_t), hipMemcpyHostToDevice,
                            h->stream));
        HIPC(hipMemcpyAsync(h->d_init_starts, h->h_init_starts, 16 * (size_t)n_init * sizeof(float),
                            hipMemcpyHostToDevice, h->stream));
        const uint32_t R1 = (uint32_t)h->RP.r1_count;
        hipLaunchKernelGGL(k_reset_regions,
                           dim3(std::min<uint32_t>(blocks_for((uint64_t)n_init * h->R2, 256), 8 * h->cu_count)),
                           dim3(256), 0, h->stream, h->rv, h->d_init_slots, n_init, R1, h->R2);
        hipLaunchKernelGGL(k_init, dim3(n_init), dim3(256), 0, h->stream, h->ctl, h->qd, h->tree, h->R,
                           h->slot_status, Q + 1, h->d_init_slots, h->d_init_starts, S, C, h->T, h->RP, h->PP);
        HIPC(hipGetLastError());
        if(h->mix > 0.0f && n_prop > 0)
        {
            const int rc = launch_propose(h, n_init);
            if(rc != KPAX_OK) return rc;
        }
        if(prof && first) HIPC(hipEventRecord(h->init_ev[1], h->stream));
        return KPAX_OK;
    };

    auto harvest = [&](const std::vector<uint32_t>& done) -> int {
        const uint32_t nd = (uint32_t)done.size();
        for(uint32_t k = 0; k < nd; ++k) h->h_path_slots[k] = done[k];
        HIPC(hipMemcpyAsync(h->d_path_slots, h->h_path_slots, nd * sizeof(uint32_t), hipMemcpyHostToDevice,
                            h->stream));
        const int m = h->cfg.model;
        hipLaunchKernelGGL(k_path, dim3(nd), dim3(64), 0, h->stream, h->ctl, h->qd, h->d_path_slots, h->tree, S, C,
                           h->T, h->cfg.dt, h->path_rows, h->path_steps, h->path_idx, h->path_cap,
                           make_int4(row_control(m, 0), row_control(m, 1), row_control(m, 2), row_control(m, 3)));
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(h->h_ctl, h->ctl, Q * sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
        const size_t pc = h->path_cap;
        for(uint32_t s : done)
        {
            HIPC(hipMemcpyAsync(h->h_path_rows + s * pc * D, h->path_rows + s * pc * D, pc * D * sizeof(float),
                                hipMemcpyDeviceToHost, h->stream));
            HIPC(hipMemcpyAsync(h->h_path_steps + s * pc, h->path_steps + s * pc, pc * sizeof(int32_t),
                                hipMemcpyDeviceToHost, h->stream));
            HIPC(hipMemcpyAsync(h->h_path_idx + s * pc, h->path_idx + s * pc, pc * sizeof(int32_t),
                                hipMemcpyDeviceToHost, h->stream));
        }
        HIPC(hipStreamSynchronize(h->stream));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for(uint32_t s : done)
        {
            const Ctl& c     = h->h_ctl[s];
            kpax_result& r   = out[slot_query[s]];
            if(c.overflow) return fail(h, KPAX_DEVICE_ERROR, "tree index past the configured capacity");
            switch(c.nxt.status)
            {
                case kDoneFound: r.status = KPAX_OK; break;
                case kDoneBudget: r.status = KPAX_BUDGET_EXHAUSTED; break;
                default: r.status = KPAX_NOT_FOUND; break;  // max_iter reached
            }
            r.iterations       = c.iterations;
            r.tree_size        = c.nxt.ts;
            r.expansions       = c.expansions;
            r.integrated_steps = c.steps_done;
            r.plan_ms          = ms;
            const uint32_t len = c.path_len;
            if(r.status == KPAX_OK && c.goal != INT_MAX)
            {
                r.goal_index = c.goal;
                r.n_rows     = len;
                r.rows       = (float*)malloc(std::max<size_t>((size_t)len * D, 1) * sizeof(float));
                r.steps      = (int32_t*)malloc(std::max<size_t>(len, 1) * sizeof(int32_t));
                r.node_index = (int32_t*)malloc(std::max<size_t>(len, 1) * sizeof(int32_t));
                if(!r.rows || !r.steps || !r.node_index) return fail(h, KPAX_NO_MEMORY, "result allocation failed");
                for(uint32_t k = 0; k < len; ++k)
                {
                    const uint32_t src = len - 1 - k;  // root first
                    memcpy(r.rows + (size_t)k * D, h->h_path_rows + (s * pc + src) * D, D * sizeof(float));
                    r.steps[k]      = h->h_path_steps[s * pc + src];
                    r.node_index[k] = h->h_path_idx[s * pc + src];
                }
            }
            h->stats.expansions += c.expansions;
            h->stats.integrated_steps += c.steps_done;
            h->stats.plans += 1;
            ranges.release(slot_base[s], slot_cap[s]);
            if(h->h_qd[s].mix > 0.0f) --h->prop_slots;
            slot_query[s] = -1;
            free_slots.push_back(s);
            --active;
        }
        // keep the lowest free slot on top
        std::sort(free_slots.begin(), free_slots.end(), std::greater<uint32_t>());
        return KPAX_OK;
    };

    auto newly_done = [&](uint64_t j, std::vector<uint32_t>& done) {
        const int32_t* st = h->h_status + (size_t)(j & 1) * (Q + 1);
        for(uint32_t s = 0; s < Q; ++s)
            if(slot_query[s] >= 0 && admit_iter[s] <= j && st[s] != kRunning &&
               std::find(done.begin(), done.end(), s) == done.end())
                done.push_back(s);
    };

    int rc = admit(true);
    if(rc != KPAX_OK) return rc;
    uint64_t accounted = 0;
    while(active > 0)
    {
        ++it;
        const int p = (int)(it & 1);
        if(graph)
        {
            HIPC(hipGraphLaunch(h->graph_exec[h->expand_lds ? 1 : 0][p], h->stream));
            h->stats.graph_launches += 1;
        }
        else if((rc = enqueue_iteration(h, p, prof)) != KPAX_OK)
            return rc;
        HIPC(hipEventRecord(h->it_ev[p], h->stream));
        h->stats.iterations += 1;
        std::vector<uint32_t> done;
        if(h->trace)
        {
            HIPC(hipStreamSynchronize(h->stream));
            HIPC(hipMemcpy(h->h_ctl, h->ctl, Q * sizeof(Ctl), hipMemcpyDeviceToHost));
            for(uint32_t s = 0; s < std::min<uint32_t>(Q, 4); ++s)
            {
                const Iter& x = h->h_ctl[s].nxt;
                std::fprintf(stderr, "kpax: slot %u after iteration %llu: next t=%u ts=%u W=%u M=%u fan=%u status=%d\n",
                             s, (unsigned long long)it, x.t, x.ts, x.W, x.M, x.fan, x.status);
            }
        }
        if(h->iter_hook)
        {
            HIPC(hipStreamSynchronize(h->stream));
            if((rc = h->iter_hook(it)) != KPAX_OK) return rc;
        }
        // One iteration stays queued behind the one whose statuses are read,
        // so the GPU never waits for the host; a slot that finished is
        // skipped by the iterations queued after it.
        if(it >= 2)
        {
            HIPC(hipEventSynchronize(h->it_ev[(it - 1) & 1]));
            if(prof && accounted < it - 1) account_iteration(h, (int)((it - 1) & 1));
            accounted = std::max(accounted, it - 1);
            newly_done(it - 1, done);
        }
        if(!done.empty())
        {
            HIPC(hipStreamSynchronize(h->stream));
            if(prof && accounted < it) account_iteration(h, p);
            accounted = it;
            newly_done(it, done);
            if((rc = harvest(done)) != KPAX_OK) return rc;
            if((rc = admit(false)) != KPAX_OK) return rc;
        }
    }
    HIPC(hipStreamSynchronize(h->stream));
    if(prof)
    {
        float ms = 0.0f;
        if(hipEventElapsedTime(&ms, h->init_ev[0], h->init_ev[1]) == hipSuccess)
        {
            h->stats.phase_ms[4] += ms;
            h->stats.phase_launches[4] += 1;
        }
    }
    // FP64 re-verification of every returned trajectory (host, after timing)
    if(h->cfg.verify)
    {
        std::vector<uint32_t> todo;
        for(uint32_t i = 0; i < n; ++i)
            if(out[i].n_rows > 0) todo.push_back(i);
        std::atomic<size_t> next{0};
        auto work = [&] {
            for(size_t k; (k = next.fetch_add(1)) < todo.size();)
            {
                kpax_result& r   = out[todo[k]];
                const QueryIn& q = qs[todo[k]];
                const Env& e     = h->envs[q.env];
                float dev        = 0.0f;
                const bool ok = verify_path_fp64(h->cfg, e.host.data(), e.n_obs, r.rows, r.steps, r.n_rows, q.goal,
                                                 &dev);
                r.verified       = ok ? 1 : 0;
                r.verify_max_dev = dev;
            }
        };
        const unsigned nt = std::min<unsigned>((unsigned)todo.size(),
                                               std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
        if(nt <= 1)
            work();
        else
        {
            std::vector<std::thread> pool;
            for(unsigned k = 0; k < nt; ++k) pool.emplace_back(work);
            for(auto& th : pool) th.join();
        }
    }
    return KPAX_OK;
}

// run_queries_impl, and on a failure: the stream drained, every slot idle
// again on the device (a slot left running would be advanced by the next
// call's iterations over node ranges given to new queries), and the rows of
// results already built released.
int run_queries(kpax_handle* h, const QueryIn* qs, uint32_t n, kpax_result* out)
{
    const int rc = run_queries_impl(h, qs, n, out);
    if(rc == KPAX_OK) return rc;
    (void)hipStreamSynchronize(h->stream);
    (void)hipGetLastError();
    const uint32_t Q = h->Qmax;
    std::vector<Ctl> c0(Q);
    for(Ctl& c : c0)
    {
        memset(&c, 0, sizeof(c));
        c.cur.status = kNotStarted;
        c.nxt.status = kNotStarted;
        c.goal       = INT_MAX;
    }
    std::vector<int32_t> s0(Q + 1, kNotStarted);
    s0[Q] = 0;
    (void)hipMemcpy(h->ctl, c0.data(), Q * sizeof(Ctl), hipMemcpyHostToDevice);
    for(int par = 0; par < 2; ++par) memcpy(h->h_status + (size_t)par * (Q + 1), s0.data(), (Q + 1) * sizeof(int32_t));
    for(uint32_t i = 0; i < n; ++i) kpax_result_free(&out[i]);
    return rc;
}

void drop_graphs(kpax_handle* h)
{
    for(auto& ev : h->graph_exec)
        for(auto& e : ev)
            if(e)
            {
                (void)hipGraphExecDestroy(e);
                e = nullptr;
            }
    h->graph_tried = false;
}

}  // namespace

int kpax::refresh_launch_state(kpax_handle* h)
{
    drop_graphs(h);
    return upload_expand_args(h);
}

// The single-query expansion's work-group size: the one of 256, 320, 512 and
// 640 threads with the most resident waves per CU under its dynamic LDS (the
// smaller on a tie).  Measured on the quadcopter forest (22 KB image): 256
// and 512 threads ran the expansion in 2.79 ms per plan, 320 in 2.88 and 640
// (one resident work-group per CU) in 3.10.
static uint32_t pick_expand_threads(kpax_handle* h)
{
    if(h->Qmax != 1) return 256u;
    uint32_t best = 256u, best_waves = 0;
    for(uint32_t t : {256u, 320u, 512u, 640u})
    {
        h->expand_threads = t;
        const uint32_t waves = expand_blocks(h) / (uint32_t)h->cu_count * (t / 64);
        if(waves > best_waves)
        {
            best       = t;
            best_waves = waves;
        }
    }
    return best;
}

int kpax::refresh_lds_launch(kpax_handle* h)
{
    uint32_t bytes = 0;
    if(h->Qmax == 1)
        for(const Env& e : h->envs)
            if(e.has_lds) bytes = std::max(bytes, e.lg.bytes);
    h->lds_bytes = bytes;
    if(!h->launch_set)
    {
        h->expand_threads = pick_expand_threads(h);
        h->expand_blocks  = expand_blocks(h);
    }
    return refresh_launch_state(h);
}

namespace {

int apply_launch(kpax_handle* h, const kpax_launch& l)
{
    if(l.expand_blocks < 1 || l.expand_blocks > (1u << 20) || l.expand_chunk < 16 || l.expand_chunk > 4096 ||
       l.expand_chunk % 16 != 0 || l.scan_blocks < 1 || l.scan_blocks > (1u << 20) || l.refill_lanes < 1 ||
       l.refill_lanes > 64 || (l.scan_tile != 1024 && l.scan_tile != 2048 && l.scan_tile != 4096) ||
       (l.lds_grid != 0 && l.lds_grid != 1) || l.expand_threads < 64 || l.expand_threads > 640 ||
       l.expand_threads % 64 != 0 || (l.wave_priority != 0 && l.wave_priority != 1))
        return fail(h, KPAX_INVALID_ARG, "launch shape out of range");
    HIPC(hipStreamSynchronize(h->stream));
    h->launch_set    = true;
    h->use_lds       = l.lds_grid != 0;
    h->expand_threads = h->Qmax == 1 ? l.expand_threads : 256u;
    h->expand_blocks = l.expand_blocks;
    h->expand_chunk  = l.expand_chunk;
    h->scan_blocks   = l.scan_blocks;
    h->use_graph     = l.use_graph != 0;
    h->order_children = l.order_children != 0;
    h->refill_lanes   = l.refill_lanes;
    h->scan_tile      = l.scan_tile;
    h->wave_priority  = l.wave_priority != 0;
    drop_graphs(h);
    const int rc = upload_expand_args(h);
    if(rc == KPAX_OK && !l.use_graph) h->graph_tried = true;  // stay on direct launches
    return rc;
}

kpax_launch current_launch(const kpax_handle* h)
{
    kpax_launch l;
    l.expand_blocks = h->expand_blocks;
    l.expand_chunk  = h->expand_chunk;
    l.scan_blocks   = h->scan_blocks;
    l.use_graph     = h->use_graph ? 1 : 0;
    l.order_children = h->order_children ? 1 : 0;
    l.refill_lanes   = h->refill_lanes;
    l.scan_tile      = h->scan_tile;
    l.lds_grid       = h->use_lds ? 1 : 0;
    l.expand_threads = h->expand_threads;
    l.wave_priority  = h->wave_priority ? 1 : 0;
    return l;
}

int check_obstacles(kpax_handle* h, const float* aabb6, uint32_t n)
{
    if(n && !aabb6) return fail(h, KPAX_INVALID_ARG, "obstacles pointer is NULL");
    for(uint32_t i = 0; i < n; ++i)
        for(int d = 0; d < 3; ++d)
        {
            const float a = aabb6[6 * (size_t)i + d], b = aabb6[6 * (size_t)i + 3 + d];
            if(!std::isfinite(a) || !std::isfinite(b) || a > b)
                return fail(h, KPAX_INVALID_ARG, "obstacle " + std::to_string(i) + " is not a finite min<=max box");
        }
    return KPAX_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int kpax_create_batch(const kpax_config* cfg, uint32_t max_queries, uint64_t node_capacity, kpax_handle** out)
{
    if(!cfg || !out) return KPAX_INVALID_ARG;
    *out = nullptr;
    if(max_queries < 1 || max_queries > kMaxSlots || node_capacity < 1 || node_capacity > (1ull << 31))
        return KPAX_INVALID_ARG;
    if(!validate_config(*cfg).empty()) return KPAX_INVALID_ARG;
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) return KPAX_NO_DEVICE;
    kpax_handle* h = new(std::nothrow) kpax_handle();
    if(!h) return KPAX_NO_MEMORY;
    h->cfg   = *cfg;
    h->trace = std::getenv("KPAX_TRACE") != nullptr;
    h->S     = state_dim(cfg->model);
    h->C     = control_dim(cfg->model);
    h->P     = make_model_params(*cfg);
    h->RP    = make_region_params(*cfg);
    h->GPm   = make_grid_params(*cfg);
    h->PP    = make_plan_params(*cfg);
    h->T     = (uint32_t)node_capacity;
    h->Qmax  = max_queries;
    // Longest-edge-first ordering pays in the single-query form of the
    // quadcopter and the Dubins airplane.  The double integrator's short
    // steps leave too little to gain for the extra counting pass (plan 1.146
    // ms with it, 1.128 without).  In the batched form it mixes the children
    // of all running queries (their environments, grids and node ranges)
    // within each wave, where child order keeps a wave's lanes on one query:
    // the serving batch ran at 2,750 plans/s without it and 2,270 with it.
    h->order_children = max_queries == 1 && cfg->model != KPAX_MODEL_DOUBLE_INTEGRATOR;
    // Issue priority by the longest live edge pays where a step is long
    // (quadcopter forest: expansion 2.77 -> 2.64 ms per plan); the double
    // integrator's steps are short and its iterations small, and the ballots
    // per trip cost more than the ordering gains (0.547 -> 0.560 ms).
    h->wave_priority = cfg->model != KPAX_MODEL_DOUBLE_INTEGRATOR;
    h->G     = (uint32_t)h->GPm.D[0] * h->GPm.D[1] * h->GPm.D[2];
    h->R2    = (uint32_t)h->RP.r1_count * h->RP.r2_per;
    h->envs.resize(kMaxEnvs);
    int rc = KPAX_OK;
#define TRY(x)               \
    do                       \
    {                        \
        rc = (x);            \
        if(rc != KPAX_OK)    \
        {                    \
            kpax_destroy(h); \
            return rc;       \
        }                    \
    } while(0)
    if(hipSetDevice(cfg->device) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    hipDeviceProp_t prop;
    if(hipGetDeviceProperties(&prop, cfg->device) == hipSuccess && prop.multiProcessorCount > 0)
        h->cu_count = prop.multiProcessorCount;
    if(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    {
        // an environment's LDS image: at most 72 KiB, so that two 640-thread
        // expansion work-groups (five waves per SIMD) fit in one CU's LDS
        size_t lds = 72u << 10;
        if(prop.sharedMemPerBlock > 0) lds = std::min<size_t>(lds, prop.sharedMemPerBlock);
        if(prop.maxSharedMemoryPerMultiProcessor > 0)
            lds = std::min<size_t>(lds, prop.maxSharedMemoryPerMultiProcessor / 2);
        h->lds_budget = (uint32_t)(lds & ~(size_t)15);
    }
    h->expand_threads = 256u;
    h->expand_blocks  = expand_blocks(h);
    TRY(noalloc_scratch(h, &h->noalloc));  // the allocation-free operators' array on this device
    // Children per work ticket: a wave owns its whole chunk, so near the end
    // of a launch one that has just taken a chunk works through it while the
    // others drain.  Single quadcopter plans measured 2.64 ms of expansion
    // per plan at 64, 2.54 at 32 and 2.59 at 16; the batch and the Dubins
    // maze ran best at 64.
    h->expand_chunk  = max_queries == 1 && cfg->model == KPAX_MODEL_QUADCOPTER ? 32u : 64u;
    // With 32-child chunks a wave's lanes idle sooner after a refill, and a
    // service pass at 8 idle lanes measured best for single quadcopter plans
    // (expansion 2.39 against 2.43 ms per plan at 16); the Dubins maze (41.8
    // against 42.7 ms) and the serving batch (2,754-2,770 against 2,661-2,719
    // plans/s) stay at 16, the double integrator is the same at either.
    h->refill_lanes  = max_queries == 1 && cfg->model == KPAX_MODEL_QUADCOPTER ? 8u : 16u;
    const uint32_t Q = h->Qmax;
    const size_t T   = h->T;
    // tiles of 4096 items run as 1024-thread blocks, two resident per CU
    h->scan_blocks   = std::min<uint32_t>(blocks_for(T, h->scan_tile) + Q, 2 * h->cu_count);
    TRY(dalloc(h, &h->tree.st, T * h->S));
    TRY(dalloc(h, &h->tree.ct, T * h->C));
    TRY(dalloc(h, &h->tree.steps, T));
    TRY(dalloc(h, &h->tree.parent, T));
    TRY(dalloc(h, &h->tree.reg, T));
    TRY(dalloc(h, &h->ch.rec, T * (kChildWords / 4)));
    TRY(dalloc(h, &h->ch.meta, T));
    TRY(dalloc(h, &h->R, T));
    const size_t r1 = (size_t)h->RP.r1_count * Q, r2 = (size_t)h->R2 * Q;
    TRY(dalloc(h, &h->rv.valid, r1));
    TRY(dalloc(h, &h->rv.total, r1));
    TRY(dalloc(h, &h->rv.covered, r1));
    TRY(dalloc(h, &h->rv.score, r1));
    TRY(dalloc(h, &h->rv.r2_state, r2));
    TRY(dalloc(h, &h->ctl, Q));
    TRY(dalloc(h, &h->qd, Q));
    TRY(dalloc(h, &h->bc, 1));
    TRY(dalloc(h, &h->tickets, (size_t)kTicketGroups * kTicketStride));
    {
        SlotRun* run = nullptr;
        TRY(dalloc(h, &run, (size_t)Q));
        h->slot_run = run;
    }
    if(hipMemset(h->tickets, 0, (size_t)kTicketGroups * kTicketStride * 4) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    TRY(dalloc(h, &h->child_off, Q + 1));
    TRY(dalloc(h, &h->atile_off, Q + 1));
    TRY(dalloc(h, &h->ftile_off, Q + 1));
    TRY(dalloc(h, &h->prep, 2 * T));
    h->n_status = (uint32_t)(T / kTile + Q + 2);
    TRY(dalloc(h, &h->st_append, h->n_status));
    TRY(dalloc(h, &h->st_frontier, h->n_status));
    TRY(dalloc(h, &h->d_init_slots, Q));
    TRY(dalloc(h, &h->d_init_starts, 16 * (size_t)Q));
    TRY(dalloc(h, &h->d_path_slots, Q));
    TRY(dalloc(h, &h->envs_dev, kMaxEnvs));
    h->path_cap     = cfg->max_iter + 2;
    const size_t pc = (size_t)Q * h->path_cap;
    const int D     = h->S + h->C + 1;
    TRY(dalloc(h, &h->path_rows, pc * D));
    TRY(dalloc(h, &h->path_steps, pc));
    TRY(dalloc(h, &h->path_idx, pc));
    TRY(dalloc(h, (char**)&h->expand_args, sizeof(ExpandArgs)));
    auto pinned = [&](void** p, size_t bytes) {
        return hipHostMalloc(p, std::max<size_t>(bytes, 1), hipHostMallocDefault) == hipSuccess ? KPAX_OK
                                                                                                : KPAX_NO_MEMORY;
    };
    TRY(pinned((void**)&h->h_status, 2 * (Q + 1) * sizeof(int32_t)));
    if(hipHostGetDevicePointer((void**)&h->slot_status, h->h_status, 0) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    TRY(pinned((void**)&h->h_ctl, Q * sizeof(Ctl)));
    TRY(pinned((void**)&h->h_qd, Q * sizeof(QueryDev)));
    TRY(pinned((void**)&h->h_init_slots, Q * sizeof(uint32_t)));
    TRY(pinned((void**)&h->h_init_starts, 16 * (size_t)Q * sizeof(float)));
    TRY(pinned((void**)&h->h_path_slots, Q * sizeof(uint32_t)));
    TRY(pinned((void**)&h->h_path_rows, pc * D * sizeof(float)));
    TRY(pinned((void**)&h->h_path_steps, pc * sizeof(int32_t)));
    TRY(pinned((void**)&h->h_path_idx, pc * sizeof(int32_t)));
    for(int i = 0; i < 2; ++i)
    {
        if(hipEventCreateWithFlags(&h->it_ev[i], hipEventDisableTiming) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
        if(hipEventCreate(&h->init_ev[i]) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
        for(int k = 0; k < 7; ++k)
            if(hipEventCreate(&h->ph_ev[i][k]) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    }
    // every slot idle; scan state at epoch 0 (never current)
    std::vector<Ctl> c0(Q);
    for(Ctl& c : c0)
    {
        memset(&c, 0, sizeof(c));
        c.cur.status = kNotStarted;
        c.nxt.status = kNotStarted;
        c.goal       = INT_MAX;
    }
    if(hipMemcpy(h->ctl, c0.data(), Q * sizeof(Ctl), hipMemcpyHostToDevice) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    std::vector<int32_t> s0(Q + 1, kNotStarted);
    s0[Q] = 0;
    for(int par = 0; par < 2; ++par) memcpy(h->h_status + (size_t)par * (Q + 1), s0.data(), (Q + 1) * sizeof(int32_t));
    if(hipMemset(h->bc, 0, sizeof(BatchCtl)) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    if(hipMemset(h->qd, 0, Q * sizeof(QueryDev)) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    if(hipMemset(h->st_append, 0, h->n_status * sizeof(uint64_t)) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    if(hipMemset(h->st_frontier, 0, h->n_status * sizeof(uint64_t)) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    if(hipMemset(h->envs_dev, 0, kMaxEnvs * sizeof(GridView)) != hipSuccess) TRY(KPAX_DEVICE_ERROR);
    TRY(upload_expand_args(h));
    // an empty environment 0 until kpax_set_obstacles
    TRY(build_grid(h, 0, nullptr, 0));
#undef TRY
    *out = h;
    return KPAX_OK;
}

int kpax_create(const kpax_config* cfg, kpax_handle** out)
{
    if(!cfg) return KPAX_INVALID_ARG;
    return kpax_create_batch(cfg, 1, cfg->max_tree, out);
}

int kpax_destroy(kpax_handle* h)
{
    if(!h) return KPAX_INVALID_ARG;
    if(h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graphs(h);
    for(void* p : h->allocs) (void)hipFree(p);
    if(h->grid_scratch) (void)hipFree(h->grid_scratch);
    for(Env& e : h->envs) free_env(e);
    void* pinned[] = {h->h_status,     h->h_ctl,       h->h_qd,         h->h_init_slots, h->h_init_starts,
                      h->h_path_slots, h->h_path_rows, h->h_path_steps, h->h_path_idx};
    for(void* p : pinned)
        if(p) (void)hipHostFree(p);
    for(int i = 0; i < 2; ++i)
    {
        if(h->it_ev[i]) (void)hipEventDestroy(h->it_ev[i]);
        if(h->init_ev[i]) (void)hipEventDestroy(h->init_ev[i]);
        for(hipEvent_t e : h->ph_ev[i])
            if(e) (void)hipEventDestroy(e);
    }
    if(h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return KPAX_OK;
}

const char* kpax_last_error(const kpax_handle* h)
{
    if(!h) return "null handle";
    return h->err.c_str();
}

int kpax_set_environment(kpax_handle* h, uint32_t env, const float* aabb6, uint32_t n)
{
    if(!h) return KPAX_INVALID_ARG;
    if(env >= (uint32_t)kMaxEnvs) return fail(h, KPAX_INVALID_ARG, "environment index out of range");
    const int rc = check_obstacles(h, aabb6, n);
    if(rc != KPAX_OK) return rc;
    HIPC(hipSetDevice(h->cfg.device));
    return build_grid(h, env, aabb6, n);
}

int kpax_set_obstacles(kpax_handle* h, const float* aabb6, uint32_t n)
{
    return kpax_set_environment(h, 0, aabb6, n);
}

int kpax_plan(kpax_handle* h, const float* start, const float* goal, uint64_t seed, kpax_result* out)
{
    if(!h || !start || !goal || !out) return fail(h, KPAX_INVALID_ARG, "null argument");
    QueryIn q{start, goal, seed, std::min<uint32_t>(h->cfg.max_tree, h->T), 0};
    const int rc = run_queries(h, &q, 1, out);
    if(rc != KPAX_OK) return rc;
    if(out->status == KPAX_INVALID_ARG) return fail(h, KPAX_INVALID_ARG, "start or goal is not finite");
    h->last_tree = out->tree_size;
    h->planned   = true;
    return KPAX_OK;
}

int kpax_plan_batch(kpax_handle* h, const kpax_query* queries, uint32_t n, kpax_result* out)
{
    if(!h || (n && (!queries || !out))) return fail(h, KPAX_INVALID_ARG, "null argument");
    if(n == 0) return KPAX_OK;
    std::vector<QueryIn> qs(n);
    for(uint32_t i = 0; i < n; ++i)
    {
        const kpax_query& q = queries[i];
        qs[i]               = QueryIn{q.start, q.goal, q.seed, q.budget ? q.budget : h->cfg.max_tree, q.env};
    }
    const int rc = run_queries(h, qs.data(), n, out);
    h->planned   = false;  // slot 0's tree is no longer one single plan's
    return rc;
}

// planBench (KPAX.cu:245-300) with writeDeviceVectorsToCSV (KPAX.cu:544-618):
// the same directory layout and CSV formats (helper.cuh:57-108), written after
// every iteration.  Rows past the tree are not written (the reference writes
// its whole zero-filled capacity); the frontier flags and repeat counts are
// those of the repeat list built for the next iteration.
int kpax_plan_dump(kpax_handle* h, const float* start, const float* goal, uint64_t seed, const char* dir,
                   int32_t bench_itr, kpax_result* out)
{
    if(!h || !start || !goal || !out || !dir) return fail(h, KPAX_INVALID_ARG, "null argument");
    const std::string root = std::string(dir) + "/";
    const std::string b    = std::to_string(bench_itr);
    const char* sub[] = {"Samples/Samples", "Parents/Parents", "TotalCountPerVertex/TotalCountPerVertex",
                         "ValidCountPerVertex/ValidCountPerVertex", "Frontier/Frontier",
                         "FrontierRepeatCount/FrontierRepeatCount", "VertexScores/VertexScores",
                         "FrontierSize/FrontierSize", "TreeSize/TreeSize", "ExpandedNodes/ExpandedNodes",
                         "ControlPathToGoal/ControlPathToGoal"};
    std::error_code ec;
    for(const char* d : sub) std::filesystem::create_directories(root + d + b, ec);
    std::filesystem::create_directories(root + "ExecutionTime", ec);
    if(ec) return fail(h, KPAX_IO_ERROR, "cannot create " + root + ": " + ec.message());
    const uint32_t cap = std::min<uint32_t>(h->cfg.max_tree, h->T);
    const int S = h->S, C = h->C, D = S + C + 1;
    const uint32_t R1 = (uint32_t)h->RP.r1_count;
    auto vec_rows = [](const std::string& f, const std::vector<std::string>& rows) {
        std::ofstream o(f, std::ios_base::app);
        for(const auto& r : rows) o << r << "\n";
        return (bool)o;
    };
    auto join = [](auto begin, auto end, int prec) {
        std::ostringstream os;
        os << std::fixed << std::setprecision(prec);
        for(auto it = begin; it != end; ++it) os << (it == begin ? "" : ",") << *it;
        return os.str();
    };
    auto value = [](const std::string& f, double v, bool is_float) {
        std::ofstream o(f, std::ios_base::app);
        if(is_float) o << std::fixed << std::setprecision(10);
        if(is_float)
            o << v << "\n";
        else
            o << (long long)v << "\n";
        return (bool)o;
    };
    h->iter_hook = [&](uint64_t it) -> int {
        Ctl c;
        HIPC(hipMemcpy(&c, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
        if(c.cur.status != kRunning) return KPAX_OK;  // queued after the plan ended
        const uint32_t itr = c.cur.t;  // h_itr_ of the iteration just run
        const bool found   = c.goal != INT_MAX;
        const uint32_t ts  = found ? c.cur.ts + c.Fn : c.nxt.ts;
        const std::string n = std::to_string(itr);
        std::vector<float> st((size_t)S * ts), ct((size_t)C * ts);
        std::vector<int32_t> steps(ts), par(ts), valid(R1), total(R1);
        std::vector<float> score(R1);
        for(int d = 0; d < S; ++d)
            HIPC(hipMemcpy(st.data() + (size_t)d * ts, h->tree.st + (size_t)d * h->T, (size_t)ts * 4,
                           hipMemcpyDeviceToHost));
        for(int d = 0; d < C; ++d)
            HIPC(hipMemcpy(ct.data() + (size_t)d * ts, h->tree.ct + (size_t)d * h->T, (size_t)ts * 4,
                           hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(steps.data(), h->tree.steps, (size_t)ts * 4, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(par.data(), h->tree.parent, (size_t)ts * 4, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(valid.data(), h->rv.valid, R1 * 4, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(total.data(), h->rv.total, R1 * 4, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(score.data(), h->rv.score, R1 * 4, hipMemcpyDeviceToHost));
        // repeat counts of the next iteration's list (empty once the goal is found)
        std::vector<int32_t> repeat(ts, 0);
        if(!found)
        {
            const uint32_t wl = std::min<uint32_t>(c.nxt.W, cap);
            std::vector<uint32_t> R(wl);
            if(wl) HIPC(hipMemcpy(R.data(), h->R, (size_t)wl * 4, hipMemcpyDeviceToHost));
            for(uint32_t k = 0; k < wl; ++k)
                if(R[k] < ts) repeat[R[k]]++;
        }
        std::vector<std::string> rows(ts);
        std::vector<float> row(D);
        for(uint32_t i = 0; i < ts; ++i)
        {
            for(int d = 0; d < S; ++d) row[d] = st[(size_t)d * ts + i];
            for(int d = 0; d < C; ++d) row[S + d] = ct[(size_t)row_control(h->cfg.model, d) * ts + i];
            row[S + C] = h->cfg.dt * (float)steps[i];
            rows[i]    = join(row.begin(), row.end(), 6);
        }
        bool ok = vec_rows(root + "Samples/Samples" + b + "/samples" + n + ".csv", rows);
        for(uint32_t i = 0; i < ts; ++i) rows[i] = std::to_string(par[i]);
        ok = ok && vec_rows(root + "Parents/Parents" + b + "/parents" + n + ".csv", rows);
        ok = ok && vec_rows(root + "TotalCountPerVertex/TotalCountPerVertex" + b + "/totalCountPerVertex.csv",
                            {join(total.begin(), total.end(), 6)});
        ok = ok && vec_rows(root + "ValidCountPerVertex/ValidCountPerVertex" + b + "/validCountPerVertex.csv",
                            {join(valid.begin(), valid.end(), 6)});
        std::vector<int32_t> flag(ts);
        uint32_t fsize = 0;
        for(uint32_t i = 0; i < ts; ++i)
        {
            flag[i] = repeat[i] > 0;
            fsize += (uint32_t)flag[i];
        }
        ok = ok && vec_rows(root + "Frontier/Frontier" + b + "/frontier.csv", {join(flag.begin(), flag.end(), 6)});
        ok = ok && vec_rows(root + "FrontierRepeatCount/FrontierRepeatCount" + b + "/frontierRepeatCount.csv",
                            {join(repeat.begin(), repeat.end(), 6)});
        ok = ok && vec_rows(root + "VertexScores/VertexScores" + b + "/vertexScores.csv",
                            {join(score.begin(), score.end(), 6)});
        ok = ok && value(root + "FrontierSize/FrontierSize" + b + "/frontierSize.csv", fsize, false);
        ok = ok && value(root + "TreeSize/TreeSize" + b + "/treeSize.csv", ts, false);
        ok = ok && value(root + "ExpandedNodes/ExpandedNodes" + b + "/expandedNodes.csv", c.cur.M, false);
        if(!ok) return fail(h, KPAX_IO_ERROR, "cannot write under " + root);
        (void)it;
        return KPAX_OK;
    };
    QueryIn q{start, goal, seed, cap, 0};
    const auto t0 = std::chrono::steady_clock::now();
    const int rc  = run_queries(h, &q, 1, out);
    h->iter_hook  = nullptr;
    if(rc != KPAX_OK) return rc;
    if(out->status == KPAX_INVALID_ARG) return fail(h, KPAX_INVALID_ARG, "start or goal is not finite");
    h->last_tree = out->tree_size;
    h->planned   = true;
    if(out->n_rows)
    {
        // goal -> root rows, as d_controlPathToGoal_
        std::vector<std::string> rows(out->n_rows);
        for(uint32_t k = 0; k < out->n_rows; ++k)
        {
            const float* r = out->rows + (size_t)(out->n_rows - 1 - k) * D;
            rows[k]        = join(r, r + D, 6);
        }
        if(!vec_rows(root + "ControlPathToGoal/ControlPathToGoal" + b + "/controlPathToGoal.csv", rows))
            return fail(h, KPAX_IO_ERROR, "cannot write the path");
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if(!value(root + "ExecutionTime/executionTime.csv", secs, true)) return fail(h, KPAX_IO_ERROR, "cannot write");
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
    if(!h->planned) return fail(h, KPAX_INVALID_ARG, "no single plan has run on this handle");
    const uint32_t ts = h->last_tree;
    const size_t T    = h->T;  // a single plan occupies slot 0 at node 0
    if(tree_size) *tree_size = ts;
    if(cap < ts && (states || controls)) return fail(h, KPAX_INVALID_ARG, "cap is smaller than the tree");
    HIPC(hipSetDevice(h->cfg.device));
    for(int d = 0; states && d < h->S; ++d)
        HIPC(hipMemcpy(states + (size_t)d * cap, h->tree.st + (size_t)d * T, (size_t)ts * 4, hipMemcpyDeviceToHost));
    for(int d = 0; controls && d < h->C; ++d)
        HIPC(hipMemcpy(controls + (size_t)d * cap, h->tree.ct + (size_t)d * T, (size_t)ts * 4,
                       hipMemcpyDeviceToHost));
    if(steps) HIPC(hipMemcpy(steps, h->tree.steps, (size_t)ts * 4, hipMemcpyDeviceToHost));
    if(parents) HIPC(hipMemcpy(parents, h->tree.parent, (size_t)ts * 4, hipMemcpyDeviceToHost));
    if(regions) HIPC(hipMemcpy(regions, h->tree.reg, (size_t)ts * 4, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_get_proposal(kpax_handle* h, uint32_t cap, float* mean_sigma, uint32_t* tree_size)
{
    if(!h || !mean_sigma) return fail(h, KPAX_INVALID_ARG, "null argument");
    if(!h->planned) return fail(h, KPAX_INVALID_ARG, "no single plan has run on this handle");
    if(!h->prop || h->mix <= 0.0f) return fail(h, KPAX_INVALID_ARG, "no learned proposal is set");
    const uint32_t ts = h->last_tree;
    if(tree_size) *tree_size = ts;
    if(cap < ts) return fail(h, KPAX_INVALID_ARG, "cap is smaller than the tree");
    HIPC(hipSetDevice(h->cfg.device));
    for(int j = 0; j < 8; ++j)
        HIPC(hipMemcpy(mean_sigma + (size_t)j * cap, h->prop + (size_t)j * h->T, (size_t)ts * 4,
                       hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_get_regions(kpax_handle* h, int32_t* valid, int32_t* total, int32_t* covered, float* scores)
{
    if(!h) return KPAX_INVALID_ARG;
    HIPC(hipSetDevice(h->cfg.device));
    const size_t n = (size_t)h->RP.r1_count * 4;  // slot 0
    if(valid) HIPC(hipMemcpy(valid, h->rv.valid, n, hipMemcpyDeviceToHost));
    if(total) HIPC(hipMemcpy(total, h->rv.total, n, hipMemcpyDeviceToHost));
    if(covered) HIPC(hipMemcpy(covered, h->rv.covered, n, hipMemcpyDeviceToHost));
    if(scores) HIPC(hipMemcpy(scores, h->rv.score, n, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

int kpax_set_profiling(kpax_handle* h, int32_t on)
{
    if(!h) return KPAX_INVALID_ARG;
    h->profiling = on != 0;
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

int kpax_get_launch(kpax_handle* h, kpax_launch* out)
{
    if(!h || !out) return KPAX_INVALID_ARG;
    *out = current_launch(h);
    return KPAX_OK;
}

int kpax_set_launch(kpax_handle* h, const kpax_launch* launch)
{
    if(!h || !launch) return KPAX_INVALID_ARG;
    HIPC(hipSetDevice(h->cfg.device));
    return apply_launch(h, *launch);
}

// Coordinate search over the launch shape: each knob in turn takes the
// candidate whose best-of-`repeats` wall time over the sample batch is
// lowest, the others held at their current best.
int kpax_autotune(kpax_handle* h, const kpax_query* queries, uint32_t n, uint32_t repeats, kpax_launch* best)
{
    if(!h || !queries || n == 0) return fail(h, KPAX_INVALID_ARG, "autotune needs sample queries");
    repeats = std::max<uint32_t>(repeats, 1);
    HIPC(hipSetDevice(h->cfg.device));
    const bool prof = h->profiling;
    h->profiling    = false;
    std::vector<kpax_result> ref(n), got(n);
    auto run = [&](const kpax_launch& l, double* ms, std::vector<kpax_result>& out) -> int {
        int rc = apply_launch(h, l);
        if(rc != KPAX_OK) return rc;
        *ms = 1e300;
        for(uint32_t r = 0; r < repeats; ++r)
        {
            for(auto& x : out) kpax_result_free(&x);
            const auto t0 = std::chrono::steady_clock::now();
            if((rc = kpax_plan_batch(h, queries, n, out.data())) != KPAX_OK) return rc;
            *ms = std::min(*ms, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        return KPAX_OK;
    };
    auto same = [&](const std::vector<kpax_result>& a, const std::vector<kpax_result>& b) {
        for(uint32_t i = 0; i < n; ++i)
            if(a[i].status != b[i].status || a[i].tree_size != b[i].tree_size ||
               a[i].iterations != b[i].iterations || a[i].expansions != b[i].expansions ||
               a[i].goal_index != b[i].goal_index)
                return false;
        return true;
    };
    const uint32_t cu = (uint32_t)h->cu_count;
    const uint32_t occ = std::max<uint32_t>(1, expand_blocks(h) / cu);
    kpax_launch cur = current_launch(h);
    cur.expand_blocks = occ * cu;
    double cur_ms     = 0.0;
    int rc            = run(cur, &cur_ms, ref);
    // one knob at a time: each candidate is the best shape so far with one
    // field changed
    struct Knob {
        void (*set)(kpax_launch&, uint32_t);
        std::vector<uint32_t> values;
    };
    std::vector<uint32_t> blocks;
    for(uint32_t k = 1; k <= occ; ++k) blocks.push_back(k * cu);
    const std::vector<Knob> knobs = {
        {[](kpax_launch& l, uint32_t v) { l.expand_blocks = v; }, blocks},
        {[](kpax_launch& l, uint32_t v) { l.expand_chunk = v; }, {32u, 64u, 128u, 256u, 512u}},
        {[](kpax_launch& l, uint32_t v) { l.scan_tile = v; }, {4096u, 2048u, 1024u}},
        {[](kpax_launch& l, uint32_t v) { l.scan_blocks = v; }, {cu, 2 * cu, 4 * cu}},
        {[](kpax_launch& l, uint32_t v) { l.use_graph = (int32_t)v; }, {1u, 0u}},
        {[](kpax_launch& l, uint32_t v) { l.order_children = (int32_t)v; }, {1u, 0u}},
        {[](kpax_launch& l, uint32_t v) { l.refill_lanes = v; }, {1u, 8u, 16u, 24u, 32u}},
        {[](kpax_launch& l, uint32_t v) { l.lds_grid = (int32_t)v; }, {1u, 0u}},
        {[](kpax_launch& l, uint32_t v) { l.wave_priority = (int32_t)v; }, {1u, 0u}},
    };
    for(size_t k = 0; rc == KPAX_OK && k < knobs.size(); ++k)
        for(const uint32_t v : knobs[k].values)
        {
            kpax_launch cand = cur;
            knobs[k].set(cand, v);
            if(memcmp(&cand, &cur, sizeof(cand)) == 0) continue;
            double ms = 0.0;
            if((rc = run(cand, &ms, got)) != KPAX_OK) break;
            if(!same(ref, got))
            {
                rc = fail(h, KPAX_DEVICE_ERROR, "a launch shape changed the planner's results");
                break;
            }
            if(ms < cur_ms)
            {
                cur    = cand;
                cur_ms = ms;
            }
        }
    for(auto& x : ref) kpax_result_free(&x);
    for(auto& x : got) kpax_result_free(&x);
    h->profiling = prof;
    if(rc != KPAX_OK) return rc;
    rc = apply_launch(h, cur);
    if(rc == KPAX_OK && best) *best = cur;
    return rc;
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
    hipLaunchKernelGGL(k_score, dim3(1), dim3(1024), 0, h->stream, (Ctl*)nullptr, rv, h->RP, h->PP.eps);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    HIPC(hipMemcpy(scores, rv.score, r1 * 4, hipMemcpyDeviceToHost));
    return KPAX_OK;
}

}  // extern "C"
