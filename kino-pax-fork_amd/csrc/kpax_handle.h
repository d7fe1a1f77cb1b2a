// kpax_handle.h — the planner handle and the helpers shared by the planner
// (kpax_planner.hip) and the hot-path operators (kpax_ops.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "kpax_device.cuh"
#include "kpax_internal.h"

namespace kpax {

// Node arrays are SoA with column stride T (the handle's node capacity);
// query slot q owns columns [base_q, base_q + cap_q).
struct TreeView {
    float* st;
    float* ct;
    int32_t* steps;
    int32_t* parent;
    int32_t* reg;
};

// Children of an iteration, written by the lane that made each one when it
// finishes: a 64-byte record (words 0-11 state, 12-15 control; four vector
// stores) and a 16-byte meta word {integrated steps | admission flags << 24,
// parent, R1 region, R2 cell}.  The append kernel reads every child's meta
// word and the record only of the children it admits.
constexpr int kChildWords = 16;
constexpr uint32_t kPrepSlotShift = 17;  // prepared child record word 2: steps (< 2^17) | slot << 17
constexpr uint32_t kMetaFlagShift = 24;  // integrated steps < 2^24 (an edge has at most 100000)
struct ChildView {
    uint4* rec;   // kChildWords / 4 uint4 per child
    uint4* meta;  // one per child
};

// Region state of slot q: counters and scores at q * R1, sub-region cells at
// q * R1 * R2_per_R1.
struct RegionView {
    int32_t* valid;
    int32_t* total;
    int32_t* covered;
    float* score;
    uint64_t* r2_state;  // first iteration << 32 | smallest slot, ~0 = never reached
};

// One environment: obstacles and their CSR grid (device), plus the host copy
// the FP64 re-verifier reads.
struct Env {
    int32_t* start  = nullptr;
    int32_t* items  = nullptr;
    float4* lo      = nullptr;
    float4* hi      = nullptr;
    float4* box     = nullptr;  // per CSR item its obstacle's box: box[2k] min corner, box[2k + 1] max corner
    uint32_t* occ   = nullptr;  // occupancy bitmap of the learned proposal's features
    uint4* lds_img  = nullptr;  // the grid as a work-group stages it into LDS (dev::LdsGrid)
    dev::LdsGrid lg{};
    bool has_lds    = false;    // lds_img holds the current obstacles' image
    uint32_t n_obs  = 0;
    uint32_t k      = 0;
    std::vector<float> host;
    // capacities (items) of the buffers above, kept across rebuilds
    size_t lo_cap = 0, start_cap = 0, items_cap = 0, box_cap = 0, lds_cap = 0;
};

// LDS image of an environment (built on the host from its boxes): the
// coarse grid dimensions that minimise the mean obstacle count per cell
// within `budget` bytes, or none (lg.bytes = 0) when no image fits or the
// best one's mean exceeds max_mean.
std::vector<uint4> build_lds_image(const kpax_config& cfg, const float* aabb6, uint32_t n, uint32_t budget,
                                   double max_mean, dev::LdsGrid* lg);

constexpr int kTicketGroups = 64, kTicketStride = 32;  // expansion chunk counters, 128 B apart
constexpr int kTile       = 1024;  // items per scan tile (256 threads x 4); the planner's smallest tile
constexpr int kLdsRegions = 4096;  // R1 counts up to this size are summed in LDS
constexpr int kMaxEnvs    = 64;    // environments of one handle (LDS table in batched expansion)

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)std::max<uint64_t>(1, (n + bs - 1) / bs); }


}  // namespace kpax

struct kpax_handle {
    kpax_config cfg{};
    int S = 0, C = 0;
    kpax::ModelParams P{};
    kpax::RegionParams RP{};
    kpax::GridParams GPm{};
    kpax::PlanParams PP{};
    std::string err;
    int cu_count       = 256;
    hipStream_t stream = nullptr;
    // capacity
    uint32_t T    = 0;  // tree nodes over all slots (column stride of the node arrays)
    uint32_t Qmax = 1;  // query slots
    uint32_t G    = 0;  // grid cells
    uint32_t R2   = 0;  // R2 cells per slot
    // planner state (device)
    kpax::TreeView tree{};
    kpax::ChildView ch{};
    uint32_t* R = nullptr;
    kpax::RegionView rv{};
    kpax::Ctl* ctl          = nullptr;  // Qmax records
    kpax::QueryDev* qd      = nullptr;  // Qmax records
    kpax::BatchCtl* bc      = nullptr;
    uint32_t* child_off     = nullptr;  // Qmax + 1 offsets, written by the advance kernel
    uint32_t* atile_off     = nullptr;
    uint32_t* ftile_off     = nullptr;
    int32_t* slot_status    = nullptr;  // device address of h_status (2 parities x (Qmax statuses + the running count))
    int status_parity       = 0;        // parity of the iteration being enqueued
    uint64_t* st_append     = nullptr;
    uint64_t* st_frontier   = nullptr;
    uint32_t n_status       = 0;
    uint32_t* d_init_slots  = nullptr;  // slots being initialised
    float* d_init_starts    = nullptr;  // 16 floats per slot being initialised
    kpax::dev::GridView* envs_dev = nullptr;  // kMaxEnvs grids
    void* expand_args       = nullptr;
    uint32_t expand_blocks  = 1;
    uint32_t expand_threads = 256;      // work-group size of the expansion (640 with an LDS-staged grid)
    uint32_t expand_chunk   = 64;
    bool use_lds            = true;     // stage the collision grid in LDS where the environment's image fits
    uint32_t lds_budget     = 0;        // largest LDS image of an environment, bytes
    uint32_t lds_bytes      = 0;        // dynamic LDS of the single-query expansion: the largest image in use
    bool launch_set         = false;    // the launch shape came from kpax_set_launch (kept across environments)
    uint64_t* noalloc       = nullptr;  // device address of the allocation-free operators' scratch array
    int32_t* grid_scratch   = nullptr;  // counts, cursor, look-back words and ticket of the grid build
    bool order_children     = true;     // hand out children longest drawn edge first (default: single-query handles)
    uint32_t* tickets       = nullptr;  // kTicketGroups expansion chunk counters
    void* slot_run          = nullptr;  // Qmax 32-byte records: what the batched refill reads of a slot
    uint32_t refill_lanes   = 16;       // idle lanes that trigger a wave's service pass
    bool wave_priority      = false;    // expansion waves set their issue priority by their longest live edge
    uint32_t scan_tile      = 4096;     // items per append/frontier tile: 1024, 2048 or 4096 (4 per thread)
    uint4* prep             = nullptr;  // per processing position: the child's prepared draws (2 x uint4)
    uint32_t scan_blocks    = 1;
    uint32_t path_cap       = 0;        // rows per slot
    float* path_rows        = nullptr;  // Qmax x path_cap x (S + C + 1), goal -> root
    int32_t* path_steps     = nullptr;
    int32_t* path_idx       = nullptr;
    uint32_t* d_path_slots  = nullptr;
    // pinned host mirrors
    int32_t* h_status       = nullptr;  // 2 x (Qmax + 1): statuses after iterations of each parity (pinned,
                                        // written by the kernels)
    kpax::Ctl* h_ctl        = nullptr;
    kpax::QueryDev* h_qd    = nullptr;
    uint32_t* h_init_slots  = nullptr;
    float* h_init_starts    = nullptr;
    uint32_t* h_path_slots  = nullptr;
    float* h_path_rows      = nullptr;
    int32_t* h_path_steps   = nullptr;
    int32_t* h_path_idx     = nullptr;
    hipEvent_t it_ev[2]     = {nullptr, nullptr};
    hipEvent_t ph_ev[2][7]  = {};       // phase boundaries of iterations of each parity (profiling):
                                        // before the preparation pass, then after it, expand, append,
                                        // propose, score, frontier
    bool ph_ran[2][2]       = {};       // per parity: the proposal pass and k_score were launched
    hipEvent_t init_ev[2]   = {nullptr, nullptr};
    // one captured iteration per status-buffer parity, replayed while profiling is off
    hipGraphExec_t graph_exec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [LDS expansion][status parity]
    bool expand_lds         = false;    // single-slot: the running query's environment has a staged LDS image
    bool graph_tried        = false;
    bool use_graph          = false;
    // learned proposal (kpax_set_proposal)
    void* mlp               = nullptr;  // bf16 network layouts
    float* prop             = nullptr;  // per node: 4 means then 4 sigmas, column stride T
    float mix               = 0.0f;
    uint32_t prop_min_budget = 0;       // queries with a smaller tree budget sample uniformly
    uint32_t prop_slots     = 1;        // admitted slots that sample from the proposal (0: skip its pass)
    // environments
    std::vector<kpax::Env> envs;
    // last single-query plan (for kpax_get_tree / kpax_get_regions)
    uint32_t last_tree = 0;
    bool planned       = false;
    // profiling; KPAX_TRACE=1 prints every iteration record to stderr
    bool profiling = false;
    bool trace     = false;
    // called after every iteration (with the stream drained) by kpax_plan_dump
    std::function<int(uint64_t)> iter_hook;
    kpax_stats stats{};
    std::vector<void*> allocs;
};

namespace kpax {

inline int fail(kpax_handle* h, int code, const std::string& msg)
{
    if(h) h->err = msg;
    return code;
}

#define HIPC(expr)                                                                                       \
    do                                                                                                   \
    {                                                                                                    \
        hipError_t e_ = (expr);                                                                          \
        if(e_ != hipSuccess)                                                                             \
            return ::kpax::fail(h, KPAX_DEVICE_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while(0)

template<class T>
int dalloc(kpax_handle* h, T** p, size_t count)
{
    void* q      = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if(e != hipSuccess) return fail(h, KPAX_NO_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    h->allocs.push_back(q);
    *p = (T*)q;
    return KPAX_OK;
}

inline dev::GridView grid_view(const kpax_handle* h, uint32_t env)
{
    dev::GridView g;
    g.gp    = h->GPm;
    g.start = h->envs[env].start;
    g.box   = h->envs[env].box;
    g.occ   = h->envs[env].occ;
    g.lg    = h->envs[env].lg;
    if(!h->envs[env].has_lds) g.lg.bytes = 0;
    g.lds   = nullptr;
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

// Builds environment `env`'s CSR grid on the device (kpax_ops.hip).
int build_grid(kpax_handle* h, uint32_t env, const float* aabb6, uint32_t n);
void free_env(Env& e);
// Occupancy bitmap of environment `env` for the proposal's features (kpax_mlp.hip).
int build_occupancy(kpax_handle* h, uint32_t env);
// Proposal pass over the nodes appended this iteration, or over the roots of
// the n_init slots listed in d_init_slots (kpax_mlp.hip).
int launch_propose(kpax_handle* h, uint32_t n_init);
// After a change of what the iteration launches: drops the captured graphs
// and rewrites the expansion's argument block (kpax_planner.hip).
int refresh_launch_state(kpax_handle* h);
// Device address of the allocation-free operators' scratch array on the
// current device (kpax_ops.hip), looked up once per device.
int noalloc_scratch(kpax_handle* h, uint64_t** out);
// After a change of the environments' LDS images: the single-query
// expansion's dynamic LDS, work-group size and (unless kpax_set_launch fixed
// them) work-group count (kpax_planner.hip).
int refresh_lds_launch(kpax_handle* h);

}  // namespace kpax
