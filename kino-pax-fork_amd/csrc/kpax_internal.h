// kpax_internal.h — structures shared by the HIP planner (kpax_planner.hip,
// kpax_ops.hip) and the host-only parts of libkpax (kpax_host.cpp).  Not part of the ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/kpax.h"

namespace kpax {

constexpr int kMaxS = 12;
constexpr int kMaxC = 4;

// Iteration status codes on the device.  Terminal codes are kept apart from
// "running" (0): KPAX_OK is also 0, so the host maps them to kpax_status.
enum : int32_t { kRunning = 0, kDoneFound = 1, kDoneNotFound = 2, kDoneBudget = 3, kNotStarted = 4 };

// Per-iteration control record.  The frontier kernel of iteration t writes
// `nxt`; the one-block advance kernel copies it to `cur` at the start of
// iteration t+1, so no kernel ever changes a field its own blocks read.
struct Iter {
    uint32_t t;       // iteration number, 1-based
    uint32_t ts;      // tree size at the start of the iteration
    uint32_t W;       // length of the frontier repeat list (may exceed its capacity)
    uint32_t M;       // children expanded in this iteration
    uint32_t fan;     // children per repeat entry
    int32_t status;   // kRunning, or a terminal code (kDone*)
};

// Per-query planner record.
struct Ctl {
    Iter cur;
    Iter nxt;
    int32_t goal;          // smallest tree index inside the goal ball, INT32_MAX if none
    uint32_t Fn;           // children appended in the current iteration
    float f_accept;        // re-admission slack of the current iteration
    float score_total;     // sum of raw region scores (current iteration)
    uint64_t expansions;   // children propagated since the plan started
    uint64_t steps_done;   // integration steps taken by those children
    uint32_t iterations;   // last iteration that ran
    uint32_t path_len;     // rows written by the path kernel
    uint32_t overflow;     // set if any kernel found an index past its buffer
    uint32_t pad[5];
};

// A query slot: where its tree lives and what it plans for.
struct QueryDev {
    uint32_t base;   // first node of the slot's tree, children and repeat-list segments
    uint32_t cap;    // tree cap of the query
    uint32_t env;    // environment (collision grid) index
    uint32_t query;  // Philox query word
    uint64_t seed;
    float goal[4];   // goal ball centre (xyz)
    float mix;       // probability of the learned proposal for this query (0: uniform controls only)
    uint32_t pad;
};

// Shared record of one iteration over all query slots (written by the
// advance kernel, read by the others).
struct BatchCtl {
    uint32_t children;        // children of all running slots
    uint32_t atiles, ftiles;  // append and frontier scan tiles (frontier: an upper bound)
    uint32_t running;         // slots running in this iteration
    uint32_t tile_unused, tile_append, tile_frontier;  // work tickets of the scans
    uint32_t epoch;           // tag of this iteration's scan status words
    uint32_t pad[4];
};

// Largest number of query slots of one handle (the advance kernel gives each
// slot one thread of a 1024-thread block).
constexpr uint32_t kMaxSlots = 1024;

// Dynamics and checks, with every constant folded on the host exactly as the
// reference's macros fold in float.
struct ModelParams {
    int32_t model;
    int32_t di_textbook;
    float dt, half, d6;  // STEP_SIZE, 0.5f * STEP_SIZE, STEP_SIZE / 6.0f
    int32_t steps_min, span;
    float ws_min[3], ws_max[3];
    float ctrl_min[4], ctrl_rng[4];
    float vmin, vmax, pmin, pmax;
    float negnu, negmu, g, mi, k1x, k1y, k1z, ixi, iyi, izi;
};

struct RegionParams {
    int32_t L[3], L2[3];
    float mn[3], ext[3], size[3];
    int32_t r1_count, r2_per;
    float vol;
};

struct GridParams {
    int32_t D[3];
    float mn[3], cell[3];
    float inv[3];  // cells per metre, D / extent: cell index = (x - mn) * inv
};

// Planner parameters that are not per-model or per-query.
struct PlanParams {
    uint32_t max_iter;
    uint32_t fanout, root_repeat, sparse_repeat, dense_repeat;
    int32_t sparse_valid_below;
    float eps;
    float goal_radius;
    int32_t prune;  // goal-progress pruning (PruneKPAX) on/off
    float prune_expl, prune_goal, prune_scale;
};

int state_dim(int model);
int control_dim(int model);

// Row format of trajectories and dumped trees: [state, control, duration]
// with the controls in the reference's storage order.  The planner keeps a
// Dubins airplane's controls in their draw order (a, yawRate, pitchRate,
// statePropagatorSpatialHash.cu:141-143, the order of kpax_config.ctrl_min);
// the reference stores (yawRate, pitchRate, a) (x1[6..8], :223-225).  Row
// control column k holds the planner's control row_control(model, k).
inline int row_control(int model, int k) { return model == KPAX_MODEL_DUBINS_AIRPLANE ? (k + 1) % 3 : k; }
std::string validate_config(const kpax_config& c);
ModelParams make_model_params(const kpax_config& c);
RegionParams make_region_params(const kpax_config& c);
GridParams make_grid_params(const kpax_config& c);
PlanParams make_plan_params(const kpax_config& c);

// Host FP64 re-verification (kpax_host.cpp).
bool verify_path_fp64(const kpax_config& c, const float* obs, uint32_t n_obs, const float* rows,
                      const int32_t* steps, uint32_t n_rows, const float* goal, float* max_dev);

}  // namespace kpax
