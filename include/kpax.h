/*
 * kpax.h — C ABI of the MI355X-native Kino-PAX planner (libkpax.so).
 *
 * This is the drop-in boundary for the reference's planner hot path.  The
 * reference (kranzowe/Kino-PAX-Fork) exposes a C++ class hierarchy compiled
 * with nvcc; every entry point below names the reference interface it
 * replaces.  Plain pointers and sizes only: host buffers in, host buffers out.
 * The library owns its device memory and one HIP stream per handle.
 *
 * Conventions
 *   S = state dimension, C = control dimension of the configured model.
 *   "SoA" arrays are dimension-major: x[d * n + i] is dimension d of item i.
 *   Every function returns a kpax_status; on failure kpax_last_error()
 *   describes the cause.  Nothing calls exit() and nothing prints.
 */
#ifndef KPAX_H
#define KPAX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KPAX_ABI_VERSION 1

/* Status codes.  The reference has none (errors are unchecked and a missing
 * obstacle file calls exit(1), src/helper/helper.cu:18-22). */
typedef enum kpax_status {
    KPAX_OK               = 0,
    KPAX_NOT_FOUND        = 1,  /* plan finished without reaching the goal   */
    KPAX_BUDGET_EXHAUSTED = 2,  /* tree cap reached before the goal          */
    KPAX_INVALID_ARG      = -1,
    KPAX_DEVICE_ERROR     = -2,
    KPAX_NO_MEMORY        = -3,
    KPAX_NO_DEVICE        = -4,
    KPAX_IO_ERROR         = -5
} kpax_status;

/* Dynamics models.  Numbering follows the reference's MODEL macro
 * (include/config/config.h:209; dispatch src/statePropagator/statePropagator.cu:375-390). */
typedef enum kpax_model {
    KPAX_MODEL_DOUBLE_INTEGRATOR = 1, /* S=6  C=3  statePropagatorSpatialHash.cu:72-134 */
    KPAX_MODEL_DUBINS_AIRPLANE   = 2, /* S=6  C=3  statePropagatorSpatialHash.cu:139-229 */
    KPAX_MODEL_QUADCOPTER        = 3  /* S=12 C=4  statePropagatorSpatialHash.cu:234-308 + statePropagator.cu:325-370 */
} kpax_model;

/* Runtime configuration.  Replaces the compile-time macros of
 * include/config/config.h:205-308 (one model per binary there). */
typedef struct kpax_config {
    int32_t model;             /* kpax_model                                            */
    int32_t di_textbook;       /* double integrator only: 0 = reference update
                                  x += (5/6)v dt + (1/3)a dt^2 (statePropagatorSpatialHash.cu:92-94),
                                  1 = textbook x += v dt + a dt^2/2                      */
    float   dt;                /* STEP_SIZE                                             */
    int32_t steps_min;         /* integration steps per edge drawn uniformly in          */
    int32_t steps_max;         /*   [steps_min, steps_max] (reference: 1 + (int)(u*10))  */
    float   ws_min[3];         /* workspace box per axis (reference: W_MIN..W_MAX cube)  */
    float   ws_max[3];
    float   ctrl_min[4];       /* control box; quad: Zc, Lc, Mc, Nc; DI: ax, ay, az;     */
    float   ctrl_max[4];       /*   Dubins: a, yawRate, pitchRate (reference draw order) */
    float   vel_min;           /* per-axis velocity box: DI vx,vy,vz; quad body u,v,w;  */
    float   vel_max;           /*   Dubins airspeed v (V_MIN/V_MAX)                      */
    float   pitch_min;         /* Dubins pitch box (DUBINS_AIRPLANE_MIN/MAX_PITCH)       */
    float   pitch_max;
    float   mass;              /* quad: MASS, IX, IY, IZ, NU, MU, GRAVITY                */
    float   inertia[3];
    float   nu;
    float   mu;
    float   gravity;
    int32_t r1_len[3];         /* R1 regions per workspace axis (W_R1_LENGTH)            */
    int32_t r2_len[3];         /* R2 sub-regions per R1 per axis (W_R2_LENGTH)           */
    int32_t grid_dims[3];      /* uniform collision grid cells per axis                  */
    uint32_t max_tree;         /* MAX_TREE_SIZE                                          */
    uint32_t max_iter;         /* MAX_ITER                                               */
    float   goal_radius;       /* GOAL_THRESH (strict <, position distance)              */
    uint32_t fanout;           /* children per frontier repeat entry (h_activeBlockSize_ = 32) */
    uint32_t root_repeat;      /* repeat count of the root (KPAX.cu:77, = 5)             */
    uint32_t sparse_repeat;    /* repeat count of a new node in a sparse region (= 15)   */
    uint32_t dense_repeat;     /* repeat count otherwise (= 1)                            */
    int32_t sparse_valid_below;/* region is sparse while valid count < this (= 10)       */
    float   eps;               /* EPSILON (= 1e-2)                                        */
    int32_t verify;            /* 1: re-propagate every returned trajectory in FP64 and
                                  re-check it against all obstacles on the host          */
    int32_t device;            /* HIP device ordinal                                      */
    /* Goal-progress pruning of PruneKPAX (src/planners/PruneKPAX.cu:466-501): an
     * admitted child is kept with probability
     *   clamp(exploration + goal_bias * clamp(progress / progress_scale + 1/2, 0, 1), eps, 1)
     * where progress = parent's goal distance - child's goal distance.  0 = off. */
    int32_t prune;
    float   prune_exploration;     /* h_explorationBias_ (default 0.3)                     */
    float   prune_goal_bias;       /* h_goalBias_ (default 0.7)                            */
    float   prune_progress_scale;  /* h_progressScale_ (default 10)                        */
} kpax_config;

/* Trajectory returned by kpax_plan.  Replaces the public fields
 * h_pathToGoal_ / h_controlPathToGoal_ / h_treeSize_ / h_itr_ of
 * include/planners/Planner.cuh:31-34 (rows there run goal -> root). */
typedef struct kpax_result {
    int32_t  status;           /* KPAX_OK, KPAX_NOT_FOUND or KPAX_BUDGET_EXHAUSTED       */
    uint32_t n_rows;           /* path nodes, root first, goal node last (0 if not found) */
    uint32_t state_dim;
    uint32_t control_dim;
    float*   rows;             /* n_rows x (S + C + 1): [state, control, duration]; the
                                  root row has zero control and duration (reference row
                                  format, KPAX.cu:485-500).  Controls in the reference's
                                  storage order: the Dubins airplane's as (yawRate,
                                  pitchRate, a) (statePropagatorSpatialHash.cu:223-225),
                                  while kpax_config.ctrl_min/max and kpax_get_tree keep
                                  the draw order (a, yawRate, pitchRate)                 */
    int32_t* steps;            /* n_rows: integer steps of the edge into each node       */
    int32_t* node_index;       /* n_rows: tree index of each node                        */
    int32_t  goal_index;       /* tree index of the goal node (h_pathToGoal_), -1 if none */
    uint32_t iterations;       /* h_itr_                                                  */
    uint32_t tree_size;        /* h_treeSize_                                             */
    uint64_t expansions;       /* children propagated in total                            */
    uint64_t integrated_steps; /* integration steps those children took                   */
    double   plan_ms;          /* wall time of kpax_plan including initialisation         */
    int32_t  verified;         /* -1 not run, 0 failed, 1 passed the FP64 re-check        */
    float    verify_max_dev;   /* max |FP64 re-propagated - stored FP32| over all rows    */
} kpax_result;

/* Per-phase device time, accumulated with HIP events on the handle's stream
 * while profiling is on (kpax_set_profiling).  Phases: 0 expand (the
 * expansion kernel alone), 1 append, 2 score, 3 frontier, 4 init, 5 propose
 * (the learned proposal's network on the appended nodes; zero while no
 * proposal is set), 6 prepare (the children's ordering counts and prepared
 * records, before the expansion). */
#define KPAX_N_PHASES 7
typedef struct kpax_stats {
    double   phase_ms[KPAX_N_PHASES];
    uint64_t phase_launches[KPAX_N_PHASES];
    uint64_t expansions;       /* children propagated by the timed expand launches       */
    uint64_t integrated_steps; /* integration steps taken by those children               */
    uint64_t plans;            /* queries finished (single or batched)                    */
    uint64_t iterations;       /* planner iterations launched                             */
    uint64_t graph_launches;   /* iterations launched as one captured HIP graph           */
} kpax_stats;

/* ---- configuration ---------------------------------------------------- */

/* Fill *out with the built-in workload `workload`:
 *   1  6D double integrator, [0,1]^3, N = 65,536
 *   2  12D quadcopter forest, [0,10]^2 x [0,5], N = 2^20
 *   3  6D Dubins airplane maze, [0,1]^3, N = 2^21
 * Replaces editing include/config/config.h and recompiling. */
int kpax_config_workload(int32_t workload, kpax_config* out);

/* Dimensions of a model: *state_dim = S, *control_dim = C. */
int kpax_model_dims(int32_t model, int32_t* state_dim, int32_t* control_dim);

/* ---- planner lifecycle (Planner / KPAX classes) ------------------------ */

typedef struct kpax_handle kpax_handle;

/* Replaces KPAX::KPAX() (src/planners/KPAX.cu:5-44) and Planner::Planner()
 * (src/planners/Planner.cu:4-35): allocates the tree and work buffers at the
 * configured capacity on cfg->device. */
int kpax_create(const kpax_config* cfg, kpax_handle** out);

/* Batched form (no counterpart in the reference, which plans one query per
 * process: Planner.cuh:19-27).  A handle with `max_queries` query slots and
 * `node_capacity` tree nodes shared by all slots: each admitted query takes a
 * contiguous range of `budget` nodes (its tree cap) and one slot, and every
 * planner iteration advances all admitted queries together.  kpax_create is
 * kpax_create_batch(cfg, 1, cfg->max_tree).  max_queries <= 1024. */
int kpax_create_batch(const kpax_config* cfg, uint32_t max_queries, uint64_t node_capacity,
                      kpax_handle** out);

/* Replaces KPAX::~KPAX() (KPAX.cu:46-49); frees everything. */
int kpax_destroy(kpax_handle* h);

/* Message of the last failure on this handle (never NULL). */
const char* kpax_last_error(const kpax_handle* h);

/* Copy n obstacles [minx,miny,minz,maxx,maxy,maxz] (host array, row format of
 * helper.cu:13-38) into the handle and build the collision grid on the GPU.
 * Replaces the caller's cudaMalloc of d_obstacles (examples/gpu/main.cu:25-26)
 * plus updateSpatialHashGrid (src/collisionCheck/spatialHash.cu:123-139),
 * which the reference re-runs every iteration (KPAX.cu:306). */
int kpax_set_obstacles(kpax_handle* h, const float* aabb6, uint32_t n);

/* Obstacles of environment `env` (< 64) for batched queries; environment 0 is
 * the one kpax_set_obstacles sets.  Every environment shares the configured
 * workspace and grid dimensions. */
int kpax_set_environment(kpax_handle* h, uint32_t env, const float* aabb6, uint32_t n);

/* Plan from start (S floats) to the goal ball goal[0..2] with the configured
 * radius.  Replaces KPAX::plan(h_initial, h_goal, d_obstacles, n, saveTree)
 * (include/planners/KPAX.cuh:14, src/planners/KPAX.cu:51-115), with an
 * explicit seed instead of the wall clock (KPAX.cu:90-91).  On return *out
 * owns its arrays until kpax_result_free. */
int kpax_plan(kpax_handle* h, const float* start, const float* goal, uint64_t seed,
              kpax_result* out);

/* One query of a batch. */
typedef struct kpax_query {
    const float* start;        /* S floats                                                */
    const float* goal;         /* goal ball centre, 3 floats (radius from the config)     */
    uint64_t     seed;         /* a query planned alone with this seed gives the same tree */
    uint32_t     budget;       /* tree cap of this query in nodes; 0 = cfg.max_tree       */
    uint32_t     env;          /* environment index (kpax_set_environment)                */
} kpax_query;

/* Plan n queries on a batched handle.  Queries are admitted in order while a
 * slot and a contiguous node range of their budget are free (first fit);
 * the rest wait until running queries finish and release theirs.  out[i]
 * receives query i's result exactly as kpax_plan would return it for the same
 * start, goal, seed, environment and a tree cap of its budget; plan_ms is the
 * time from the call to the query's completion (queueing included).  A query
 * that can never run (budget above the node capacity, unknown environment,
 * non-finite start) gets status KPAX_INVALID_ARG and the others still run.
 * Every out[i] must be released with kpax_result_free. */
int kpax_plan_batch(kpax_handle* h, const kpax_query* queries, uint32_t n, kpax_result* out);

/* kpax_plan plus the per-iteration dumps of KPAX::planBench
 * (src/planners/KPAX.cu:245-300, writeDeviceVectorsToCSV :544-618): under
 * `dir` (the reference's "Data"), Samples/Samples<bench_itr>/samples<itr>.csv,
 * Parents/..., TotalCountPerVertex/..., ValidCountPerVertex/..., Frontier/...,
 * FrontierRepeatCount/..., VertexScores/..., FrontierSize/..., TreeSize/...,
 * ExpandedNodes/..., ControlPathToGoal/... and ExecutionTime/executionTime.csv,
 * in the reference's CSV formats (helper.cuh:57-108), so its visualisers
 * (viz/KGMT_3D*.m) read them.  Each iteration drains the stream first, as
 * planBench does. */
int kpax_plan_dump(kpax_handle* h, const float* start, const float* goal, uint64_t seed,
                   const char* dir, int32_t bench_itr, kpax_result* out);

/* Free the arrays of a result filled by kpax_plan or kpax_plan_batch. */
void kpax_result_free(kpax_result* r);

/* Copy the tree of the last plan: tree_size nodes, SoA states (S x cap),
 * controls (C x cap), steps, parents, R1 region of each node.  Any pointer
 * may be NULL.  `cap` is the column stride of the SoA outputs and must be
 * >= tree_size.  Replaces the per-iteration dumps of KPAX::planBench
 * (KPAX.cu:545-628). */
int kpax_get_tree(kpax_handle* h, uint32_t cap, float* states, float* controls,
                  int32_t* steps, int32_t* parents, int32_t* regions, uint32_t* tree_size);

/* Copy the region graph of the last plan: valid and total child counts per R1
 * region, covered-R2 count per R1 region, normalised scores per R1 region.
 * Replaces reading graph_.d_validCounterArray_ / d_counterArray_ /
 * d_vertexScoreArray_ (include/graphs/Graph.cuh:25-26). */
int kpax_get_regions(kpax_handle* h, int32_t* valid, int32_t* total, int32_t* covered,
                     float* scores);

/* Profiling of device phases with HIP events (planDebug, KPAX.cu:117-241). */
int kpax_set_profiling(kpax_handle* h, int32_t on);
int kpax_get_stats(kpax_handle* h, kpax_stats* out);
int kpax_reset_stats(kpax_handle* h);

/* ---- launch shape and its autotuner (no counterpart in the reference, whose
 * launch sizes are compile-time constants: KPAX.cu:20-25, config.h:206) ---- */

typedef struct kpax_launch {
    uint32_t expand_blocks;    /* work-groups of the expansion kernel (its waves are persistent) */
    uint32_t expand_chunk;     /* children a wave takes per work ticket (multiple of 16, 16..4096) */
    uint32_t scan_blocks;      /* work-groups of the append and frontier scans                   */
    int32_t  use_graph;        /* 1: replay each iteration as one captured HIP graph (default 0:
                                  direct launches, measured faster)                             */
    int32_t  order_children;   /* 1: hand out children longest drawn edge first (default 1 for a
                                  single-query handle, 0 for a batch handle)                    */
    uint32_t refill_lanes;     /* idle lanes (1..64) before a wave writes out finished children
                                  and takes new ones; 1 services every lane as soon as it idles */
    uint32_t scan_tile;        /* items per append/frontier scan tile: 1024, 2048 or 4096
                                  (blocks of 256, 512 or 1024 threads, 4 items each)            */
    int32_t  lds_grid;         /* 1 (default): single-query expansions and the collision / propagation
                                  operators stage the environment's collision grid in LDS when its image
                                  fits (kpax_get_lds_grid); 0: they query the grid in global memory    */
    uint32_t expand_threads;   /* work-group size of the single-query expansion: a multiple of 64, at most
                                  640 (the batched expansion always runs 256)                          */
    int32_t  wave_priority;    /* 1: an expansion wave raises its issue priority with the steps its
                                  longest live edge has left (s_setprio), so the waves on a launch's
                                  critical path issue first; 0: every wave at the same priority   */
} kpax_launch;

/* Current launch shape of the handle. */
int kpax_get_launch(kpax_handle* h, kpax_launch* out);

/* Set the launch shape.  Results do not depend on it (every plan is the same
 * bit for bit); only the time does. */
int kpax_set_launch(kpax_handle* h, const kpax_launch* launch);

/* Time the n sample queries (as kpax_plan_batch) under candidate launch
 * shapes (expansion work-groups, children per ticket, scan work-groups, graph
 * replay, child ordering), one knob at a time, keeping the fastest of `repeats` runs of each;
 * the best shape is applied to the handle and written to *best.  Every
 * candidate must reproduce the first run's results, or the call fails with
 * KPAX_DEVICE_ERROR. */
int kpax_autotune(kpax_handle* h, const kpax_query* queries, uint32_t n, uint32_t repeats,
                  kpax_launch* best);

/* ---- learned control proposal (no counterpart in the reference, whose
 * controls are uniform draws: statePropagatorSpatialHash.cu:236-241) ------ */

/* Network: 80 features (see kpax_mlp_features) -> 128 GELU -> 128 GELU -> 8
 * outputs: the mean of the 4 controls normalised to [-1, 1], then their log
 * sigma.  Parameters are fp32 in the flat order W1[80][128], b1[128],
 * W2[128][128], b2[128], W3[128][8], b3[8] (W[in][out]); the GPU computes
 * with bf16 copies and fp32 accumulation.  Quadcopter (C = 4) only. */
#define KPAX_MLP_FEATURES 80
#define KPAX_MLP_OUTPUTS  8
#define KPAX_MLP_PARAMS   27912

/* He-uniform initial parameters from a seed (log sigma biases at log 1/2). */
int kpax_mlp_init_params(uint64_t seed, float* params);

/* Use the network as a proposal: each child draws its control from its
 * parent's Gaussian (clamped to the control box) with probability `mix` and
 * uniformly otherwise.  mix = 0 or params = NULL restores the plain planner,
 * whose results are then unchanged bit for bit. */
int kpax_set_proposal(kpax_handle* h, const float* params, float mix);

/* The proposal's Gaussian of every node of the last kpax_plan's tree: 4
 * normalised means then 4 sigmas, SoA with column stride cap (>= tree size),
 * as the expansion samples them (the network's outputs after its mapping).
 * For replaying a proposal plan elsewhere, e.g. on a CPU restatement. */
int kpax_get_proposal(kpax_handle* h, uint32_t cap, float* mean_sigma, uint32_t* tree_size);

/* Budget policy of the proposal: queries whose tree budget (kpax_query.budget,
 * or max_tree for kpax_plan) is below min_budget sample uniformly, exactly as
 * the plain planner; the others use the mixture.  0 (the default) applies it
 * to every query.  Takes effect for queries admitted after the call.  (With
 * the forest-trained network, held-out queries at 2^16-node budgets lose
 * success under any mixture weight, while at 2^17 and above the mixture wins.) */
int kpax_set_proposal_min_budget(kpax_handle* h, uint32_t min_budget);

/* Features of n states (S x n SoA) with their goals (3 x n) in environment
 * `env`: 3 scaled positions, 3 angles / pi, 3 body velocities / vel bound,
 * 3 rates / 4, 3 goal-relative positions, 64 occupancy bits of the 4x4x4
 * window of a 16^3 lattice around the position, one zero.  out: n x 80. */
int kpax_mlp_features(kpax_handle* h, uint32_t env, const float* states_soa,
                      const float* goals_soa, uint32_t n, float* out);

/* Raw network outputs (n x 8) of n feature rows (n x 80) under `params`. */
int kpax_mlp_infer(kpax_handle* h, const float* params, const float* features, uint32_t n,
                   float* out);

/* Controls (n x 4) from given means and sigmas (n x 8: 4 means, 4 sigmas) and
 * four 32-bit random words per item (n x 4): Box-Muller normals, clamp to
 * [-1, 1], map onto the control box.  The sampler the expansion uses. */
int kpax_propose_sample(kpax_handle* h, const float* mean_sigma, const uint32_t* words,
                        uint32_t n, float* controls);

/* Trainer: the network's forward and backward passes and Adam on the GPU. */
typedef struct kpax_trainer kpax_trainer;
int kpax_trainer_create(int32_t device, uint32_t max_batch, kpax_trainer** out);
int kpax_trainer_destroy(kpax_trainer* t);
const char* kpax_trainer_last_error(const kpax_trainer* t);
int kpax_trainer_set_params(kpax_trainer* t, const float* params);
int kpax_trainer_get_params(kpax_trainer* t, float* params);
/* Loss (mean over rows of the Gaussian negative log-likelihood of the
 * targets Y, n x 4 normalised controls, log sigma floored at -4) and its
 * gradient with respect to every parameter, for features X (n x 80). */
int kpax_trainer_gradients(kpax_trainer* t, const float* X, const float* Y, uint32_t n,
                           float* grads, float* loss);
/* Adam over `epochs` shuffled passes of full batches; epoch_loss[epochs]. */
int kpax_trainer_fit(kpax_trainer* t, const float* X, const float* Y, uint32_t n, uint32_t epochs,
                     uint32_t batch, float lr, uint64_t seed, float* epoch_loss);

/* ---- hot-path operators (parity entry points, host buffers) ------------- */

/* Propagate n edges with injected controls and step counts: the body of
 * propagateAndCheckSpatialHash (statePropagatorSpatialHash.cu:72-308) with
 * the cuRAND draws replaced by u_soa (C x n) and steps (n).  x0_soa is S x n.
 * Outputs: x1_soa (S x n, the state after the last step taken; on failure the
 * failing step's state as in the reference), valid (n), fail_step (n: number
 * of the step that failed, 1-based, or 0 when valid). */
int kpax_propagate_batch(kpax_handle* h, const float* x0_soa, const float* u_soa,
                         const int32_t* steps, uint32_t n, float* x1_soa, uint8_t* valid,
                         int32_t* fail_step);

/* Segment-vs-obstacle verdicts through the collision grid for n segment
 * boxes: free[i] = 1 iff the box [bbmin_i, bbmax_i] is strictly separated from
 * every obstacle on some axis.  Replaces isMotionValidSpatialHash
 * (spatialHash.cu:145-194) / isMotionValid (collisionCheck.cu:49-62). */
int kpax_collide_batch(kpax_handle* h, const float* bbmin_soa, const float* bbmax_soa,
                       uint32_t n, uint8_t* free_out);

/* R1 and R2 region index of n states (S x n SoA).  Replaces getRegion /
 * getSubRegion (src/graphs/Graph.cu:99-205). */
int kpax_region_batch(kpax_handle* h, const float* states_soa, uint32_t n, int32_t* r1,
                      int32_t* r2);

/* Region counters of n children taken as one iteration on an empty coverage
 * map: total[r1] += 1 for every child, valid[r1] += 1 for valid ones, and
 * covered[r1] += the number of distinct R2 cells of r1 that valid children
 * reach.  Counters accumulate into the arrays passed in (R1 entries each).
 * Replaces the atomics of propagateFrontier_kernel1 (KPAX.cu:386-396). */
int kpax_region_histogram(kpax_handle* h, const float* states_soa, const uint8_t* valid,
                          uint32_t n, int32_t* total, int32_t* valid_count, int32_t* covered);

/* Normalised R1 scores from counters (R1 entries each).  Replaces
 * Graph::updateVertices (src/graphs/Graph.cu:207-358). */
int kpax_scores(kpax_handle* h, const int32_t* valid_count, const int32_t* total,
                const int32_t* covered, float* scores);

/* Weighted repeat-expansion: out lists index i weights[i] times, in increasing
 * i; *total = sum of weights.  out must hold the total.  Replaces the frontier
 * scan + findInd + repeatInd chain (KPAX.cu:310-321, Planner.cu:52-80). */
int kpax_repeat_expand(kpax_handle* h, const uint32_t* weights, uint32_t n, uint32_t* out,
                       uint32_t out_cap, uint32_t* total);

/* kpax_repeat_expand without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Inputs and outputs are DEVICE buffers: d_weights (n), d_out (out_cap; at
 * most out_cap entries are written), *d_total (the sum of the weights, which
 * must be below 2^32).  The work is enqueued on the handle's stream and not
 * waited for.  The look-back status words live in one library-owned device
 * array per GPU, sized for n <= KPAX_NOALLOC_MAX_ITEMS (larger n is rejected
 * with KPAX_INVALID_ARG): calls that overlap on one GPU, from any handle,
 * race on it. */
#define KPAX_NOALLOC_MAX_ITEMS (1u << 24)
int kpax_repeat_expand_noalloc(kpax_handle* h, const uint32_t* d_weights, uint32_t n, uint32_t* d_out,
                               uint32_t out_cap, uint32_t* d_total);

/* Stream compaction of a flag array: out lists every index i with
 * flags[i] != 0, in increasing i; *count = their number (out must hold
 * out_cap entries; at most out_cap are written).  Replaces the frontier mask
 * scan + findInd (KPAX.cu:310-321 with its +1 fix-up for a set last element,
 * Planner.cu:52-58) and the next-frontier compaction (KPAX.cu:523-525, which
 * drops a set last element; here it is kept).  The repeat-expansion above
 * with weights 0 / 1, on the same kernel. */
int kpax_scan_compact(kpax_handle* h, const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t out_cap,
                      uint32_t* count);

/* Copy the collision grid built by kpax_set_obstacles: cell_start holds
 * G + 1 offsets (cells x fastest, gridToIndex of spatialHash.cuh:57-60),
 * items holds the obstacle ids of every cell in ascending order.  Replaces
 * the fixed 16-slot SpatialHashGrid (spatialHash.cuh:25-37), which drops
 * obstacles past the 16th of a cell (spatialHash.cu:41-48). */
int kpax_get_grid(kpax_handle* h, int32_t* cell_start, int32_t* items, uint32_t items_cap,
                  uint32_t* n_cells, uint32_t* n_items);

/* LDS form of environment `env`'s collision grid: a coarse uniform grid of
 * the workspace whose cells list 16-bit obstacle ids, with the obstacles'
 * boxes, chosen to minimise the mean number of obstacles per cell within the
 * LDS budget (spatialHash.cu:9-54 builds the reference's grid, which its
 * query reads from global memory, :145-194).  dims = {0, 0, 0} and bytes = 0
 * when the environment's image does not fit: its queries then read the CSR
 * grid in global memory. */
int kpax_get_lds_grid(kpax_handle* h, uint32_t env, int32_t* dims, uint32_t* bytes);

/* ---- environments (input formats either side of the path) -------------- */

/* Read an obstacle CSV (rows of 6 comma-separated floats) like
 * readObstaclesFromCSV (src/helper/helper.cu:13-38).  Call with out = NULL to
 * get the count, then with an array of 6 * n floats. */
int kpax_read_obstacles_csv(const char* path, float* out, uint32_t cap, uint32_t* n);

/* Seeded environment generators for the built-in workloads:
 *   1  random boxes: n AABBs, centres U[ws], edges U[0.05,0.15], clipped to the
 *      workspace, rejected if they touch the start point or the goal ball
 *   2  forest: n vertical columns of footprint U[0.1,0.3] over the full height
 *   3  maze: a dividing wall of tiles with one passage a tenth of the workspace
 *      wide (0.1 in the unit cube) plus slabs; lengths scale with the workspace
 * `start` and `goal` (xyz + radius) are kept clear.  Call with out = NULL to
 * get the count. */
int kpax_generate_environment(int32_t kind, uint64_t seed, uint32_t n, const float* ws_min,
                              const float* ws_max, const float* start, const float* goal,
                              float* out, uint32_t cap, uint32_t* n_out);

/* Host FP64 re-verification of a trajectory (rows as in kpax_result): each
 * edge is re-propagated in double from its parent row's stored state and
 * re-checked against all obstacles; the last row must lie in the goal ball.
 * *max_dev = max |FP64 state - stored FP32 state|.  Returns KPAX_OK when the
 * trajectory passes, KPAX_NOT_FOUND when it does not. */
int kpax_verify_path(const kpax_config* cfg, const float* obstacles, uint32_t n_obs,
                     const float* rows, const int32_t* steps, uint32_t n_rows,
                     const float* goal, float* max_dev);

#ifdef __cplusplus
}
#endif
#endif /* KPAX_H */
