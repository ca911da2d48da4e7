/*
 * kgmt.h — C-ABI of the MI355X-native KGMT expansion path.
 *
 * Plain C types only (no HIP, no torch): pointers, sizes and status codes.
 * Every function returns KGMT_OK (0) or a negative KGMT_E_* code; the text
 * of the last error on the calling thread is returned by kgmt_last_error().
 *
 * Reference interfaces replaced (paths relative to the KGMT repository):
 *   kgmt_config            include/config/config.h:209-307 (compile-time macros
 *                          MODEL, MAX_TREE_SIZE, MAX_ITER, STEP_SIZE, GOAL_THRESH,
 *                          W_MIN/W_MAX, V_MIN/V_MAX, W_R1_LENGTH, W_R2_LENGTH,
 *                          EPSILON, quad parameters) become one runtime struct.
 *   kgmt_create            KGMT::KGMT            src/planners/KGMT.cu:4-39
 *                          Planner::Planner      src/planners/Planner.cu:4-35
 *                          Graph::Graph          src/graphs/Graph.cu:5-42
 *   kgmt_plan              KGMT::plan            src/planners/KGMT.cu:41-105
 *                          (virtual Planner::plan, include/planners/Planner.cuh:26)
 *   kgmt_get_tree          public members d_treeSamples_, d_treeSamplesParentIdxs_,
 *                          d_treeSampleCosts_    include/planners/Planner.cuh:37-38
 *   kgmt_get_path          h_controlPathToGoal_  include/planners/Planner.cuh:34,
 *                          filled at src/planners/KGMT.cu:92
 *   kgmt_propagate_collide propagateAndCheck*    src/statePropagator/statePropagator.cu:4-370
 *                          + isMotionValid       src/collisionCheck/collisionCheck.cu:3-10,49-62
 *                          + getRegion/getSubRegion src/graphs/Graph.cu:99-205
 *   kgmt_draw_controls     the curand_uniform draws at statePropagator.cu:78-81,
 *                          149-152, 248-253 and KGMT.cu:254,294,372
 *   kgmt_region_update     counter/flag section of propagateFrontier_kernel1/2
 *                          src/planners/KGMT.cu:249-256, 289-296
 *   kgmt_score             Graph::updateVertices + score kernels src/graphs/Graph.cu:207-358
 *   kgmt_compact           thrust::exclusive_scan + findInd src/planners/KGMT.cu:170-173,
 *                          383-385; src/planners/Planner.cu:52-58
 *   kgmt_repeat            thrust::exclusive_scan + repeatInd src/planners/KGMT.cu:176-181;
 *                          src/planners/Planner.cu:68-80
 *   kgmt_frontier_update   KGMT::updateFrontier + updateFrontier_kernel
 *                          src/planners/KGMT.cu:305-403
 *   kgmt_propagate_collide_planner  as kgmt_propagate_collide, through the
 *                          planner's own per-wave check
 *   kgmt_get_tree_device   public device members d_treeSamples_ptr_,
 *                          d_treeSamplesParentIdxs_ptr_, d_treeSampleCosts_ptr_
 *                          include/planners/Planner.cuh:37-43
 *   kgmt_batch_*           no single counterpart: many KGMT::plan calls served by
 *                          one launch sequence per iteration.
 */
#ifndef KGMT_H
#define KGMT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KGMT_ABI_VERSION 5

/* ---- status codes -------------------------------------------------------- */
enum {
    KGMT_OK            = 0,
    KGMT_E_ARG         = -1, /* bad argument (null pointer, size out of range)  */
    KGMT_E_HIP         = -2, /* a HIP runtime call failed                       */
    KGMT_E_NOMEM       = -3, /* device or host allocation failed                */
    KGMT_E_CAPACITY    = -4, /* input larger than the context was created for   */
    KGMT_E_UNSUPPORTED = -5, /* model / option not built                        */
    KGMT_E_NODEVICE    = -6  /* no HIP device visible                           */
};

/* ---- dynamics models (numbering follows config.h MODEL) ------------------ */
enum {
    KGMT_MODEL_UNICYCLE          = 0, /* statePropagator.cu:13-71 (planar, W_DIM 2) */
    KGMT_MODEL_DOUBLE_INTEGRATOR = 1, /* statePropagator.cu:76-142             */
    KGMT_MODEL_DUBINS_AIRPLANE   = 2, /* statePropagator.cu:147-241 (6D, v state) */
    KGMT_MODEL_QUADCOPTER        = 3, /* statePropagator.cu:246-370            */
    KGMT_MODEL_DUBINS_5D         = 4  /* 5D Dubins airplane, speed is a control */
};

/* double-integrator position update */
enum {
    KGMT_DI_REFERENCE = 0, /* statePropagator.cu:97-99 literally: x += (5v + 2ah)h/6 */
    KGMT_DI_EXACT     = 1  /* textbook RK4 of x'' = a: x += v h + a h^2 / 2          */
};

/* presets for kgmt_config_preset */
enum {
    KGMT_PRESET_REFERENCE_QUAD = 0, /* config.h:209-307 as shipped (quad, W 0..100) */
    KGMT_PRESET_DI6D           = 1, /* 6D double integrator, unit cube, 2^17 nodes  */
    KGMT_PRESET_DUBINS5D       = 2, /* 5D Dubins airplane, unit cube, 2^18 nodes    */
    KGMT_PRESET_QUAD12D        = 3, /* 12D quadcopter, unit cube, 2^19 nodes        */
    KGMT_PRESET_UNICYCLE2D     = 4  /* planar unicycle, unit square, 2^17 nodes     */
};

/* Sample layout (one tree node / one row of the path), as in the reference
 * (statePropagator.cu:139-140, 229-238, 314-320):
 *   [ state (S) | control (C) | duration ]   D = S + C + 1 floats
 *   double integrator : S=6  (x y z vx vy vz)        C=3 (ax ay az)
 *   dubins airplane   : S=6  (x y z yaw pitch v)     C=3 (yawRate pitchRate a)
 *   quadcopter        : S=12 (x y z phi theta psi u v w p q r) C=4 (Zc Lc Mc Nc)
 *   dubins 5D         : S=5  (x y z yaw pitch)       C=3 (yawRate pitchRate speed)
 *   unicycle          : S=4  (x y theta v)           C=2 (a steering)
 *
 * Workspace dimension W: 2 for the unicycle (positions x y), 3 otherwise
 * (x y z). Obstacles are axis-aligned boxes, one row of 2W floats each
 * (min corner then max corner: xmin ymin [zmin] xmax ymax [zmax]), the
 * layout isMotionValid reads (collisionCheck.cu:53-58). Regions tile the
 * workspace only: W_R1_LENGTH^W regions of W_R2_LENGTH^W subregions each.
 */
#define KGMT_MAX_STATE   12
#define KGMT_MAX_CONTROL 4
#define KGMT_MAX_SAMPLE  17

typedef struct kgmt_config {
    int32_t  model;              /* KGMT_MODEL_*                                   */
    int32_t  integrator;         /* KGMT_DI_* (double integrator only)             */
    uint32_t capacity;           /* MAX_TREE_SIZE: tree nodes per query            */
    uint32_t max_iter;           /* MAX_ITER                                       */
    float    dt;                 /* STEP_SIZE: integration step                    */
    int32_t  steps_min;          /* edge duration = steps * dt, steps uniform in   */
    int32_t  steps_max;          /*   [steps_min, steps_max]                       */
    float    goal_thresh;        /* GOAL_THRESH                                    */
    float    ws_min, ws_max;     /* W_MIN, W_MAX: workspace cube bounds            */
    float    v_min, v_max;       /* V_MIN, V_MAX: velocity bounds (DI, dubins, quad body vel) */
    float    ctrl_min[4];        /* per-control uniform draw range                 */
    float    ctrl_max[4];
    float    pitch_min, pitch_max; /* dubins pitch bounds                          */
    int32_t  r1_len;             /* W_R1_LENGTH: regions per workspace axis        */
    int32_t  r2_len;             /* W_R2_LENGTH: subregions per region axis        */
    float    epsilon;            /* EPSILON                                        */
    uint32_t children_per_entry; /* h_activeBlockSize_ (32)                        */
    uint32_t root_repeat;        /* repeat count of the root (5)                   */
    uint32_t repeat_sparse;      /* repeat of a new node in a sparse region (15)   */
    uint32_t repeat_dense;       /* repeat of a new node otherwise (1)             */
    int32_t  sparse_valid_below; /* region is sparse while valid count < this (10) */
    float    quad_mass, quad_ix, quad_iy, quad_iz;
    float    quad_nu, quad_mu, gravity;
    uint32_t query_id;           /* Philox counter word: queries sharing a seed     */
    float    uni_length;         /* UNI_LENGTH: unicycle wheel base (unicycle only) */
    uint32_t reserved;           /* must be 0                                       */
    uint64_t seed;               /* replaces the wall-clock seed of KGMT.cu:80-81  */
} kgmt_config;

typedef struct kgmt_result {
    int32_t  success;     /* 1 if a node within goal_thresh of the goal was added   */
    int32_t  goal_index;  /* tree index of that node (h_pathToGoal_), -1 if none     */
    uint32_t iterations;  /* h_itr_                                                  */
    uint32_t tree_size;   /* h_treeSize_                                             */
    uint32_t path_rows;   /* rows of the goal->root path (depth + 1), 0 if none      */
    float    path_cost;   /* sum of edge durations along the path                    */
    float    path_length; /* Euclidean length of the path's position polyline        */
    float    time_ms;     /* wall time of the plan call (host clock)                  */
    uint64_t expanded;    /* children propagated over all iterations                 */
} kgmt_result;

/* Per-kernel device time accumulated while profiling is on (HIP events
 * around every launch of the plain, non-graph iteration loop). Kernel order:
 * 0 propagate, 1 update (region scores, tree append, re-activation, goal
 * test and next expansion list in one launch), 2 finish (closes the
 * iteration of every query and builds the next iteration's work lists; a
 * single-query context has no finish launch: its last update job does
 * that); slots 3 and 4 are unused. */
#define KGMT_PROFILE_KERNELS 5
typedef struct kgmt_profile {
    uint64_t launches[KGMT_PROFILE_KERNELS];
    double   ms[KGMT_PROFILE_KERNELS];
    uint64_t children;   /* children propagated by the profiled launches */
    uint64_t iterations; /* profiled iterations */
    uint64_t flagged;    /* of those children, the ones flagged for the tree (the
                            propagate kernel writes their rows only) */
} kgmt_profile;

/* Trajectory refinement (differentiable RK4 rollout of the quadcopter ODE).
 * For trajectory k with T steps of dt, controls u_t (t < T), states
 * x_{t+1} = RK4(x_t, u_t), positions p_t = x_t[0..3):
 *   J = w_effort * sum_t |u_t - u_ref|^2
 *     + w_obstacle * sum_{t=1..T} sum_boxes tau * softplus((margin - sd(p_t, box)) / tau)
 *     + w_goal * |p_T - goal|^2
 * sd = signed distance from a point to an axis-aligned box. */
typedef struct kgmt_adjoint_params {
    float    w_effort, w_obstacle, w_goal;
    float    margin, tau;
    float    u_ref[4];     /* control the effort is measured from (e.g. hover thrust) */
    float    lr, beta1, beta2, adam_eps;  /* Adam step of kgmt_refine */
    uint32_t iterations;   /* Adam steps of kgmt_refine */
    int32_t  clamp;        /* clamp controls to cfg ctrl_min/max after each step */
} kgmt_adjoint_params;

typedef struct kgmt_ctx kgmt_ctx;
typedef struct kgmt_batch kgmt_batch;

/* ---- library ------------------------------------------------------------- */
int         kgmt_abi_version(void);
const char* kgmt_last_error(void);
/* Dimensions of a model: S, C, D = S + C + 1. */
int kgmt_model_dims(int32_t model, int32_t* state_dim, int32_t* control_dim, int32_t* sample_dim);
/* Fill *out with one of the KGMT_PRESET_* configurations. */
int kgmt_config_preset(int32_t preset, kgmt_config* out);
/* Validate a configuration without touching the device. */
int kgmt_config_check(const kgmt_config* cfg);

/* ---- single-query planner (KGMT::plan) ---------------------------------- */
int kgmt_create(const kgmt_config* cfg, kgmt_ctx** out);
int kgmt_destroy(kgmt_ctx* ctx);
/* Seed of the next plans (replaces the wall-clock seed of KGMT.cu:80-81). */
int kgmt_set_seed(kgmt_ctx* ctx, uint64_t seed);
/* Replay iterations from a hipGraph captured at the first plan (on != 0) or
 * launch them one by one (0, the default; needed for kgmt_set_profiling). */
int kgmt_set_graph(kgmt_ctx* ctx, int32_t on);
/* Launch shape of the planner's persistent kernels: threads per propagate
 * workgroup (128, 256 or 512), lanes per child (1; or 2 or 4 with 256
 * threads: the lanes split the child's box tests), and the workgroup counts
 * of the propagate and update kernels (0 keeps the current value).
 * Results do not depend on it. No counterpart in the reference (fixed launch
 * shapes at KGMT.cu:183-211). */
int kgmt_set_launch(kgmt_ctx* ctx, uint32_t propagate_block, uint32_t lanes_per_child, uint32_t propagate_grid,
                    uint32_t chunk_grid);
/* Use an existing hipStream_t for every launch of this context (NULL: own stream). */
int kgmt_set_stream(kgmt_ctx* ctx, void* hip_stream);
/* Waits until every launch of this context has finished (including the few
 * that a plan leaves to drain after it returns). No counterpart in the
 * reference (its plan() synchronises the device). */
int kgmt_synchronize(kgmt_ctx* ctx);
/* start, goal: host arrays of D floats (only state[0..S) of start and the
 * W position floats of goal are read).  obstacles: 2W*n_obs floats, one box
 * per row (xmin ymin zmin xmax ymax zmax; xmin ymin xmax ymax for W = 2); a
 * device pointer when obstacles_on_device is non-zero (as in the reference),
 * else host memory. Blocking: returns with the result and the goal path
 * (kgmt_get_path) complete. The device writes them into host memory when the
 * last iteration closes, so the call does not wait for the few launches that
 * were queued past it; they exit at once on the context's stream, and the
 * tree and region getters wait for them. The obstacle rows are not read after
 * the call returns. Device obstacle rows are read on the context's own
 * (non-blocking) stream: work that fills them asynchronously on another
 * stream must be complete before the call (synchronise that stream, or share
 * it with kgmt_set_stream). */
int kgmt_plan(kgmt_ctx* ctx, const float* start, const float* goal, const float* obstacles,
              uint32_t n_obs, int32_t obstacles_on_device, kgmt_result* out);
/* Step-by-step planning, as KGMT::planBench (KGMT.cu:107-165) runs the
 * iterations and dumps the state after each (KGMT.cu:149-153, 405-488):
 * kgmt_plan_begin prepares a plan with the arguments of kgmt_plan without
 * running an iteration; each kgmt_plan_step runs one iteration, waits for it
 * and fills *out with the plan so far (iterations run, tree size, success);
 * *done becomes 1 once the plan has ended (goal reached, full tree or
 * max_iter), and further steps do nothing. The getters (tree, path, regions,
 * frontier) read the state after the last step. Results equal kgmt_plan's. */
int kgmt_plan_begin(kgmt_ctx* ctx, const float* start, const float* goal, const float* obstacles,
                    uint32_t n_obs, int32_t obstacles_on_device);
int kgmt_plan_step(kgmt_ctx* ctx, kgmt_result* out, int32_t* done);
/* Frontier flags (d_frontier_) and repeat counts (d_activeFrontierRepeatCount_:
 * 15 or 1 for a new node, 1 for a re-activated one, 0 otherwise,
 * KGMT.cu:336-339, 373) of the tree nodes after the last iteration. Any
 * pointer may be NULL; *n receives the tree size. */
int kgmt_get_frontier(kgmt_ctx* ctx, uint8_t* frontier, uint32_t* repeat, uint32_t cap, uint32_t* n);
/* Host copies of the tree of the last plan: samples as rows of D floats
 * (row-major, like d_treeSamples_), parents (-1 for the root), costs
 * (distance of the node position to the goal). Any pointer may be NULL.
 * cap: rows available; *n receives the tree size. */
int kgmt_get_tree(kgmt_ctx* ctx, float* samples, int32_t* parents, float* costs,
                  uint32_t cap, uint32_t* n);
/* The tree of the last plan in device memory, in the reference's layout
 * (public members d_treeSamples_ (rows of D floats), d_treeSamplesParentIdxs_
 * and d_treeSampleCosts_, include/planners/Planner.cuh:37-43): library-owned
 * buffers of `capacity` entries, entries past tree_size 0, -1 and 0 as the
 * reference's fills leave them (KGMT.cu:62-64). The rows are written by a
 * device kernel on the first call after a plan; the call returns when they are
 * complete, and they stay valid until the next plan or kgmt_destroy. */
typedef struct kgmt_device_tree {
    float*   samples;    /* [capacity][D] */
    int32_t* parents;    /* [capacity], -1 for the root */
    float*   costs;      /* [capacity], distance of the node position to the goal */
    uint32_t tree_size;
    uint32_t capacity;
    int32_t  sample_dim; /* D */
    uint32_t reserved;
} kgmt_device_tree;
int kgmt_get_tree_device(kgmt_ctx* ctx, kgmt_device_tree* out);
/* Rows goal->root of the last successful plan (D floats each). */
int kgmt_get_path(kgmt_ctx* ctx, float* rows, uint32_t cap_rows, uint32_t* n_rows);
/* Per-region counters of the last plan: total[R1], valid[R1], score[R1]. */
int kgmt_get_regions(kgmt_ctx* ctx, int32_t* total, int32_t* valid, float* score, uint32_t cap,
                     uint32_t* n_regions);

/* Event timing of the plain iteration loop: 0 off, 1 propagate launches only
 * (two events per iteration), 2 every kernel. Resets the accumulators. */
int kgmt_set_profiling(kgmt_ctx* ctx, int32_t on);
int kgmt_get_profile(kgmt_ctx* ctx, kgmt_profile* out);

/* ---- batched planner ----------------------------------------------------- */
/* n_q queries, each with its own configuration (capacity, model, ...) and its
 * own obstacle set. obstacle_offsets[q] (in boxes) indexes `obstacles`;
 * obstacle_counts[q] boxes belong to query q. starts/goals: n_q rows of
 * KGMT_MAX_SAMPLE floats. All inputs are host memory. Every query of a batch
 * has the same workspace dimension W (rows of 2W floats); kgmt_batch_create
 * rejects a mix. */
int kgmt_batch_create(const kgmt_config* cfgs, uint32_t n_q, kgmt_batch** out);
int kgmt_batch_destroy(kgmt_batch* b);
int kgmt_batch_set_stream(kgmt_batch* b, void* hip_stream);
/* use_graph != 0: one iteration is captured once as a hipGraph and replayed. */
int kgmt_batch_plan(kgmt_batch* b, const float* starts, const float* goals, const float* obstacles,
                    const uint32_t* obstacle_offsets, const uint32_t* obstacle_counts,
                    int32_t use_graph, kgmt_result* out /* n_q */, float* finish_ms /* n_q or NULL */);
int kgmt_batch_get_path(kgmt_batch* b, uint32_t q, float* rows, uint32_t cap_rows, uint32_t* n_rows);
int kgmt_batch_get_tree(kgmt_batch* b, uint32_t q, float* samples, int32_t* parents, float* costs,
                        uint32_t cap, uint32_t* n);
/* Launch shape of the persistent kernels, as kgmt_set_launch. */
int kgmt_batch_set_launch(kgmt_batch* b, uint32_t propagate_block, uint32_t lanes_per_child,
                          uint32_t propagate_grid, uint32_t chunk_grid);
int kgmt_batch_set_profiling(kgmt_batch* b, int32_t on);
int kgmt_batch_get_profile(kgmt_batch* b, kgmt_profile* out);

/* ---- stage-level entry points (device pointers, asynchronous on `stream`) -
 * These run one stage of an iteration on caller-owned device buffers so that
 * each stage can be checked on identical inputs. SoA layout: element (d, i)
 * of an [X][n] array lives at d * n + i. */

/* Philox draws of iteration `itr` for children u = u0 .. u0+n-1 of query
 * `query`: controls [C][n], steps [n], frontier uniform [n]. */
int kgmt_draw_controls(const kgmt_config* cfg, uint32_t query, uint32_t itr, uint32_t u0, uint32_t n,
                       float* d_controls, int32_t* d_steps, float* d_ufront, void* stream);
/* Obstacle scene as the propagate kernels read it, built from n rows of 6
 * floats (xmin ymin zmin xmax ymax zmax, collisionCheck.cu:53-58; the planner
 * stores a planar box as z range [-1, 1], which contains every planar
 * position's z = 0, so the 3D test gives the planar verdict): boxes in
 * Morton order of their centres in groups of 8, n_groups group bounds first;
 * each entry is a (lo, hi) pair of 4-float vectors, so the scene takes
 * 8 * n_groups * 9 floats. Returns KGMT_E_CAPACITY (n_groups still set) when
 * out_floats is too small. No counterpart in the reference (it tests every
 * box in input order). */
int kgmt_build_scene(const float* d_rows, uint32_t n, float* d_out, uint32_t out_floats, uint32_t* n_groups,
                     void* stream);
/* Propagate + collide: child i starts from parent_states[:, parent_idx[i]]
 * ([S][n_parents]), applies controls[:, i] for steps[i] steps. Writes child
 * samples [D][n] (state, control, duration), valid[n], r1[n], r2[n].
 * d_obstacles: n_obs rows of 2W floats, as kgmt_plan. */
int kgmt_propagate_collide(const kgmt_config* cfg, const int32_t* d_parent_idx,
                           const float* d_parent_states, uint32_t n_parents,
                           const float* d_controls, const int32_t* d_steps, uint32_t n_children,
                           const float* d_obstacles, uint32_t n_obs, float* d_child_samples,
                           uint8_t* d_valid, int32_t* d_r1, int32_t* d_r2, void* stream);
/* kgmt_propagate_collide without any device allocation: no counterpart in
 * the reference; for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). Same arguments,
 * outputs and bits as kgmt_propagate_collide; the obstacle scene goes into a
 * library-held device array of one scene of at most 1500 boxes per GPU
 * (54 KB, kept until the process ends), and n_obs > 1500 is rejected with
 * KGMT_E_CAPACITY. Calls on one GPU share that array: calls that overlap on
 * one GPU race. Asynchronous on `stream`. */
int kgmt_propagate_collide_noalloc(const kgmt_config* cfg, const int32_t* d_parent_idx,
                                   const float* d_parent_states, uint32_t n_parents,
                                   const float* d_controls, const int32_t* d_steps, uint32_t n_children,
                                   const float* d_obstacles, uint32_t n_obs, float* d_child_samples,
                                   uint8_t* d_valid, int32_t* d_r1, int32_t* d_r2, void* stream);
/* Checks the planner's propagate path can decide a child's verdict with
 * (kgmt_propagate_collide_planner reports which one ran). */
enum {
    KGMT_CHECK_SCENE      = 0, /* every step against every box of the scene          */
    KGMT_CHECK_CANDIDATES = 1, /* two-phase: the wave's candidate list, each lane its  */
                               /* own candidates inside its edge box                  */
    KGMT_CHECK_DRY_RUN    = 2, /* two-phase: no box near the wave, the dry run decides */
    KGMT_CHECK_ONE_PASS   = 3  /* quadcopter: one pass against every box             */
};
/* Propagate + collide through the planner's own per-child path, for checking
 * it on identical inputs: children in waves of 64 consecutive indices, sorted
 * within the wave by step count (longest first) as the planner's propagate
 * kernel sorts them, then the check the planner runs for the model: the
 * two-phase check for the double integrator, both Dubins models and the
 * unicycle (an edge box from a dry run, or for the Dubins models from an
 * interval bound of the edge; the wave's candidate boxes = boxes meeting some
 * lane's edge box; each lane tests only its own candidates; the whole scene
 * when a box is not finite or the candidates exceed 128), the one-pass check
 * for the quadcopter. Same arguments and outputs as kgmt_propagate_collide
 * (isMotionValid, collisionCheck.cu:49-62, per step), plus two optional
 * outputs (NULL: not written): d_check[n], the KGMT_CHECK_* that decided
 * each child, and d_edge_box [6][n], each child's edge box (lo xyz, hi xyz;
 * not written for the quadcopter). */
int kgmt_propagate_collide_planner(const kgmt_config* cfg, const int32_t* d_parent_idx,
                                   const float* d_parent_states, uint32_t n_parents,
                                   const float* d_controls, const int32_t* d_steps, uint32_t n_children,
                                   const float* d_obstacles, uint32_t n_obs, float* d_child_samples,
                                   uint8_t* d_valid, int32_t* d_r1, int32_t* d_r2, uint8_t* d_check,
                                   float* d_edge_box, void* stream);
/* Frontier update of one iteration (updateFrontier + updateFrontier_kernel,
 * KGMT.cu:305-403) through the planner's own update kernel, on
 * caller buffers (device memory, capacity = cfg->capacity):
 *   tree: samples [D][capacity] (SoA), parents, costs [capacity], frontier
 *   flags [capacity] (u8); the first tree_size entries are the tree;
 *   the iteration's children: samples [D][n_children], parent (tree index),
 *   next-frontier flag (u8) [n_children];
 *   region valid counts and scores [R1] after the iteration's propagate;
 *   goal: host array of W position floats; itr: the iteration (1-based).
 * Writes: the flagged children in child order, clamped to the free capacity,
 * appended at tree_size.. (sample, parent, cost = distance to the goal,
 * frontier 1); frontier 1 for every old node with frontier 0 whose draw U <=
 * score[region] + fAccept; rep_idx [capacity]: the next expansion multiset
 * (re-activated nodes ascending, then each new node repeated 15 times when
 * its region's valid count < 10, else once), its first capacity entries;
 * info [4] (int32): appended count, multiset size, the smallest new index
 * within goal_thresh or -1, re-activated count. Synchronous on `stream`. */
int kgmt_frontier_update(const kgmt_config* cfg, uint32_t itr, uint32_t tree_size, float* d_tree_samples,
                         int32_t* d_parents, float* d_costs, uint8_t* d_frontier, const float* d_child_samples,
                         const int32_t* d_child_parent, const uint8_t* d_fnext, uint32_t n_children,
                         const int32_t* d_valid_count, const float* d_score, const float* goal, uint32_t* d_rep_idx,
                         int32_t* d_info, void* stream);
/* Region bookkeeping of one iteration: total[r1]++ for every child,
 * valid[r1]++ and active_next[r2]=1 for valid ones, and
 * fnext[i] = valid && (ufront[i] < score[r1] || !active_prev[r2]).
 * active_* are bitsets of R2 bits (uint32 words); active_next must hold
 * active_prev on entry. total/valid accumulate. */
int kgmt_region_update(const kgmt_config* cfg, const int32_t* d_r1, const int32_t* d_r2,
                       const uint8_t* d_valid, const float* d_ufront, uint32_t n,
                       const float* d_score, const uint32_t* d_active_prev, uint32_t* d_active_next,
                       int32_t* d_total, int32_t* d_valid_count, uint8_t* d_fnext, void* stream);
/* Region scores from total/valid counters and the subregion bitset. */
int kgmt_score(const kgmt_config* cfg, const int32_t* d_total, const int32_t* d_valid_count,
               const uint32_t* d_active, float* d_score, void* stream);
/* Stream compaction: out[k] = k-th index i (ascending) with flags[i] != 0;
 * *d_count = number of such indices. */
int kgmt_compact(const uint8_t* d_flags, uint32_t n, uint32_t* d_out, uint32_t* d_count, void* stream);
/* Segmented repeat: node ids[i] written counts[i] times, in order, into out;
 * *d_total = sum of counts. */
int kgmt_repeat(const uint32_t* d_ids, const uint32_t* d_counts, uint32_t n, uint32_t* d_out,
                uint32_t out_cap, uint32_t* d_total, void* stream);

/* ---- trajectory refinement (quadcopter RK4 adjoint) ----------------------
 * No counterpart in the reference planner: an extension that differentiates
 * its quadcopter rollout (statePropagator.cu:246-370) to refine control
 * sequences. Device buffers, SoA over trajectories k < n:
 *   steps [n] int32 (T_k <= max_steps), x0 [12][n], goal [3][n],
 *   u / grad [max_steps][4][n] (element (t, c, k) at (t*4 + c)*n + k),
 *   obstacles: 6 floats per box (shared by all trajectories).
 * cfg supplies the quad parameters, dt and (for clamping) ctrl_min/max. */
/* J_k -> cost[n] and dJ_k/du -> grad (zero past T_k). */
int kgmt_rk4_adjoint(const kgmt_config* cfg, const kgmt_adjoint_params* ap, uint32_t n, uint32_t max_steps,
                     const int32_t* d_steps, const float* d_x0, const float* d_goal, const float* d_u,
                     const float* d_obstacles, uint32_t n_obs, float* d_cost, float* d_grad, void* stream);
/* ap->iterations Adam steps on u in place; cost_history [(iterations+1)][n]
 * holds J before every step and after the last. */
int kgmt_refine(const kgmt_config* cfg, const kgmt_adjoint_params* ap, uint32_t n, uint32_t max_steps,
                const int32_t* d_steps, const float* d_x0, const float* d_goal, float* d_u,
                const float* d_obstacles, uint32_t n_obs, float* d_cost_history, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* KGMT_H */
