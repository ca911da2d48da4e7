/*
 * kgmt_oracle.h — CPU restatement of KGMT's per-iteration expansion path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker that the HIP path is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it. The product (kgmt_amd/, libkgmt) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the KGMT repository). Where the reference is racy or undefined, the oracle
 * implements the deterministic semantics documented beside the function.
 */
#ifndef KGMT_ORACLE_H
#define KGMT_ORACLE_H

#include <stdint.h>
#include "../include/kgmt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* counter-based RNG ------------------------------------------------------- */
void  orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_u01(uint32_t x);
void  orc_sincosf(float x, float* s, float* c);

/* per-child draws --------------------------------------------------------- */
void orc_draw_child(const kgmt_config* cfg, uint32_t query, uint32_t itr, uint32_t u,
                    float* ctrl, int32_t* steps, float* ufront);
float orc_draw_react(const kgmt_config* cfg, uint32_t query, uint32_t itr, uint32_t node);

/* geometry ---------------------------------------------------------------- */
int orc_segment_valid(const float* p0, const float* p1, const float* obs, uint32_t n_obs);
int orc_segment_valid2(const float* p0, const float* p1, const float* obs, uint32_t n_obs);
int orc_wdim(const kgmt_config* cfg);
int orc_region(const kgmt_config* cfg, const float* pos);
int orc_subregion(const kgmt_config* cfg, const float* pos, int r1);
int orc_num_regions(const kgmt_config* cfg, int32_t* r1, int32_t* r2, int32_t* per_r1);

/* dynamics: propagate one child from x0 (S floats) with given controls ----- */
int orc_propagate(const kgmt_config* cfg, const float* x0, const float* ctrl, int32_t steps,
                  const float* obs, uint32_t n_obs, float* x1 /* D floats */);
int orc_propagate_f64(const kgmt_config* cfg, const double* x0, const double* ctrl, int32_t steps,
                      const double* obs, uint32_t n_obs, double* x1);
void orc_quad_ode_f64(const kgmt_config* cfg, const double* x, const double* ctrl, double* xdot);

/* region bookkeeping ------------------------------------------------------ */
void orc_region_update(const kgmt_config* cfg, const int32_t* r1, const int32_t* r2, const uint8_t* valid,
                       const float* ufront, uint32_t n, const float* score, const uint8_t* active_prev,
                       uint8_t* active_next, int32_t* total, int32_t* valid_count, uint8_t* fnext);
void orc_score(const kgmt_config* cfg, const int32_t* total, const int32_t* valid_count,
               const uint8_t* active, float* score);
float orc_faccept(const kgmt_config* cfg, uint32_t itr, uint32_t tree_size, uint32_t next);

/* frontier update of one iteration (updateFrontier + updateFrontier_kernel,
 * KGMT.cu:305-403) on the planner's state after propagate and score:
 * samples [capacity][D], parents, costs, frontier [capacity] (the first
 * tree_size entries are the tree; accepted children are appended after
 * them), children child [n_children][D], parentU, fnext [n_children], region
 * valid counts and scores [R1]. Outputs: the next expansion multiset's first
 * capacity entries in M (re-activated nodes ascending, then every new node
 * repeated 15 or 1 times) and its length nM, the full multiset size R, the
 * appended count, the smallest new node index within the goal threshold (or
 * -1) and the number of re-activated nodes. */
int orc_frontier_update(const kgmt_config* cfg, uint32_t itr, uint32_t tree_size, float* samples, int32_t* parents,
                        float* costs, uint8_t* frontier, const float* child, const int32_t* parentU,
                        const uint8_t* fnext, uint32_t n_children, const int32_t* valid_count, const float* score,
                        const float* goal, uint32_t* M, uint32_t* nM, uint64_t* R, uint32_t* next,
                        int32_t* goal_index, uint32_t* n_react);

/* full planner ------------------------------------------------------------ */
typedef struct orc_stats {
    uint32_t iterations;
    uint32_t tree_size;
    int32_t  goal_index;
    uint32_t path_rows;
    float    path_cost;
    float    path_length;
    uint64_t expanded;
} orc_stats;

/* Random draws use cfg->query_id as the query word of the Philox counter.
 * samples: capacity*D floats (row-major), parents/costs: capacity entries,
 * path: (max_iter+1)*D floats. iter_log (optional, may be NULL): per
 * iteration 4 uint32 {R, n_children, next, n_react}. */
int orc_plan(const kgmt_config* cfg, const float* start, const float* goal,
             const float* obs, uint32_t n_obs, float* samples, int32_t* parents, float* costs,
             float* path, int32_t* total_out, int32_t* valid_out, float* score_out,
             uint32_t* iter_log, orc_stats* st);

/* Plan n_q queries (configuration cfgs[q]) on `threads` OpenMP
 * threads, one query per thread at a time; the CPU baseline. starts/goals:
 * n_q rows of KGMT_MAX_SAMPLE floats. Per-query wall ms in ms_out. */
int orc_plan_many(const kgmt_config* cfgs, uint32_t n_q, const float* starts, const float* goals,
                  const float* obs, const uint32_t* obs_off, const uint32_t* obs_cnt, int threads,
                  orc_stats* st_out, double* ms_out);

/* trajectory cost and its gradient (FP64): the quadcopter ODE's
 * vector-Jacobian product, the cost of one control sequence and its
 * gradient by the discrete adjoint of the RK4 rollout. u: T x 4 row-major. */
void orc_quad_vjp_f64(const kgmt_config* cfg, const double* s, const double* c, const double* a, double* gs,
                      double* gc);
double orc_traj_cost_f64(const kgmt_config* cfg, const kgmt_adjoint_params* ap, int32_t T, const double* x0,
                         const double* u, const double* goal, const double* obs, uint32_t n_obs, double* xT);
double orc_traj_grad_f64(const kgmt_config* cfg, const kgmt_adjoint_params* ap, int32_t T, const double* x0,
                         const double* u, const double* goal, const double* obs, uint32_t n_obs, double* grad);

/* batch conveniences (SoA layouts of the stage-level C-ABI) */
void orc_sincosf_n(const float* x, uint32_t n, float* s, float* c);
void orc_philox_n(const uint32_t* ctr, const uint32_t* key, uint32_t n, uint32_t* out);
void orc_draw_children(const kgmt_config* cfg, uint32_t query, uint32_t itr, uint32_t u0, uint32_t n,
                       float* controls, int32_t* steps, float* ufront);
int orc_propagate_children(const kgmt_config* cfg, const int32_t* parent_idx, const float* parent_states,
                           uint32_t n_parents, const float* controls, const int32_t* steps, uint32_t n,
                           const float* obs, uint32_t n_obs, float* child, uint8_t* valid, int32_t* r1,
                           int32_t* r2);
int orc_propagate_children_f64(const kgmt_config* cfg, const double* x0s, const double* ctrls,
                               const int32_t* steps, uint32_t n, const double* obs, uint32_t n_obs,
                               double* x1s, uint8_t* valid);
void orc_segment_valid_n(const float* p0, const float* p1, uint32_t n, const float* obs, uint32_t n_obs,
                         uint8_t* out);
void orc_region_n(const kgmt_config* cfg, const float* pos, uint32_t n, int32_t* r1, int32_t* r2);

#ifdef __cplusplus
}
#endif
#endif
