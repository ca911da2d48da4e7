// kgmt_launch.hpp — host-side launchers of the kernels in kgmt_kernels.hip.
// All launchers are asynchronous on `stream` and return the launch status.
#pragma once

#include <hip/hip_runtime.h>

#include "kgmt_device.hpp"

namespace kgmt
{

struct Launch
{
    uint32_t propGrid  = 1024;  // persistent workgroups of the propagate kernel
    uint32_t propBlock = kTile; // threads per propagate workgroup (128, 256, 512)
    uint32_t lanes     = 1;     // lanes per child (1; 2 or 4 with 256 threads); children per tile = propBlock / lanes
    uint32_t chunkGrid = 1024;  // workgroups of the update kernel
};

// Scenes: boxes sorted along a Morton curve of their centres, grouped by
// kGroup, written as nGrp group bounds followed by the boxes, each a (lo, hi)
// float4 pair. Built on the device from rows of 6 floats: one scene per query
// (rows + 6 * rowOff, nObs boxes, at pool pair obsOff), or one scene of n
// boxes with nGrp = scene_groups(n) groups.
uint32_t scene_groups(uint32_t n);
uint32_t scene_pairs(uint32_t n);
// rows of 2 * wdim floats (wdim 2: planar boxes, z range [-1, 1])
hipError_t launch_build_scene(const float* rows, uint32_t n, uint32_t nGrp, float4* out, int wdim, hipStream_t s);

// propagate launch shapes that are built
bool launch_supported(uint32_t block, uint32_t lanes);

// planner iteration kernels; `model` = KGMT_MODEL_* when every query shares
// it, -1 otherwise; `parity` = (iteration - 1) & 1 selects the work lists
// scenes of the queries built from `rows` first (null: no obstacles)
hipError_t launch_init(const QueryDesc* qds, QueryState* sts, const Pools& P, const WorkLists& W, const float* rows,
                       hipStream_t s);
// the same for a single query whose descriptor travels as a kernel argument
// (stored into *dq by the kernel)
hipError_t launch_init_one(const QueryDesc& qd, QueryDesc* dq, QueryState* sts, const Pools& P, const WorkLists& W,
                           const float* rows, hipStream_t s);
hipError_t launch_propagate(int model, const QueryDesc* qds, QueryState* sts, const Pools& P, const WorkLists& W,
                            uint32_t parity, uint32_t maxPairs, const Launch& L, hipStream_t s);
// own: the region scores from the region counts (the planner), or given in
// P.score (the frontier update stage)
hipError_t launch_update(const QueryDesc* qds, QueryState* sts, const Pools& P, const WorkLists& W,
                         uint32_t parity, const Launch& L, bool own, hipStream_t s);
hipError_t launch_finish(const QueryDesc* qds, QueryState* sts, const WorkLists& W, uint32_t parity, hipStream_t s);
// goal paths of the solved queries, once a plan has ended
hipError_t launch_paths(const QueryDesc* qds, QueryState* sts, const Pools& P, uint32_t nq, hipStream_t s);

// a query's tree as rows of D floats, parents and costs (entries past the
// tree: 0, -1, 0)
hipError_t launch_tree_rows(const float* treeS, const int32_t* parent, const float* cost, uint32_t cap, int D,
                            uint32_t treeSize, float* rows, int32_t* parents, float* costs, hipStream_t s);

// stage-level kernels (one query, caller buffers)
hipError_t launch_stage_draw(const QueryDesc* qd, int C, uint32_t query, uint32_t itr, uint32_t u0, uint32_t n,
                             float* controls, int32_t* steps, float* ufront, hipStream_t s);
hipError_t launch_stage_propagate(int model, const QueryDesc* qd, const int32_t* parentIdx, const float* parentStates,
                                  uint32_t nParents, const float* controls, const int32_t* steps, uint32_t n,
                                  const float4* obs, uint32_t nGrp, uint32_t nObs, float* child, uint8_t* valid,
                                  int32_t* r1, int32_t* r2, hipStream_t s);
// the planner's per-child path (wave sort, two-phase or one-pass check)
hipError_t launch_stage_propagate_planner(int model, const QueryDesc* qd, const int32_t* parentIdx,
                                          const float* parentStates, uint32_t nParents, const float* controls,
                                          const int32_t* steps, uint32_t n, const float4* obs, uint32_t nGrp,
                                          float* child, uint8_t* valid, int32_t* r1, int32_t* r2, uint8_t* check,
                                          float* box, hipStream_t s);
// frontier update of one iteration on a one-query pool (the caller's tree
// already in P.treeS / parent / cost / frontier, region counts and scores in
// P.validc / P.score): staging kernels, then the planner's k_update;
// info[4] = {appended, multiset size, goal node or -1, re-activated}
hipError_t launch_stage_frontier_update(const QueryDesc* qd, QueryState* st, const Pools& P, const WorkLists& W,
                                        uint32_t itr, uint32_t tree, const float* child, const int32_t* childParent,
                                        const uint8_t* fnext, uint32_t nC, int32_t* info, hipStream_t s);
// allocation-free propagate stage (kgmt_noalloc.hip): the scene in a
// file-scope device array of scratch_pairs() (lo, hi) pairs, whose address on
// the current device scratch_address gives; the descriptor passed by value
uint32_t scratch_pairs();
hipError_t scratch_address(float4** out);
hipError_t launch_stage_propagate_by_value(int model, const QueryDesc& qd, const int32_t* parentIdx,
                                           const float* parentStates, uint32_t nParents, const float* controls,
                                           const int32_t* steps, uint32_t n, const float4* obs, uint32_t nGrp,
                                           float* child, uint8_t* valid, int32_t* r1, int32_t* r2, hipStream_t s);
hipError_t launch_stage_region_update(const QueryDesc* qd, const int32_t* r1, const int32_t* r2, const uint8_t* valid,
                                      const float* ufront, uint32_t n, const float* score, const uint32_t* actPrev,
                                      uint32_t* actNext, int32_t* total, int32_t* validc, uint8_t* fnext,
                                      hipStream_t s);
hipError_t launch_stage_score(const QueryDesc* qd, const int32_t* total, const int32_t* validc, const uint32_t* act,
                              float* score, hipStream_t s);
hipError_t launch_stage_compact(const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t* count, uint32_t* tmp,
                                hipStream_t s);
hipError_t launch_stage_repeat(const uint32_t* ids, const uint32_t* counts, uint32_t n, uint32_t* out,
                               uint32_t outCap, uint32_t* total, uint32_t* tmp, hipStream_t s);

// trajectory refinement (kgmt_adjoint.hip)
hipError_t launch_pack_boxes(const float* rows, uint32_t n, float4* out, hipStream_t s);
hipError_t launch_adjoint(bool refine, const QueryDesc* qd, const kgmt_adjoint_params& ap, uint32_t n, uint32_t maxT,
                          const int32_t* steps, const float* x0, const float* goal, float* u, float* grad, float* ckpt,
                          float* mom, float* vel, const float4* obs, uint32_t nObs, float* cost, hipStream_t s);

}  // namespace kgmt
