// kgmt_planner.hpp — C++ drop-in for the reference planner classes, over the
// C-ABI of kgmt.h (link with libkgmt.so; header-only).
//
// Mirrors include/planners/Planner.cuh:19-44 and include/planners/KGMT.cuh:5-36:
// same class names, same plan(float*, float*, float*, uint) signature taking a
// DEVICE pointer to the obstacle rows, same public result members
// (h_treeSize_, h_itr_, h_pathToGoal_, h_controlPathToGoal_, h_costToGoal_).
// Differences a caller can see:
//   * the configuration is a runtime kgmt_config (config.h's macros); the
//     default constructor uses config.h's quadcopter block;
//   * failures throw kgmt::Error instead of continuing silently;
//   * h_controlPathToGoal_ holds h_pathRows_ rows (the exact path depth);
//   * tree arrays are copied to the host on request (treeSamples() ...),
//     the device buffers are owned by the library.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "kgmt.h"

namespace kgmt
{
struct Error : std::runtime_error
{
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc)
{
    if(rc != KGMT_OK) throw Error(rc, std::string("kgmt: ") + kgmt_last_error());
}
}  // namespace kgmt

typedef unsigned int uint;

// include/planners/Planner.cuh:19-44
class Planner
{
public:
    virtual ~Planner() = default;
    virtual void plan(float* h_initial, float* h_goal, float* d_obstacles_ptr, uint h_obstaclesCount) = 0;

    uint h_treeSize_ = 0, h_itr_ = 0;
    float h_costToGoal_ = 0.0f;
    int h_pathToGoal_ = 0;
    std::vector<float> h_controlPathToGoal_;  // rows goal -> root, SAMPLE_DIM floats each
    uint h_pathRows_ = 0;
};

// include/planners/KGMT.cuh:5-36
class KGMT : public Planner
{
public:
    KGMT() : KGMT(preset(KGMT_PRESET_REFERENCE_QUAD)) {}
    explicit KGMT(const kgmt_config& cfg) : cfg_(cfg)
    {
        kgmt::check(kgmt_model_dims(cfg.model, &S_, &C_, &D_));
        kgmt::check(kgmt_create(&cfg_, &ctx_));
    }
    ~KGMT() override { kgmt_destroy(ctx_); }
    KGMT(const KGMT&)            = delete;
    KGMT& operator=(const KGMT&) = delete;

    static kgmt_config preset(int32_t p)
    {
        kgmt_config c;
        kgmt::check(kgmt_config_preset(p, &c));
        return c;
    }

    // KGMT::plan, src/planners/KGMT.cu:41-105. d_obstacles_ptr: device memory,
    // 2 * W_DIM floats per box (xmin ymin zmin xmax ymax zmax; xmin ymin xmax
    // ymax for the planar unicycle).
    void plan(float* h_initial, float* h_goal, float* d_obstacles_ptr, uint h_obstaclesCount) override
    {
        kgmt::check(kgmt_plan(ctx_, h_initial, h_goal, d_obstacles_ptr, h_obstaclesCount, 1, &result_));
        h_treeSize_   = result_.tree_size;
        h_itr_        = result_.iterations;
        h_pathToGoal_ = result_.success ? result_.goal_index : 0;
        h_costToGoal_ = result_.success ? result_.path_cost : 0.0f;
        h_pathRows_   = result_.path_rows;
        h_controlPathToGoal_.assign((size_t)h_pathRows_ * D_, 0.0f);
        if(h_pathRows_)
            {
                uint32_t n = 0;
                kgmt::check(kgmt_get_path(ctx_, h_controlPathToGoal_.data(), h_pathRows_, &n));
            }
    }

    // KGMT::planBench, src/planners/KGMT.cu:107-165 (without the per-iteration
    // CSV dump; use treeSamples()/treeParents() after the plan instead)
    void planBench(float* h_initial, float* h_goal, float* d_obstacles_ptr, uint h_obstaclesCount, int /*benchItr*/)
    {
        plan(h_initial, h_goal, d_obstacles_ptr, h_obstaclesCount);
    }

    void setSeed(uint64_t seed) { kgmt::check(kgmt_set_seed(ctx_, seed)); }
    // plan() returns with its results complete; this also waits for the few
    // launches it leaves to drain on the planner's stream
    void synchronize() { kgmt::check(kgmt_synchronize(ctx_)); }

    // host copies of d_treeSamples_ (row-major, D floats per node),
    // d_treeSamplesParentIdxs_, d_treeSampleCosts_
    std::vector<float> treeSamples() const
    {
        std::vector<float> s((size_t)h_treeSize_ * D_);
        uint32_t n = 0;
        kgmt::check(kgmt_get_tree(ctx_, s.data(), nullptr, nullptr, h_treeSize_, &n));
        return s;
    }
    std::vector<int32_t> treeParents() const
    {
        std::vector<int32_t> p(h_treeSize_);
        uint32_t n = 0;
        kgmt::check(kgmt_get_tree(ctx_, nullptr, p.data(), nullptr, h_treeSize_, &n));
        return p;
    }
    std::vector<float> treeCosts() const
    {
        std::vector<float> c(h_treeSize_);
        uint32_t n = 0;
        kgmt::check(kgmt_get_tree(ctx_, nullptr, nullptr, c.data(), h_treeSize_, &n));
        return c;
    }

    const kgmt_result& result() const { return result_; }
    const kgmt_config& config() const { return cfg_; }
    int sampleDim() const { return D_; }

private:
    kgmt_config cfg_;
    kgmt_ctx* ctx_ = nullptr;
    kgmt_result result_{};
    int32_t S_ = 0, C_ = 0, D_ = 0;
};
