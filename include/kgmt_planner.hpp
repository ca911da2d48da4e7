// kgmt_planner.hpp — C++ drop-in for the reference planner classes, over the
// C-ABI of kgmt.h (link with libkgmt.so; header-only).
//
// Mirrors include/planners/Planner.cuh:19-44 and include/planners/KGMT.cuh:5-36:
// same class names, same plan(float*, float*, float*, uint) signature taking a
// DEVICE pointer to the obstacle rows, same public result members
// (h_treeSize_, h_itr_, h_pathToGoal_, h_costToGoal_, h_controlPathToGoal_
// as a float* the object owns, d_treeSamples_ptr_, d_treeSamplesParentIdxs_ptr_,
// d_treeSampleCosts_ptr_ in device memory), plus the helpers a reference
// caller uses: readObstaclesFromCSV (src/helper/helper.cu:13-39) and the
// dimension macros of config.h (STATE_DIM, CONTROL_DIM, SAMPLE_DIM, W_DIM).
// Differences a caller can see:
//   * the configuration is a runtime kgmt_config (config.h's macros); the
//     default constructor uses the preset named by KGMT_DROPIN_PRESET (0, the
//     reference's quadcopter block of config.h:209-307, unless defined before
//     this header), and the dimension macros follow the same preset;
//   * failures throw kgmt::Error instead of continuing silently (the
//     reference's CSV reader calls exit(1); an uncaught kgmt::Error also ends
//     the program with an error);
//   * h_controlPathToGoal_ holds h_pathRows_ rows (the exact path depth);
//   * the device tree members are raw pointers into library-owned buffers
//     (there are no thrust vectors), valid after plan() until the next plan;
//     setDeviceTree(false) skips writing them.
#pragma once

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "kgmt.h"

// config.h dimension macros of the drop-in's preset (KGMT_PRESET_* number)
#ifndef KGMT_DROPIN_PRESET
#define KGMT_DROPIN_PRESET 0
#endif
#if KGMT_DROPIN_PRESET == 0 || KGMT_DROPIN_PRESET == 3
#define KGMT_DROPIN_STATE_DIM 12
#define KGMT_DROPIN_CONTROL_DIM 4
#define KGMT_DROPIN_W_DIM 3
#elif KGMT_DROPIN_PRESET == 1
#define KGMT_DROPIN_STATE_DIM 6
#define KGMT_DROPIN_CONTROL_DIM 3
#define KGMT_DROPIN_W_DIM 3
#elif KGMT_DROPIN_PRESET == 2
#define KGMT_DROPIN_STATE_DIM 5
#define KGMT_DROPIN_CONTROL_DIM 3
#define KGMT_DROPIN_W_DIM 3
#elif KGMT_DROPIN_PRESET == 4
#define KGMT_DROPIN_STATE_DIM 4
#define KGMT_DROPIN_CONTROL_DIM 2
#define KGMT_DROPIN_W_DIM 2
#else
#error "KGMT_DROPIN_PRESET must be a KGMT_PRESET_* number (0..4)"
#endif
#ifndef STATE_DIM
#define STATE_DIM KGMT_DROPIN_STATE_DIM
#endif
#ifndef CONTROL_DIM
#define CONTROL_DIM KGMT_DROPIN_CONTROL_DIM
#endif
#ifndef SAMPLE_DIM
#define SAMPLE_DIM (STATE_DIM + CONTROL_DIM + 1)
#endif
#ifndef W_DIM
#define W_DIM KGMT_DROPIN_W_DIM
#endif

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

namespace kgmt
{
// include/helper/helper.cuh:56-85: rows x cols values, comma separated, fixed
// six decimals for floating point (integers as they are), appended or new
template <typename T>
inline void writeVectorToCSV(const std::vector<T>& vec, const std::string& filename, int rows, int cols,
                             bool append = false)
{
    std::ofstream file;
    if(append)
        file.open(filename, std::ios_base::app);
    else
        file.open(filename);
    file << std::fixed << std::setprecision(6);
    for(int i = 0; i < rows; i++)
        {
            for(int j = 0; j < cols; j++)
                {
                    file << vec[(size_t)i * cols + j];
                    if(j < cols - 1) file << ",";
                }
            file << std::endl;
        }
}

// include/helper/helper.cuh:95-109: one value appended, ten decimals for
// floating point
template <typename T>
inline void writeValueToCSV(const T& value, const std::string& filename)
{
    std::ofstream file;
    file.open(filename, std::ios_base::app);
    if constexpr(std::is_floating_point_v<std::decay_t<T>>) file << std::fixed << std::setprecision(10);
    file << value << std::endl;
}
}  // namespace kgmt

// src/helper/helper.cu:13-39: every number of the file in order (a comma
// after a number is skipped; a line stops at the first token that is not a
// number), numObstacles = values / (2 * workspaceDim). A file that cannot be
// opened throws kgmt::Error (the reference prints and calls exit(1)).
inline std::vector<float> readObstaclesFromCSV(const std::string& filename, int& numObstacles, int workspaceDim)
{
    std::vector<float> obstacles;
    std::ifstream file(filename);
    if(!file.is_open())
        {
            std::cerr << "Error opening file: " << filename << std::endl;
            throw kgmt::Error(KGMT_E_ARG, "cannot open obstacle file " + filename);
        }
    std::string line;
    while(std::getline(file, line))
        {
            std::istringstream ss(line);
            float value;
            while(ss >> value)
                {
                    obstacles.push_back(value);
                    if(ss.peek() == ',') ss.ignore();
                }
        }
    numObstacles = (int)(obstacles.size() / (size_t)(2 * workspaceDim));
    return obstacles;
}

// include/planners/Planner.cuh:19-44
class Planner
{
public:
    Planner()                          = default;
    Planner(const Planner&)            = delete;
    Planner& operator=(const Planner&) = delete;
    virtual ~Planner() { delete[] h_controlPathToGoal_; }
    virtual void plan(float* h_initial, float* h_goal, float* d_obstacles_ptr, uint h_obstaclesCount) = 0;

    // --- host fields ---
    uint h_treeSize_ = 0, h_itr_ = 0;
    float h_costToGoal_ = 0.0f;
    int h_pathToGoal_ = 0;
    float* h_controlPathToGoal_ = nullptr;  // rows goal -> root, SAMPLE_DIM floats each (h_pathRows_ of them)
    uint h_pathRows_ = 0;

    // --- device fields (rows of SAMPLE_DIM floats, parent indices, costs) ---
    float* d_treeSamples_ptr_ = nullptr;
    int* d_treeSamplesParentIdxs_ptr_ = nullptr;
    float* d_treeSampleCosts_ptr_ = nullptr;
};

// include/planners/KGMT.cuh:5-36
class KGMT : public Planner
{
public:
    KGMT() : KGMT(preset(KGMT_DROPIN_PRESET)) {}
    explicit KGMT(const kgmt_config& cfg) : cfg_(cfg)
    {
        kgmt::check(kgmt_model_dims(cfg.model, &S_, &C_, &D_));
        kgmt::check(kgmt_create(&cfg_, &ctx_));
        // a path has at most one row per iteration plus the root
        h_controlPathToGoal_ = new float[(size_t)D_ * (cfg_.max_iter + 1)]();
    }
    ~KGMT() override { kgmt_destroy(ctx_); }

    static kgmt_config preset(int32_t p)
    {
        kgmt_config c;
        kgmt::check(kgmt_config_preset(p, &c));
        return c;
    }

    // KGMT::plan, src/planners/KGMT.cu:41-105. d_obstacles_ptr: device memory,
    // 2 * W_DIM floats per box (xmin ymin zmin xmax ymax zmax; xmin ymin xmax
    // ymax for the planar unicycle), complete before the call (it is read on
    // the planner's own stream).
    void plan(float* h_initial, float* h_goal, float* d_obstacles_ptr, uint h_obstaclesCount) override
    {
        kgmt::check(kgmt_plan(ctx_, h_initial, h_goal, d_obstacles_ptr, h_obstaclesCount, 1, &result_));
        publish();
    }

    // result members from the last plan or plan step
    void publish()
    {
        h_treeSize_   = result_.tree_size;
        h_itr_        = result_.iterations;
        h_pathToGoal_ = result_.success ? result_.goal_index : 0;
        h_costToGoal_ = result_.success ? result_.path_cost : 0.0f;
        h_pathRows_   = result_.path_rows;
        if(h_pathRows_)
            {
                uint32_t n = 0;
                kgmt::check(kgmt_get_path(ctx_, h_controlPathToGoal_, cfg_.max_iter + 1, &n));
            }
        if(deviceTree_)
            {
                kgmt_device_tree t;
                kgmt::check(kgmt_get_tree_device(ctx_, &t));
                d_treeSamples_ptr_           = t.samples;
                d_treeSamplesParentIdxs_ptr_ = t.parents;
                d_treeSampleCosts_ptr_       = t.costs;
            }
    }

    // KGMT::planBench, src/planners/KGMT.cu:107-165: the plan's iterations one
    // at a time, each followed by writeDeviceVectorsToCSV(benchItr) into
    // Data/ under the current directory (KGMT.cu:405-488); prints each
    // iteration and the execution time as the reference does
    void planBench(float* h_initial, float* h_goal, float* d_obstacles_ptr, uint h_obstaclesCount, int benchItr)
    {
        const auto t0 = std::chrono::steady_clock::now();
        kgmt::check(kgmt_plan_begin(ctx_, h_initial, h_goal, d_obstacles_ptr, h_obstaclesCount, 1));
        h_treeSize_ = 1, h_itr_ = 0, h_pathToGoal_ = 0, h_costToGoal_ = 0.0f, h_pathRows_ = 0;
        h_frontierSize_ = 0;
        uint64_t expandedBefore = 0;
        int32_t done            = 0;
        uint frontierNow        = 1;  // the root
        while(!done)
            {
                std::printf("Iteration: %u, Tree Size: %u, Frontier Size: %u\n", h_itr_ + 1, h_treeSize_,
                            h_frontierSize_);
                h_frontierSize_ = frontierNow;
                kgmt::check(kgmt_plan_step(ctx_, &result_, &done));
                publish();
                h_expandedNodes_ = (uint)(result_.expanded - expandedBefore);
                expandedBefore   = result_.expanded;
                writeDeviceVectorsToCSV(benchItr);
                frontierNow = 0;
                for(uint8_t f : frontier_) frontierNow += f;
            }
        if(h_pathToGoal_) std::printf("Goal Reached\n");
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "KGMT execution time: " << s << " seconds. Iterations: " << h_itr_
                  << ". Tree Size: " << h_treeSize_ << std::endl;
    }

    // KGMT::writeDeviceVectorsToCSV, src/planners/KGMT.cu:405-488: the state
    // after the last iteration under Data/
    void writeDeviceVectorsToCSV(int itr)
    {
        namespace fs    = std::filesystem;
        const uint cap  = cfg_.capacity;
        const bool app  = h_itr_ != 0;
        const auto dir  = [&](const char* sub) {
            const std::string d = std::string("Data/") + sub + "/" + sub + std::to_string(itr);
            fs::create_directories(d);
            return d + "/";
        };
        std::vector<float> samples((size_t)cap * D_, 0.0f), costs(h_treeSize_);
        std::vector<int> parents(cap, -1);
        uint32_t n = 0;
        kgmt::check(kgmt_get_tree(ctx_, samples.data(), parents.data(), costs.data(), cap, &n));
        int32_t nr = 0;
        uint32_t nreg = 0;
        kgmt::check(kgmt_get_regions(ctx_, nullptr, nullptr, nullptr, 0, &nreg));
        std::vector<int> total(nreg), valid(nreg);
        std::vector<float> score(nreg);
        kgmt::check(kgmt_get_regions(ctx_, total.data(), valid.data(), score.data(), nreg, &nreg));
        nr = (int32_t)nreg;
        frontier_.assign(h_treeSize_, 0);
        std::vector<uint32_t> rep(h_treeSize_);
        kgmt::check(kgmt_get_frontier(ctx_, frontier_.data(), rep.data(), h_treeSize_, &n));
        std::vector<int> fr(cap, 0);
        std::vector<uint> repeat(cap, 0);
        for(uint i = 0; i < h_treeSize_; ++i) fr[i] = frontier_[i], repeat[i] = rep[i];
        kgmt::writeVectorToCSV(samples, dir("Samples") + "samples" + std::to_string(h_itr_) + ".csv", cap, D_, app);
        kgmt::writeVectorToCSV(parents, dir("Parents") + "parents" + std::to_string(h_itr_) + ".csv", cap, 1, app);
        kgmt::writeVectorToCSV(total, dir("TotalCountPerVertex") + "totalCountPerVertex.csv", 1, nr, app);
        kgmt::writeVectorToCSV(valid, dir("ValidCountPerVertex") + "validCountPerVertex.csv", 1, nr, app);
        kgmt::writeVectorToCSV(fr, dir("Frontier") + "frontier.csv", 1, cap, app);
        kgmt::writeVectorToCSV(repeat, dir("FrontierRepeatCount") + "frontierRepeatCount.csv", 1, cap, app);
        kgmt::writeVectorToCSV(score, dir("VertexScores") + "vertexScores.csv", 1, nr, app);
        kgmt::writeValueToCSV(h_frontierSize_, dir("FrontierSize") + "frontierSize.csv");
        kgmt::writeValueToCSV(h_treeSize_, dir("TreeSize") + "treeSize.csv");
        kgmt::writeValueToCSV(h_expandedNodes_, dir("ExpandedNodes") + "expandedNodes.csv");
        const std::string pathFile = dir("ControlPathToGoal") + "controlPathToGoal.csv";
        if(h_pathToGoal_)
            {
                std::vector<float> path(h_controlPathToGoal_, h_controlPathToGoal_ + (size_t)h_pathRows_ * D_);
                kgmt::writeVectorToCSV(path, pathFile, (int)h_pathRows_, D_, app);
            }
    }

    void setSeed(uint64_t seed) { kgmt::check(kgmt_set_seed(ctx_, seed)); }
    // write the device tree members after every plan (default) or not
    void setDeviceTree(bool on)
    {
        deviceTree_ = on;
        if(!on) d_treeSamples_ptr_ = nullptr, d_treeSamplesParentIdxs_ptr_ = nullptr, d_treeSampleCosts_ptr_ = nullptr;
    }
    // plan() returns with its results complete; this also waits for the few
    // launches it leaves to drain on the planner's stream
    void synchronize() { kgmt::check(kgmt_synchronize(ctx_)); }

    // host copies of the tree (row-major, D floats per node), parents, costs
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

    uint h_frontierSize_ = 0;   // frontier size at the start of the last iteration (planBench)
    uint h_expandedNodes_ = 0;  // children propagated by the last iteration (planBench)

    const kgmt_result& result() const { return result_; }
    const kgmt_config& config() const { return cfg_; }
    int sampleDim() const { return D_; }
    kgmt_ctx* context() const { return ctx_; }

private:
    kgmt_config cfg_;
    kgmt_ctx* ctx_ = nullptr;
    kgmt_result result_{};
    int32_t S_ = 0, C_ = 0, D_ = 0;
    bool deviceTree_ = true;
    std::vector<uint8_t> frontier_;
};
