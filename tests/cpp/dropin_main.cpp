// Drop-in check of include/kgmt_planner.hpp, written the way the reference's
// examples/gpu/executionTimeMain.cu drives its planner: readObstaclesFromCSV
// with W_DIM, start and goal arrays of SAMPLE_DIM floats, the obstacles copied
// to device memory, plan() with the device pointer, then the public result
// members, including h_controlPathToGoal_ (float*) and the device tree
// pointers d_treeSamples_ptr_ / d_treeSamplesParentIdxs_ptr_ /
// d_treeSampleCosts_ptr_.
//
// usage: dropin_main <obstacles.csv> <capacity> <max_iter> <seed> <runs>
//        prints one JSON object per plan;
//        dropin_main <obstacles.csv> <capacity> <max_iter> <seed> bench <dir>
//        runs planBench(..., 0) with <dir> as the working directory (its
//        Data/ dump lands there).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kgmt_planner.hpp"

static uint64_t bit_sum(const float* v, size_t n)
{
    uint64_t s = 0;
    for(size_t i = 0; i < n; ++i)
        {
            uint32_t b;
            std::memcpy(&b, &v[i], 4);
            s += b;
        }
    return s;
}

int main(int argc, char** argv)
{
    if(argc < 6)
        {
            std::fprintf(stderr, "usage: %s obstacles.csv capacity max_iter seed runs\n", argv[0]);
            return 2;
        }
    int numObstacles = 0;
    float* d_obstacles;
    // --- Load Workspace Obstacles ---
    std::vector<float> obstacles = readObstaclesFromCSV(argv[1], numObstacles, W_DIM);
    // --- Transfer Obstacles to device ---
    if(hipMalloc(&d_obstacles, std::max(numObstacles, 1) * 2 * W_DIM * sizeof(float)) != hipSuccess) return 3;
    if(numObstacles && hipMemcpy(d_obstacles, obstacles.data(), numObstacles * 2 * W_DIM * sizeof(float),
                                 hipMemcpyHostToDevice) != hipSuccess)
        return 3;

    // executionTimeMain.cu: start and goal in the 100-unit workspace
    float h_initial[SAMPLE_DIM] = {10.0, 8, 5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
          h_goal[SAMPLE_DIM]    = {80, 95.0, 90.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    try
        {
            // the default planner is the reference's configuration (config.h's
            // quadcopter block); the parity runs take smaller sizes
            KGMT reference_default;
            if(reference_default.sampleDim() != SAMPLE_DIM || reference_default.config().capacity != 400000)
                {
                    std::fprintf(stderr, "default configuration is not config.h's\n");
                    return 4;
                }
            kgmt_config cfg = KGMT::preset(KGMT_PRESET_REFERENCE_QUAD);
            cfg.capacity    = (uint32_t)std::atoi(argv[2]);
            cfg.max_iter    = (uint32_t)std::atoi(argv[3]);
            cfg.seed        = (uint64_t)std::atoll(argv[4]);
            KGMT kgmt(cfg);
            if(std::string(argv[5]) == "bench" && argc > 6)
                {
                    std::filesystem::current_path(argv[6]);
                    kgmt.planBench(h_initial, h_goal, d_obstacles, numObstacles, 0);
                    (void)hipFree(d_obstacles);
                    return 0;
                }
            const int runs = std::atoi(argv[5]);
            for(int i = 0; i < runs; ++i)
                {
                    // --- Execute planner ---
                    kgmt.plan(h_initial, h_goal, d_obstacles, numObstacles);
                    const uint n   = kgmt.h_treeSize_;
                    const uint cap = cfg.capacity;
                    std::vector<float> rows((size_t)cap * SAMPLE_DIM), costs(cap);
                    std::vector<int> parents(cap);
                    if(hipMemcpy(rows.data(), kgmt.d_treeSamples_ptr_, rows.size() * sizeof(float),
                                 hipMemcpyDeviceToHost) != hipSuccess ||
                       hipMemcpy(parents.data(), kgmt.d_treeSamplesParentIdxs_ptr_, cap * sizeof(int),
                                 hipMemcpyDeviceToHost) != hipSuccess ||
                       hipMemcpy(costs.data(), kgmt.d_treeSampleCosts_ptr_, cap * sizeof(float),
                                 hipMemcpyDeviceToHost) != hipSuccess)
                        return 5;
                    // entries past the tree as the reference's fills leave them
                    bool tail = true;
                    for(size_t k = (size_t)n * SAMPLE_DIM; k < rows.size(); ++k) tail = tail && rows[k] == 0.0f;
                    for(uint k = n; k < cap; ++k) tail = tail && parents[k] == -1 && costs[k] == 0.0f;
                    rows.resize((size_t)n * SAMPLE_DIM), costs.resize(n), parents.resize(n);
                    // the device members agree with the host getters
                    const bool same = rows == kgmt.treeSamples() && parents == kgmt.treeParents() &&
                                      costs == kgmt.treeCosts();
                    long long psum = 0;
                    for(int p : parents) psum += p;
                    std::printf("{\"run\": %d, \"tree_size\": %u, \"iterations\": %u, \"goal\": %d, \"path_rows\": %u, "
                                "\"parent_sum\": %lld, \"sample_bits\": %llu, \"cost_bits\": %llu, "
                                "\"path_bits\": %llu, \"device_tree_matches_host\": %s, \"device_tree_tail_ok\": %s, "
                                "\"time_ms\": %.3f}\n",
                                i, kgmt.h_treeSize_, kgmt.h_itr_, kgmt.h_pathToGoal_, kgmt.h_pathRows_, psum,
                                (unsigned long long)bit_sum(rows.data(), rows.size()),
                                (unsigned long long)bit_sum(costs.data(), costs.size()),
                                (unsigned long long)bit_sum(kgmt.h_controlPathToGoal_,
                                                            (size_t)kgmt.h_pathRows_ * SAMPLE_DIM),
                                same ? "true" : "false", tail ? "true" : "false", kgmt.result().time_ms);
                }
        }
    catch(const kgmt::Error& e)
        {
            std::fprintf(stderr, "%s\n", e.what());
            return 1;
        }
    // --- Free memory ---
    (void)hipFree(d_obstacles);
    return 0;
}
