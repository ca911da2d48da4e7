// Drop-in check of include/kgmt_planner.hpp, written the way the reference's
// examples/gpu/executionTimeMain.cu drives its planner: read an obstacle CSV,
// copy the boxes to device memory, construct KGMT, call plan() with the device
// pointer, read the public result members.
//
// usage: dropin_main <obstacles.csv> <capacity> <max_iter> <seed> <runs>
// prints one JSON object per plan.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/kgmt_planner.hpp"

static std::vector<float> read_csv(const char* path)
{
    std::ifstream f(path);
    std::vector<float> v;
    std::string line;
    while(std::getline(f, line))
        {
            for(char& ch : line)
                if(ch == ',') ch = ' ';
            std::stringstream ss(line);
            float x;
            while(ss >> x) v.push_back(x);
        }
    v.resize(v.size() / 6 * 6);
    return v;
}

int main(int argc, char** argv)
{
    if(argc < 6)
        {
            std::fprintf(stderr, "usage: %s obstacles.csv capacity max_iter seed runs\n", argv[0]);
            return 2;
        }
    std::vector<float> obs = read_csv(argv[1]);
    const uint n_obs       = (uint)(obs.size() / 6);
    kgmt_config cfg        = KGMT::preset(KGMT_PRESET_REFERENCE_QUAD);
    cfg.capacity           = (uint32_t)std::atoi(argv[2]);
    cfg.max_iter           = (uint32_t)std::atoi(argv[3]);
    cfg.seed               = (uint64_t)std::atoll(argv[4]);
    const int runs         = std::atoi(argv[5]);

    float* d_obstacles = nullptr;
    if(hipMalloc(&d_obstacles, std::max<size_t>(obs.size(), 1) * sizeof(float)) != hipSuccess) return 3;
    if(n_obs && hipMemcpy(d_obstacles, obs.data(), obs.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return 3;

    // executionTimeMain.cu:43-44: start and goal in the 100-unit workspace
    float h_initial[KGMT_MAX_SAMPLE] = {10.0f, 8.0f, 5.0f};
    float h_goal[KGMT_MAX_SAMPLE]    = {80.0f, 95.0f, 90.0f};
    try
        {
            KGMT kgmt(cfg);
            for(int i = 0; i < runs; ++i)
                {
                    kgmt.plan(h_initial, h_goal, d_obstacles, n_obs);
                    const auto parents = kgmt.treeParents();
                    long long psum     = 0;
                    for(int32_t p : parents) psum += p;
                    std::printf("{\"run\": %d, \"tree_size\": %u, \"iterations\": %u, \"goal\": %d, \"path_rows\": %u, "
                                "\"parent_sum\": %lld, \"time_ms\": %.3f}\n",
                                i, kgmt.h_treeSize_, kgmt.h_itr_, kgmt.h_pathToGoal_, kgmt.h_pathRows_, psum,
                                kgmt.result().time_ms);
                }
        }
    catch(const kgmt::Error& e)
        {
            std::fprintf(stderr, "%s\n", e.what());
            return 1;
        }
    (void)hipFree(d_obstacles);
    return 0;
}
