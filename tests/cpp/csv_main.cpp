// readObstaclesFromCSV of include/kgmt_planner.hpp on one file: prints the
// obstacle count and every value read, one per line (%.9g).
// usage: csv_main <file> <workspace_dim>
#include <cstdio>
#include <cstdlib>

#include "../../include/kgmt_planner.hpp"

int main(int argc, char** argv)
{
    if(argc < 3) return 2;
    int n = -1;
    try
        {
            const std::vector<float> v = readObstaclesFromCSV(argv[1], n, std::atoi(argv[2]));
            std::printf("%d %zu\n", n, v.size());
            for(float x : v) std::printf("%.9g\n", x);
        }
    catch(const kgmt::Error& e)
        {
            std::printf("error %d\n", e.code);
            return 1;
        }
    return 0;
}
