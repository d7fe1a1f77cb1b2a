// plan_main.cpp — command-line planner over the C++ mirror, the counterpart
// of the reference's examples/gpu/executionTimeMain.cu: load an obstacle
// CSV (or generate the workload's environment), plan N times, print one line
// per run and the p50 time.
//
//   plan_main <workload 1|2|3> [runs] [obstacles.csv]
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "kpax.hpp"

int main(int argc, char** argv)
{
    const int k    = argc > 1 ? std::atoi(argv[1]) : 2;
    const int runs = argc > 2 ? std::atoi(argv[2]) : 5;
    try
    {
        kpax::Config cfg = kpax::Config::workload(k);
        std::vector<float> start(12, 0.0f), goal(4, 0.0f);
        const float s[3][3] = {{0.05f, 0.05f, 0.05f}, {0.5f, 0.5f, 1.0f}, {0.1f, 0.5f, 0.5f}};
        const float g[3][3] = {{0.95f, 0.95f, 0.95f}, {9.5f, 9.5f, 1.0f}, {0.9f, 0.5f, 0.5f}};
        for(int d = 0; d < 3; ++d)
        {
            start[d] = s[k - 1][d];
            goal[d]  = g[k - 1][d];
        }
        if(k == 3) start[5] = 0.15f;
        goal[3] = cfg.goal_radius;
        std::vector<float> obs;
        if(argc > 3)
        {
            int n = 0;
            obs   = kpax::readObstaclesFromCSV(argv[3], n);
        }
        else
        {
            const uint32_t want[3] = {20, 400, 2000};
            uint32_t n             = 0;
            kpax::check(kpax_generate_environment(k, 0, want[k - 1], cfg.ws_min, cfg.ws_max, start.data(),
                                                  goal.data(), nullptr, 0, &n),
                        "kpax_generate_environment");
            obs.resize(6 * (size_t)n);
            kpax::check(kpax_generate_environment(k, 0, want[k - 1], cfg.ws_min, cfg.ws_max, start.data(),
                                                  goal.data(), obs.data(), n, &n),
                        "kpax_generate_environment");
        }
        kpax::KPAX planner(cfg);
        std::vector<double> ms;
        for(int r = 0; r < runs; ++r)
        {
            const kpax::Result res = planner.plan(start, goal, obs, 1000 + r);
            std::printf("run %d: status %d  %.3f ms  iterations %u  tree %u  path rows %zu  verified %d\n", r,
                        res.status, res.plan_ms, res.itr, res.tree_size, res.steps.size(), res.verified);
            ms.push_back(res.plan_ms);
        }
        std::sort(ms.begin(), ms.end());
        std::printf("p50 %.3f ms over %d runs\n", ms[ms.size() / 2], runs);
    }
    catch(const kpax::Error& e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
