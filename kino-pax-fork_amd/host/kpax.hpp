// kpax.hpp — C++ host mirror of the reference's planner classes over the
// libkpax C ABI (include/kpax.h).  Header-only; link with -lkpax.
//
// The reference's user code looks like (examples/gpu/main.cu:16-29):
//     KPAX kpax;
//     kpax.plan(h_initial, h_goal, d_obstacles, numObstacles);
//     // read kpax.h_pathToGoal_, kpax.h_controlPathToGoal_, kpax.h_treeSize_, kpax.h_itr_
// The same program against this header:
//     kpax::KPAX planner(kpax::Config::workload(2));
//     kpax::Result r = planner.plan(start, goal, obstacles);
// Errors raise kpax::Error (the reference calls exit(1) or does not check).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/kpax.h"

namespace kpax {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(int rc, const char* what, const kpax_handle* h = nullptr)
{
    if(rc != KPAX_OK) throw Error(rc, std::string(what) + " failed (" + std::to_string(rc) + ")" +
                                          (h ? std::string(": ") + kpax_last_error(h) : std::string()));
}

struct Config : kpax_config {
    // replaces editing include/config/config.h and recompiling
    static Config workload(int k)
    {
        Config c;
        check(kpax_config_workload(k, &c), "kpax_config_workload");
        return c;
    }
    int state_dim() const { return model == KPAX_MODEL_QUADCOPTER ? 12 : 6; }
    int control_dim() const { return model == KPAX_MODEL_QUADCOPTER ? 4 : 3; }
};

// What the reference leaves in Planner's public fields (Planner.cuh:31-34).
struct Result {
    int status = KPAX_NOT_FOUND;
    int path_to_goal = 0;                      // h_pathToGoal_: goal node index, 0 if none
    std::vector<float> control_path_to_goal;   // h_controlPathToGoal_: rows goal -> root
    std::vector<float> rows;                   // rows root -> goal, [state, control, duration]
    std::vector<int32_t> steps;
    uint32_t tree_size = 0;                    // h_treeSize_
    uint32_t itr = 0;                          // h_itr_
    uint64_t expansions = 0;
    double plan_ms = 0.0;
    int verified = -1;
    bool found() const { return status == KPAX_OK; }
};

// Obstacles in the reference's row format [minx, miny, minz, maxx, maxy, maxz].
inline std::vector<float> readObstaclesFromCSV(const std::string& filename, int& numObstacles)
{
    uint32_t n = 0;
    check(kpax_read_obstacles_csv(filename.c_str(), nullptr, 0, &n), "kpax_read_obstacles_csv");
    std::vector<float> v(6 * (size_t)n);
    check(kpax_read_obstacles_csv(filename.c_str(), v.data(), n, &n), "kpax_read_obstacles_csv");
    numObstacles = (int)n;
    return v;
}

class KPAX {
public:
    explicit KPAX(const Config& cfg = Config::workload(2)) : cfg_(cfg)
    {
        check(kpax_create(&cfg_, &h_), "kpax_create");
    }
    ~KPAX()
    {
        if(h_) kpax_destroy(h_);
    }
    KPAX(const KPAX&)            = delete;
    KPAX& operator=(const KPAX&) = delete;

    void setObstacles(const float* obstacles, uint32_t n)
    {
        check(kpax_set_obstacles(h_, obstacles, n), "kpax_set_obstacles", h_);
    }

    // KPAX::plan(h_initial, h_goal, d_obstacles, h_obstaclesCount) (KPAX.cuh:14).
    // Obstacles are host memory here; the grid is rebuilt only when they change.
    Result plan(const float* h_initial, const float* h_goal, const float* obstacles, uint32_t n_obstacles,
                uint64_t seed = 0)
    {
        if(obstacles != last_obs_ || n_obstacles != last_n_)
        {
            setObstacles(obstacles, n_obstacles);
            last_obs_ = obstacles;
            last_n_   = n_obstacles;
        }
        kpax_result r{};
        check(kpax_plan(h_, h_initial, h_goal, seed, &r), "kpax_plan", h_);
        Result out;
        const int D      = cfg_.state_dim() + cfg_.control_dim() + 1;
        out.status       = r.status;
        out.path_to_goal = r.goal_index > 0 ? r.goal_index : 0;
        out.rows.assign(r.rows, r.rows + (size_t)r.n_rows * D);
        out.steps.assign(r.steps, r.steps + r.n_rows);
        for(uint32_t i = r.n_rows; i-- > 0;)
            out.control_path_to_goal.insert(out.control_path_to_goal.end(), r.rows + (size_t)i * D,
                                            r.rows + (size_t)(i + 1) * D);
        out.tree_size  = r.tree_size;
        out.itr        = r.iterations;
        out.expansions = r.expansions;
        out.plan_ms    = r.plan_ms;
        out.verified   = r.verified;
        kpax_result_free(&r);
        return out;
    }

    Result plan(const std::vector<float>& initial, const std::vector<float>& goal,
                const std::vector<float>& obstacles, uint64_t seed = 0)
    {
        return plan(initial.data(), goal.data(), obstacles.data(), (uint32_t)(obstacles.size() / 6), seed);
    }

    kpax_handle* handle() const { return h_; }
    const Config& config() const { return cfg_; }

private:
    Config cfg_;
    kpax_handle* h_        = nullptr;
    const float* last_obs_ = nullptr;
    uint32_t last_n_       = ~0u;
};

}  // namespace kpax
