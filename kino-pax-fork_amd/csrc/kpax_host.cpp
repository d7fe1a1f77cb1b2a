// kpax_host.cpp — host-only half of libkpax: built-in workloads, config
// validation and constant folding, the obstacle CSV reader, seeded
// environment generators and the FP64 trajectory re-verifier.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "kpax_internal.h"

namespace kpax {

int state_dim(int model) { return model == KPAX_MODEL_QUADCOPTER ? 12 : 6; }
int control_dim(int model) { return model == KPAX_MODEL_QUADCOPTER ? 4 : 3; }

std::string validate_config(const kpax_config& c)
{
    if(c.model < 1 || c.model > 3) return "model must be 1, 2 or 3";
    if(!(c.dt > 0.0f) || !std::isfinite(c.dt)) return "dt must be positive";
    if(c.steps_min < 1 || c.steps_max < c.steps_min || c.steps_max > 100000) return "bad step range";
    for(int d = 0; d < 3; ++d)
    {
        if(!std::isfinite(c.ws_min[d]) || !std::isfinite(c.ws_max[d]) || !(c.ws_min[d] < c.ws_max[d]))
            return "bad workspace box";
        if(c.r1_len[d] < 1 || c.r2_len[d] < 1 || c.grid_dims[d] < 1) return "region and grid lengths must be >= 1";
    }
    for(int i = 0; i < control_dim(c.model); ++i)
        if(!std::isfinite(c.ctrl_min[i]) || !std::isfinite(c.ctrl_max[i]) || c.ctrl_min[i] > c.ctrl_max[i])
            return "bad control box";
    if(!(c.vel_min <= c.vel_max)) return "bad velocity box";
    if(c.model == KPAX_MODEL_DUBINS_AIRPLANE && !(c.pitch_min <= c.pitch_max)) return "bad pitch box";
    if(c.model == KPAX_MODEL_QUADCOPTER &&
       (!(c.mass > 0.0f) || !(c.inertia[0] > 0.0f) || !(c.inertia[1] > 0.0f) || !(c.inertia[2] > 0.0f)))
        return "quadcopter mass and inertia must be positive";
    const uint64_t g = (uint64_t)c.grid_dims[0] * c.grid_dims[1] * c.grid_dims[2];
    if(g > (1ull << 27)) return "grid has too many cells";
    const uint64_t r1 = (uint64_t)c.r1_len[0] * c.r1_len[1] * c.r1_len[2];
    const uint64_t r2 = r1 * c.r2_len[0] * c.r2_len[1] * c.r2_len[2];
    if(r1 > (1u << 20) || r2 > (1u << 28)) return "too many regions";
    if(c.max_tree < 2 || c.max_tree > (1u << 27)) return "max_tree must be in [2, 2^27]";
    if(c.max_iter < 1 || c.max_iter > 100000) return "max_iter must be in [1, 100000]";
    if(!(c.goal_radius > 0.0f)) return "goal radius must be positive";
    if(c.fanout < 1 || c.fanout > 1024) return "fanout must be in [1, 1024]";
    if(c.root_repeat < 1 || c.root_repeat > c.max_tree) return "bad root repeat";
    if(c.sparse_repeat > 1024 || c.dense_repeat > 1024) return "repeat counts must be <= 1024";
    if(!(c.eps > 0.0f)) return "eps must be positive";
    if(c.device < 0) return "bad device";
    if(c.prune && !(c.prune_progress_scale > 0.0f && std::isfinite(c.prune_exploration) &&
                    std::isfinite(c.prune_goal_bias)))
        return "pruning needs a positive progress scale and finite biases";
    return "";
}

ModelParams make_model_params(const kpax_config& c)
{
    ModelParams p{};
    p.model       = c.model;
    p.di_textbook = c.di_textbook;
    p.dt          = c.dt;
    p.half        = 0.5f * c.dt;
    p.d6          = c.dt / 6.0f;
    p.steps_min   = c.steps_min;
    p.span        = c.steps_max - c.steps_min + 1;
    for(int d = 0; d < 3; ++d)
    {
        p.ws_min[d] = c.ws_min[d];
        p.ws_max[d] = c.ws_max[d];
    }
    for(int i = 0; i < 4; ++i)
    {
        p.ctrl_min[i] = c.ctrl_min[i];
        p.ctrl_rng[i] = c.ctrl_max[i] - c.ctrl_min[i];
    }
    p.vmin = c.vel_min;
    p.vmax = c.vel_max;
    p.pmin = c.pitch_min;
    p.pmax = c.pitch_max;
    // folded as the reference's macros fold them (config.h:300-308)
    const float ix = c.inertia[0], iy = c.inertia[1], iz = c.inertia[2];
    p.negnu = -c.nu;
    p.negmu = -c.mu;
    p.g     = c.gravity;
    p.mi    = c.model == KPAX_MODEL_QUADCOPTER ? 1.0f / c.mass : 0.0f;
    if(c.model == KPAX_MODEL_QUADCOPTER)
    {
        p.k1x = (iy - iz) / ix;
        p.k1y = (iz - ix) / iy;
        p.k1z = (ix - iy) / iz;
        p.ixi = 1.0f / ix;
        p.iyi = 1.0f / iy;
        p.izi = 1.0f / iz;
    }
    return p;
}

RegionParams make_region_params(const kpax_config& c)
{
    RegionParams r{};
    for(int d = 0; d < 3; ++d)
    {
        r.L[d]    = c.r1_len[d];
        r.L2[d]   = c.r2_len[d];
        r.mn[d]   = c.ws_min[d];
        r.ext[d]  = c.ws_max[d] - c.ws_min[d];
        r.size[d] = r.ext[d] / (float)r.L[d];
    }
    r.r1_count = r.L[0] * r.L[1] * r.L[2];
    r.r2_per   = r.L2[0] * r.L2[1] * r.L2[2];
    r.vol      = r.size[0] * r.size[1] * r.size[2];
    return r;
}

GridParams make_grid_params(const kpax_config& c)
{
    GridParams g{};
    for(int d = 0; d < 3; ++d)
    {
        g.D[d]    = c.grid_dims[d];
        g.mn[d]   = c.ws_min[d];
        g.cell[d] = (c.ws_max[d] - c.ws_min[d]) / (float)c.grid_dims[d];
        g.inv[d]  = (float)c.grid_dims[d] / (c.ws_max[d] - c.ws_min[d]);
    }
    return g;
}

PlanParams make_plan_params(const kpax_config& c)
{
    PlanParams p{};
    p.max_iter           = c.max_iter;
    p.fanout             = c.fanout;
    p.root_repeat        = c.root_repeat;
    p.sparse_repeat      = c.sparse_repeat;
    p.dense_repeat       = c.dense_repeat;
    p.sparse_valid_below = c.sparse_valid_below;
    p.eps                = c.eps;
    p.goal_radius        = c.goal_radius;
    p.prune              = c.prune;
    p.prune_expl         = c.prune_exploration;
    p.prune_goal         = c.prune_goal_bias;
    p.prune_scale        = c.prune_progress_scale;
    return p;
}

// ---------------------------------------------------------------------------
// FP64 re-verification.  Each edge restarts from its parent row's stored
// state, integrates in double with the same scheme and checks, and collides
// against every obstacle (brute force, collisionCheck.cu:49-62).
// ---------------------------------------------------------------------------
namespace {

bool inb(double v, double lo, double hi) { return v >= lo && v <= hi; }

bool free_fp64(const kpax_config& c, const float* obs, uint32_t n, const double p0[3], const double p1[3])
{
    for(int d = 0; d < 3; ++d)
        if(!inb(p1[d], c.ws_min[d], c.ws_max[d])) return false;
    double lo[3], hi[3];
    for(int d = 0; d < 3; ++d)
    {
        lo[d] = std::min(p0[d], p1[d]);
        hi[d] = std::max(p0[d], p1[d]);
    }
    for(uint32_t i = 0; i < n; ++i)
    {
        const float* o = obs + 6 * (size_t)i;
        bool sep       = false;
        for(int d = 0; d < 3 && !sep; ++d) sep = hi[d] <= o[d] || o[3 + d] <= lo[d];
        if(!sep) return false;
    }
    return true;
}

void quad_deriv(const kpax_config& c, const double* x, const double* u, double* f)
{
    const double phi = x[3], th = x[4], psi = x[5], uu = x[6], v = x[7], w = x[8], p = x[9], q = x[10], r = x[11];
    const double sf = std::sin(phi), cf = std::cos(phi), st = std::sin(th), ct = std::cos(th);
    const double ss = std::sin(psi), cs = std::cos(psi);
    const double ix = c.inertia[0], iy = c.inertia[1], iz = c.inertia[2], g = c.gravity, m = c.mass;
    f[0]            = ct * cs * uu + (sf * st * cs - cf * ss) * v + (cf * st * cs + sf * ss) * w;
    f[1]            = ct * ss * uu + (sf * st * ss + cf * cs) * v + (cf * st * ss - sf * cs) * w;
    f[2]            = -st * uu + sf * ct * v + cf * ct * w;
    f[3]            = p + (q * sf + r * cf) * std::tan(th);
    f[4]            = q * cf - r * sf;
    f[5]            = (q * sf + r * cf) / ct;
    const double xyz = -(double)c.nu * std::sqrt(uu * uu + v * v + w * w);
    f[6]             = (r * v - q * w) - g * st + xyz * uu / m;
    f[7]             = (p * w - r * uu) + g * ct * sf + xyz * v / m;
    f[8]             = (q * uu - p * v) + g * ct * cf + xyz * w / m + u[0] / m;
    const double lmn = -(double)c.mu * std::sqrt(p * p + q * q + r * r);
    f[9]             = (iy - iz) / ix * q * r + lmn * p / ix + u[1] / ix;
    f[10]            = (iz - ix) / iy * p * r + lmn * q / iy + u[2] / iy;
    f[11]            = (ix - iy) / iz * p * q + lmn * r / iz + u[3] / iz;
}

// Returns true if the edge is valid; x holds the state after the last step.
bool edge_fp64(const kpax_config& c, const float* obs, uint32_t n, double* x, const double* u, int steps)
{
    const double dt = c.dt;
    for(int i = 0; i < steps; ++i)
    {
        const double p0[3] = {x[0], x[1], x[2]};
        if(c.model == KPAX_MODEL_DOUBLE_INTEGRATOR)
        {
            for(int d = 0; d < 3; ++d)
            {
                if(c.di_textbook)
                    x[d] += x[3 + d] * dt + 0.5 * u[d] * dt * dt;
                else
                    x[d] += (5.0 / 6.0) * x[3 + d] * dt + u[d] * dt * dt / 3.0;
                x[3 + d] += u[d] * dt;
            }
            for(int d = 0; d < 3; ++d)
                if(!inb(x[3 + d], c.vel_min, c.vel_max)) return false;
        }
        else if(c.model == KPAX_MODEL_DUBINS_AIRPLANE)
        {
            // yaw, pitch and v are linear in time; x, y, z by Simpson's rule,
            // which is what the reference's four-stage update reduces to.
            const double a = u[0], yr = u[1], pr = u[2];
            const double yaw = x[3], pitch = x[4], v = x[5];
            auto vel = [&](double s, double* o) {
                const double vv = v + a * s, pp = pitch + pr * s, yy = yaw + yr * s;
                o[0]            = vv * std::cos(pp) * std::cos(yy);
                o[1]            = vv * std::cos(pp) * std::sin(yy);
                o[2]            = vv * std::sin(pp);
            };
            double k1[3], k2[3], k4[3];
            vel(0.0, k1);
            vel(0.5 * dt, k2);
            vel(dt, k4);
            for(int d = 0; d < 3; ++d) x[d] += dt / 6.0 * (k1[d] + 4.0 * k2[d] + k4[d]);
            x[3] = yaw + dt * yr;
            x[4] = pitch + dt * pr;
            x[5] = v + dt * a;
            if(!inb(x[5], c.vel_min, c.vel_max) || !inb(x[4], c.pitch_min, c.pitch_max)) return false;
        }
        else
        {
            double k1[12], k2[12], k3[12], k4[12], tmp[12];
            quad_deriv(c, x, u, k1);
            for(int j = 0; j < 12; ++j) tmp[j] = x[j] + 0.5 * dt * k1[j];
            quad_deriv(c, tmp, u, k2);
            for(int j = 0; j < 12; ++j) tmp[j] = x[j] + 0.5 * dt * k2[j];
            quad_deriv(c, tmp, u, k3);
            for(int j = 0; j < 12; ++j) tmp[j] = x[j] + dt * k3[j];
            quad_deriv(c, tmp, u, k4);
            for(int j = 0; j < 12; ++j) x[j] += dt / 6.0 * (k1[j] + 2.0 * k2[j] + 2.0 * k3[j] + k4[j]);
            for(int d = 6; d < 9; ++d)
                if(!inb(x[d], c.vel_min, c.vel_max)) return false;
        }
        const double p1[3] = {x[0], x[1], x[2]};
        if(!free_fp64(c, obs, n, p0, p1)) return false;
    }
    return true;
}

}  // namespace

bool verify_path_fp64(const kpax_config& c, const float* obs, uint32_t n_obs, const float* rows, const int32_t* steps,
                      uint32_t n_rows, const float* goal, float* max_dev)
{
    const int S = state_dim(c.model), C = control_dim(c.model), D = S + C + 1;
    double dev = 0.0;
    bool ok    = n_rows >= 1;
    double last[12];
    for(int d = 0; d < S && n_rows; ++d) last[d] = rows[d];
    for(uint32_t r = 1; ok && r < n_rows; ++r)
    {
        double x[12], u[4];
        for(int d = 0; d < S; ++d) x[d] = rows[(size_t)(r - 1) * D + d];
        for(int d = 0; d < C; ++d) u[row_control(c.model, d)] = rows[(size_t)r * D + S + d];
        if(steps[r] < 1 || !edge_fp64(c, obs, n_obs, x, u, steps[r])) ok = false;
        for(int d = 0; d < S; ++d)
        {
            const double e = std::fabs(x[d] - (double)rows[(size_t)r * D + d]);
            dev            = std::max(dev, e);
            if(!(e <= 1e-4 * (1.0 + std::fabs(x[d])))) ok = false;
        }
        for(int d = 0; d < S; ++d) last[d] = x[d];
    }
    if(ok)
    {
        const double dx = last[0] - goal[0], dy = last[1] - goal[1], dz = last[2] - goal[2];
        ok              = std::sqrt(dx * dx + dy * dy + dz * dz) < (double)c.goal_radius;
    }
    if(max_dev) *max_dev = (float)dev;
    return ok;
}

}  // namespace kpax

// ===========================================================================
// C ABI (host-only entry points)
// ===========================================================================
using namespace kpax;

namespace {

// splitmix64 (Steele, Lea, Flood, OOPSLA'14) for the environment generators.
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * 0x1p-53; }
    double uni(double a, double b) { return a + (b - a) * uni(); }
};

double box_point_dist(const float* b, const float* p)
{
    double s = 0.0;
    for(int d = 0; d < 3; ++d)
    {
        const double lo = b[d], hi = b[3 + d], v = p[d];
        const double e  = v < lo ? lo - v : (v > hi ? v - hi : 0.0);
        s += e * e;
    }
    return std::sqrt(s);
}

// Keeps boxes off the start point and the goal ball with a margin.
bool clear_of(const float* b, const float* start, const float* goal, double margin)
{
    if(start && box_point_dist(b, start) <= margin) return false;
    if(goal && box_point_dist(b, goal) <= (double)goal[3] + margin) return false;
    return true;
}

}  // namespace

extern "C" {

int kpax_model_dims(int32_t model, int32_t* s, int32_t* c)
{
    if(model < 1 || model > 3) return KPAX_INVALID_ARG;
    if(s) *s = state_dim(model);
    if(c) *c = control_dim(model);
    return KPAX_OK;
}

int kpax_config_workload(int32_t workload, kpax_config* out)
{
    if(!out) return KPAX_INVALID_ARG;
    kpax_config c;
    memset(&c, 0, sizeof(c));
    c.fanout             = 32;   // h_activeBlockSize_, KPAX.cu:31
    c.root_repeat        = 5;    // KPAX.cu:77
    c.sparse_repeat      = 15;   // KPAX.cu:477
    c.dense_repeat       = 1;    // KPAX.cu:479
    c.sparse_valid_below = 10;   // KPAX.cu:476
    c.eps                = 1e-2f;  // EPSILON, config.h:263
    c.max_iter           = 300;  // MAX_ITER, config.h:212
    c.verify             = 1;
    c.prune                = 0;      // PruneKPAX defaults, PruneKPAX.cu:34-36
    c.prune_exploration    = 0.3f;
    c.prune_goal_bias      = 0.7f;
    c.prune_progress_scale = 10.0f;
    c.steps_min          = 5;
    c.steps_max          = 20;
    c.gravity            = -9.81f;  // GRAVITY, config.h:306
    c.nu                 = 10e-3f;  // NU, config.h:300
    c.mu                 = 2e-6f;   // MU, config.h:301
    c.mass               = 1.0f;
    c.inertia[0] = c.inertia[1] = c.inertia[2] = 1.0f;
    switch(workload)
    {
        case 1:  // 6D double integrator, [0,1]^3, 20 random boxes
            c.model = KPAX_MODEL_DOUBLE_INTEGRATOR;
            c.dt    = 0.01f;
            for(int d = 0; d < 3; ++d)
            {
                c.ws_min[d]    = 0.0f;
                c.ws_max[d]    = 1.0f;
                c.ctrl_min[d]  = -1.0f;
                c.ctrl_max[d]  = 1.0f;
                c.r1_len[d]    = 8;
                c.r2_len[d]    = 4;
                c.grid_dims[d] = 16;
            }
            c.vel_min     = -1.0f;
            c.vel_max     = 1.0f;
            c.max_tree    = 65536;
            c.goal_radius = 0.05f;
            break;
        case 2:  // 12D quadcopter, forest in [0,10]^2 x [0,5]
            c.model = KPAX_MODEL_QUADCOPTER;
            c.dt    = 0.02f;
            for(int d = 0; d < 3; ++d)
            {
                c.ws_min[d]    = 0.0f;
                c.ws_max[d]    = d == 2 ? 5.0f : 10.0f;
                c.r1_len[d]    = 16;
                c.r2_len[d]    = 4;
                c.grid_dims[d] = 64;
            }
            c.ctrl_min[0] = 0.0f;  // thrust U[0, 2 m g]
            c.ctrl_max[0] = 19.62f;
            for(int i = 1; i < 4; ++i)
            {
                c.ctrl_min[i] = -0.1f;  // body torques
                c.ctrl_max[i] = 0.1f;
            }
            c.vel_min    = -3.0f;
            c.vel_max    = 3.0f;
            c.mass       = 1.0f;
            c.inertia[0] = 0.01f;
            c.inertia[1] = 0.01f;
            c.inertia[2] = 0.02f;
            c.max_tree   = 1u << 20;
            c.goal_radius = 0.5f;
            break;
        case 3:  // 6D Dubins airplane, maze in [0,1]^3
            c.model = KPAX_MODEL_DUBINS_AIRPLANE;
            c.dt    = 0.01f;
            for(int d = 0; d < 3; ++d)
            {
                c.ws_min[d]    = 0.0f;
                c.ws_max[d]    = 1.0f;
                c.r1_len[d]    = 16;
                c.r2_len[d]    = 4;
                c.grid_dims[d] = 128;
            }
            // (a, yawRate, pitchRate) within the reference bounds, config.h:135-136,274-277
            c.ctrl_min[0] = -0.3f;
            c.ctrl_max[0] = 0.3f;
            c.ctrl_min[1] = -0.78539816f;
            c.ctrl_max[1] = 0.78539816f;
            c.ctrl_min[2] = -0.78539816f;
            c.ctrl_max[2] = 0.78539816f;
            c.vel_min     = 0.0f;
            c.vel_max     = 0.3f;
            c.pitch_min   = -1.04719755f;
            c.pitch_max   = 1.04719755f;
            c.max_tree    = 1u << 21;
            c.goal_radius = 0.05f;
            break;
        default: return KPAX_INVALID_ARG;
    }
    *out = c;
    return KPAX_OK;
}

int kpax_read_obstacles_csv(const char* path, float* out, uint32_t cap, uint32_t* n)
{
    if(!path || !n) return KPAX_INVALID_ARG;
    std::ifstream f(path);
    if(!f.is_open()) return KPAX_IO_ERROR;
    std::vector<float> v;
    std::string line;
    while(std::getline(f, line))
    {
        for(char& ch : line)
            if(ch == ',') ch = ' ';
        std::stringstream ss(line);
        float x;
        while(ss >> x) v.push_back(x);
        if(!ss.eof()) return KPAX_IO_ERROR;  // a token that is not a number
    }
    if(v.size() % 6 != 0) return KPAX_IO_ERROR;
    *n = (uint32_t)(v.size() / 6);
    if(out)
    {
        if(cap < *n) return KPAX_INVALID_ARG;
        std::copy(v.begin(), v.end(), out);
    }
    return KPAX_OK;
}

int kpax_generate_environment(int32_t kind, uint64_t seed, uint32_t n, const float* ws_min, const float* ws_max,
                              const float* start, const float* goal, float* out, uint32_t cap, uint32_t* n_out)
{
    if(!ws_min || !ws_max || !n_out) return KPAX_INVALID_ARG;
    for(int d = 0; d < 3; ++d)
        if(!(ws_min[d] < ws_max[d])) return KPAX_INVALID_ARG;
    Rng rng{seed * 0x2545F4914F6CDD1Dull + (uint64_t)kind};
    std::vector<float> boxes;
    auto push = [&](const float* b) { boxes.insert(boxes.end(), b, b + 6); };
    const int max_tries = 1000;
    if(kind == 1)
    {
        for(uint32_t i = 0; i < n; ++i)
            for(int tr = 0; tr < max_tries; ++tr)
            {
                float b[6];
                for(int d = 0; d < 3; ++d)
                {
                    const double c = rng.uni(ws_min[d], ws_max[d]), e = rng.uni(0.05, 0.15);
                    b[d]           = (float)std::max<double>(ws_min[d], c - 0.5 * e);
                    b[3 + d]       = (float)std::min<double>(ws_max[d], c + 0.5 * e);
                }
                if(clear_of(b, start, goal, 0.02))
                {
                    push(b);
                    break;
                }
            }
    }
    else if(kind == 2)
    {
        for(uint32_t i = 0; i < n; ++i)
            for(int tr = 0; tr < max_tries; ++tr)
            {
                const double fx = rng.uni(0.1, 0.3), fy = rng.uni(0.1, 0.3);
                const double cx = rng.uni(ws_min[0] + 0.5 * fx, ws_max[0] - 0.5 * fx);
                const double cy = rng.uni(ws_min[1] + 0.5 * fy, ws_max[1] - 0.5 * fy);
                const float b[6] = {(float)(cx - 0.5 * fx), (float)(cy - 0.5 * fy), ws_min[2],
                                    (float)(cx + 0.5 * fx), (float)(cy + 0.5 * fy), ws_max[2]};
                // keep a column's footprint off the start and goal in plan view
                float flat[6];
                memcpy(flat, b, sizeof(b));
                float s2[3], g2[4];
                bool ok = true;
                if(start)
                {
                    s2[0] = start[0], s2[1] = start[1], s2[2] = 0.5f * (ws_min[2] + ws_max[2]);
                    ok    = ok && box_point_dist(flat, s2) > 0.3;
                }
                if(goal)
                {
                    g2[0] = goal[0], g2[1] = goal[1], g2[2] = 0.5f * (ws_min[2] + ws_max[2]), g2[3] = goal[3];
                    ok    = ok && box_point_dist(flat, g2) > (double)goal[3] + 0.3;
                }
                if(ok)
                {
                    push(b);
                    break;
                }
            }
    }
    else if(kind == 3)
    {
        // A wall across x at the middle of the workspace, tiled 40 x 40 in
        // (y, z), with one square passage a tenth of the workspace wide; the
        // rest of the budget goes to thin slabs scattered on both sides.
        // Lengths scale with the workspace extent (unit cube: a 0.1 passage).
        const int T     = 40;
        const double ex = ws_max[0] - ws_min[0], ey = ws_max[1] - ws_min[1], ez = ws_max[2] - ws_min[2];
        const double ext[3] = {ex, ey, ez};
        const double sc     = std::min(ex, std::min(ey, ez));
        const double xm = 0.5 * (ws_min[0] + ws_max[0]), half = 0.01 * ex;
        const double ty = ey / T, tz = ez / T;
        const double py = rng.uni(ws_min[1] + 0.25 * ey, ws_max[1] - 0.25 * ey);
        const double pz = rng.uni(ws_min[2] + 0.25 * ez, ws_max[2] - 0.25 * ez);
        const double pwy = 0.1 * ey, pwz = 0.1 * ez;
        for(int iy = 0; iy < T; ++iy)
            for(int iz = 0; iz < T; ++iz)
            {
                const double y0 = ws_min[1] + iy * ty, y1 = y0 + ty, z0 = ws_min[2] + iz * tz, z1 = z0 + tz;
                // tiles overlapping the passage window are cut back to its edges
                const double wy0 = py - 0.5 * pwy, wy1 = py + 0.5 * pwy, wz0 = pz - 0.5 * pwz, wz1 = pz + 0.5 * pwz;
                const bool hit = y1 > wy0 && y0 < wy1 && z1 > wz0 && z0 < wz1;
                if(!hit)
                {
                    const float b[6] = {(float)(xm - half), (float)y0, (float)z0, (float)(xm + half), (float)y1, (float)z1};
                    push(b);
                    continue;
                }
                // split the tile into the parts outside the window
                const double ys[4] = {y0, std::max(y0, wy0), std::min(y1, wy1), y1};
                const double zs[4] = {z0, std::max(z0, wz0), std::min(z1, wz1), z1};
                for(int a = 0; a < 3; ++a)
                    for(int bb = 0; bb < 3; ++bb)
                    {
                        if(a == 1 && bb == 1) continue;  // the window itself
                        if(ys[a + 1] - ys[a] <= 1e-9 || zs[bb + 1] - zs[bb] <= 1e-9) continue;
                        const float b[6] = {(float)(xm - half), (float)ys[a], (float)zs[bb],
                                            (float)(xm + half), (float)ys[a + 1], (float)zs[bb + 1]};
                        push(b);
                    }
            }
        const uint32_t wall = (uint32_t)(boxes.size() / 6);
        for(uint32_t i = wall; i < n; ++i)
            for(int tr = 0; tr < max_tries; ++tr)
            {
                float b[6];
                const int thin = (int)(rng.next() % 3);
                for(int d = 0; d < 3; ++d)
                {
                    const double e = (d == thin ? 0.01 : rng.uni(0.05, 0.12)) * ext[d];
                    const double c = rng.uni(ws_min[d] + 0.5 * e, ws_max[d] - 0.5 * e);
                    b[d]           = (float)(c - 0.5 * e);
                    b[3 + d]       = (float)(c + 0.5 * e);
                }
                // away from the dividing wall's passage and the endpoints
                const float pc[3] = {(float)xm, (float)py, (float)pz};
                if(box_point_dist(b, pc) > 0.15 * sc && clear_of(b, start, goal, 0.05 * sc) &&
                   (b[3] < xm - half - 0.02 * ex || b[0] > xm + half + 0.02 * ex))
                {
                    push(b);
                    break;
                }
            }
    }
    else
        return KPAX_INVALID_ARG;
    *n_out = (uint32_t)(boxes.size() / 6);
    if(out)
    {
        if(cap < *n_out) return KPAX_INVALID_ARG;
        std::copy(boxes.begin(), boxes.end(), out);
    }
    return KPAX_OK;
}

int kpax_verify_path(const kpax_config* cfg, const float* obstacles, uint32_t n_obs, const float* rows,
                     const int32_t* steps, uint32_t n_rows, const float* goal, float* max_dev)
{
    if(!cfg || !rows || !steps || !goal || (n_obs && !obstacles)) return KPAX_INVALID_ARG;
    if(!validate_config(*cfg).empty()) return KPAX_INVALID_ARG;
    return verify_path_fp64(*cfg, obstacles, n_obs, rows, steps, n_rows, goal, max_dev) ? KPAX_OK : KPAX_NOT_FOUND;
}

}  // extern "C"
