// kpax_oracle.cpp — CPU oracle of the Kino-PAX hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load liboracle_kpax.so.  The product (libkpax.so) never links or calls it.
//
// It restates, in plain sequential C++17, the reference's planner iteration
// (kranzowe/Kino-PAX-Fork; citations are path:line in that repository):
//   propagators   src/statePropagator/statePropagatorSpatialHash.cu:72-308,
//                 src/statePropagator/statePropagator.cu:325-370 (quad ode)
//   collision     src/collisionCheck/collisionCheck.cu:3-10,49-62 (broad phase)
//                 src/collisionCheck/spatialHash.cu:9-54,145-194 (grid)
//                 include/collisionCheck/spatialHash.cuh:42-60
//   regions       src/graphs/Graph.cu:99-205, 53-97
//   scores        src/graphs/Graph.cu:233-358 (the evident intent, see below)
//   compaction    src/planners/Planner.cu:52-80, KPAX.cu:303-351
//   iteration     src/planners/KPAX.cu:51-115, 358-543
//   goal metric   include/helper/helper.cuh:138-152
// Every floating-point expression keeps the reference's operation order and is
// compiled without contraction (-ffp-contract=off), templated on float (the
// precision the planner runs in) and double (the truth the float path is
// measured against).
//
// Deliberate, documented departures from the reference (each one makes the
// planner deterministic or removes undefined behaviour):
//   * RNG: Philox4x32-10 (Salmon et al., SC'11) keyed by the seed and counted
//     by (slot, iteration, query, purpose), replacing per-slot cuRAND XORWOW
//     state seeded from the wall clock (Planner.cu:37-50, KPAX.cu:90-91).
//   * steps = steps_min + floor(u * (steps_max - steps_min + 1)), u in [0,1)
//     (reference: 1 + (int)(u * 10) with u in (0,1], statePropagatorSpatialHash.cu:77).
//   * float sin/cos: the project's own Cody-Waite + Taylor routine below (the
//     reference calls CUDA's cosf/sinf, whose last-ulp results cannot be
//     reproduced on any other machine).  tan(t) = sin(t)/cos(t).
//   * Non-finite states fail the bound checks (NaN compares let them pass in
//     the reference, statePropagator.cu:282-293).
//   * Sub-region novelty (KPAX.cu:394-395, a read-then-exchange race) is
//     decided as if the expansion threads ran one at a time in slot order:
//     only the first valid child to reach an uncovered R2 cell counts as new.
//   * Score total: the sum of all R1 scores in a fixed order (1024 strided
//     partial sums, then a pairwise tree), reset every iteration (the
//     reference's reduction is undefined behaviour, Graph.cu:260-267,285,314).
//   * Goal: the smallest tree index among new nodes inside the goal ball
//     (the reference keeps whichever thread writes last, KPAX.cu:482-501).
//   * Every expanded parent leaves the frontier (kernel2 clears arbitrary
//     indices, KPAX.cu:411); children never overflow the tree cap
//     (KPAX.cu:327-333 can).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>
#include <cstdio>
#include <cstdlib>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/kpax.h"

namespace {

// ---------------------------------------------------------------------------
// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11).
// ---------------------------------------------------------------------------
struct U4 {
    uint32_t v[4];
};

inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
    for(int r = 0; r < 10; ++r)
    {
        if(r)
        {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.v[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.v[2];
        U4 n;
        n.v[0] = (uint32_t)(p1 >> 32) ^ c.v[1] ^ k0;
        n.v[1] = (uint32_t)p1;
        n.v[2] = (uint32_t)(p0 >> 32) ^ c.v[3] ^ k1;
        n.v[3] = (uint32_t)p0;
        c      = n;
    }
    return c;
}

inline float u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

enum : uint32_t { PURPOSE_CONTROL = 0, PURPOSE_STEPS_ADMIT = 1, PURPOSE_READMIT = 2 };

inline U4 draw(uint64_t seed, uint32_t slot, uint32_t iter, uint32_t query, uint32_t purpose)
{
    U4 c = {{slot, iter, query, purpose}};
    return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// ---------------------------------------------------------------------------
// float sin/cos: reduction by pi/2 with a three-part Cody-Waite split applied
// through fmaf, then Taylor polynomials of degree 13 (sin) and 12 (cos) on
// [-pi/4, pi/4], evaluated by Horner's rule with fmaf.  Coefficients are the
// float roundings of (-1)^k / n!.  The truncation error of both series on
// [-pi/4, pi/4] is below 1e-11, far under one float ulp.
// ---------------------------------------------------------------------------
inline void sincos_f32(float x, float* s, float* c)
{
    const float k  = std::rint(x * 0x1.45f306p-1f);
    const int   q  = (int)k;
    float       r  = std::fma(-k, 0x1.921fb6p+0f, x);
    r              = std::fma(-k, -0x1.777a5cp-25f, r);
    r              = std::fma(-k, -0x1.ee59dap-50f, r);
    const float r2 = r * r;
    float ps       = std::fma(r2, 0x1.612462p-33f, -0x1.ae6456p-26f);
    ps             = std::fma(r2, ps, 0x1.71de3ap-19f);
    ps             = std::fma(r2, ps, -0x1.a01a02p-13f);
    ps             = std::fma(r2, ps, 0x1.111112p-7f);
    ps             = std::fma(r2, ps, -0x1.555556p-3f);
    const float sp = std::fma(r * r2, ps, r);
    float pc       = std::fma(r2, 0x1.1eed8ep-29f, -0x1.27e4fcp-22f);
    pc             = std::fma(r2, pc, 0x1.a01a02p-16f);
    pc             = std::fma(r2, pc, -0x1.6c16c2p-10f);
    pc             = std::fma(r2, pc, 0x1.555556p-5f);
    pc             = std::fma(r2, pc, -0.5f);
    const float cp = std::fma(r2, pc, 1.0f);
    switch(q & 3)
    {
        case 0: *s = sp; *c = cp; break;
        case 1: *s = cp; *c = -sp; break;
        case 2: *s = -sp; *c = -cp; break;
        default: *s = -cp; *c = sp; break;
    }
}

template<class T> struct Math;
template<> struct Math<float> {
    static void sincos(float x, float* s, float* c) { sincos_f32(x, s, c); }
    static float tan_from(float s, float sec, float) { return s * sec; }
    static float sqrt(float x) { return std::sqrt(x); }
    static float fma(float a, float b, float c) { return std::fma(a, b, c); }
    // sin/cos of a + d from those of a (the kernel's sincos_shift): angle-sum
    // identities with Taylor series of sin d and cos d through d^11 / d^10
    // for |d| <= 1, else sincos_f32 of the rounded sum.
    static void sincos_shift(float sa, float ca, float a, float d, float* s, float* c)
    {
        if(std::fabs(d) <= 1.0f)
        {
            const float d2 = d * d;
            const float ps = fma(
              d2, fma(d2, fma(d2, fma(d2, -2.5052108e-8f, 2.7557319e-6f), -1.9841270e-4f), 8.3333333e-3f),
              -1.6666667e-1f);
            const float pc =
              fma(d2, fma(d2, fma(d2, fma(d2, -2.7557319e-7f, 2.4801587e-5f), -1.3888889e-3f), 4.1666667e-2f), -0.5f);
            const float sd = fma(d * d2, ps, d);
            const float cd = fma(d2, pc, 1.0f);
            *s             = fma(sa, cd, ca * sd);
            *c             = fma(ca, cd, -(sa * sd));
        }
        else
            sincos_f32(a + d, s, c);
    }
};
template<> struct Math<double> {
    static void sincos(double x, double* s, double* c)
    {
        *s = std::sin(x);
        *c = std::cos(x);
    }
    static double tan_from(double, double, double x) { return std::tan(x); }
    static double sqrt(double x) { return std::sqrt(x); }
    static double fma(double a, double b, double c) { return std::fma(a, b, c); }
    static void sincos_shift(double, double, double a, double d, double* s, double* c) { sincos(a + d, s, c); }
};

// Float truncation toward zero clamped to [0, len-1]; NaN maps to 0.  Equal
// to the reference's (int) cast + clamp for every finite value
// (Graph.cu:107-109, spatialHash.cuh:44-51), defined for all inputs.
inline int clamp_index(float f, int len)
{
    if(!(f >= 0.0f)) return 0;
    if(f >= (float)len) return len - 1;
    return (int)f;
}

// ---------------------------------------------------------------------------
// Collision: broad phase against one obstacle (collisionCheck.cu:3-10) and the
// motion check over all obstacles (collisionCheck.cu:49-62).  This is the
// verdict specification: touching faces count as free.
// ---------------------------------------------------------------------------
template<class T>
bool segment_free(const T bbmin[3], const T bbmax[3], const float* obs, uint32_t n)
{
    for(uint32_t i = 0; i < n; ++i)
    {
        const float* o = obs + 6 * (size_t)i;
        bool separated = false;
        for(int d = 0; d < 3; ++d)
            if(bbmax[d] <= (T)o[d] || (T)o[3 + d] <= bbmin[d])
            {
                separated = true;
                break;
            }
        if(!separated) return false;
    }
    return true;
}


// Uniform grid over the workspace: the reference's spatial hash (build
// spatialHash.cu:9-54, query :145-194) as CSR lists without its 16-per-cell
// cap, with per-axis bounds.  The float planner's collision check walks the
// cells of the segment's box and applies the broad-phase verdict to every
// obstacle listed there.  The verdict equals the brute-force one: an obstacle
// that is not separated from the box shares a point with it, and the cell map
// (x - min) * cells-per-metre, truncated and clamped, is monotone, so both
// cell ranges hold that point's cell.
struct Grid {
    int D[3];
    float mn[3], inv[3];
    std::vector<int32_t> start, items;
    Grid(const kpax_config& c, const float* obs, uint32_t n)
    {
        for(int d = 0; d < 3; ++d)
        {
            D[d]   = c.grid_dims[d];
            mn[d]  = c.ws_min[d];
            inv[d] = (float)D[d] / (c.ws_max[d] - c.ws_min[d]);
        }
        const size_t G = (size_t)D[0] * D[1] * D[2];
        start.assign(G + 1, 0);
        for(int pass = 0; pass < 2; ++pass)
        {
            std::vector<int32_t> fill;
            if(pass)
            {
                for(size_t g = 0; g < G; ++g) start[g + 1] += start[g];
                items.resize(start[G]);
                fill.assign(start.begin(), start.end() - 1);
            }
            for(uint32_t i = 0; i < n; ++i)
            {
                int a[3], b[3];
                range(obs + 6 * (size_t)i, obs + 6 * (size_t)i + 3, a, b);
                for(int z = a[2]; z <= b[2]; ++z)
                    for(int y = a[1]; y <= b[1]; ++y)
                        for(int x = a[0]; x <= b[0]; ++x)
                        {
                            const size_t g = x + (size_t)D[0] * (y + (size_t)D[1] * z);
                            if(pass)
                                items[fill[g]++] = (int32_t)i;
                            else
                                start[g + 1]++;
                        }
            }
        }
    }
    void range(const float* lo, const float* hi, int a[3], int b[3]) const
    {
        for(int d = 0; d < 3; ++d)
        {
            a[d] = clamp_index((lo[d] - mn[d]) * inv[d], D[d]);
            b[d] = clamp_index((hi[d] - mn[d]) * inv[d], D[d]);
        }
    }
    bool free(const float bbmin[3], const float bbmax[3], const float* obs) const
    {
        int a[3], b[3];
        range(bbmin, bbmax, a, b);
        for(int z = a[2]; z <= b[2]; ++z)
            for(int y = a[1]; y <= b[1]; ++y)
                for(int x = a[0]; x <= b[0]; ++x)
                {
                    const size_t g = x + (size_t)D[0] * (y + (size_t)D[1] * z);
                    for(int32_t k = start[g]; k < start[g + 1]; ++k)
                    {
                        const float* o = obs + 6 * (size_t)items[k];
                        bool separated = false;
                        for(int d = 0; d < 3 && !separated; ++d)
                            separated = bbmax[d] <= o[d] || o[3 + d] <= bbmin[d];
                        if(!separated) return false;
                    }
                }
        return true;
    }
};

struct Env {
    const kpax_config* cfg;
    const float* obs;
    uint32_t n_obs;
    const Grid* grid = nullptr;  // float path only; nullptr = brute force over all obstacles
};

template<class T>
bool in_box(T v, T lo, T hi)
{
    return v >= lo && v <= hi;  // false for NaN
}

template<class T>
bool step_checks_ws_and_collision(const Env& e, const T p0[3], const T p1[3])
{
    const kpax_config& c = *e.cfg;
    for(int d = 0; d < 3; ++d)
        if(!in_box<T>(p1[d], (T)c.ws_min[d], (T)c.ws_max[d])) return false;
    T bbmin[3], bbmax[3];
    for(int d = 0; d < 3; ++d)
    {
        if(p0[d] > p1[d])
        {
            bbmin[d] = p1[d];
            bbmax[d] = p0[d];
        }
        else
        {
            bbmin[d] = p0[d];
            bbmax[d] = p1[d];
        }
    }
    if constexpr(std::is_same<T, float>::value)
        if(e.grid) return e.grid->free(bbmin, bbmax, e.obs);
    return segment_free<T>(bbmin, bbmax, e.obs, e.n_obs);
}

// Double integrator, statePropagatorSpatialHash.cu:72-134 (and the MATLAB
// restatement viz/KGMT_3D_GPU.m:283-288).  Returns the 1-based failing step or 0.
template<class T>
int propagate_di(const Env& e, const T* x0, const T* u, int steps, T* x1)
{
    const kpax_config& c = *e.cfg;
    const T dt = (T)c.dt, vmin = (T)c.vel_min, vmax = (T)c.vel_max;
    T x = x0[0], y = x0[1], z = x0[2], vx = x0[3], vy = x0[4], vz = x0[5];
    const T ax = u[0], ay = u[1], az = u[2];
    int fail = 0;
    for(int i = 0; i < steps; ++i)
    {
        const T p0[3] = {x, y, z};
        if(!c.di_textbook)
        {
            x += (vx + (vx + T(2) * (vx + ax * dt / T(2)) + (vx + ax * dt))) * dt / T(6);
            y += (vy + (vy + T(2) * (vy + ay * dt / T(2)) + (vy + ay * dt))) * dt / T(6);
            z += (vz + (vz + T(2) * (vz + az * dt / T(2)) + (vz + az * dt))) * dt / T(6);
            vx += (ax + T(2) * ax + T(2) * ax + ax) * dt / T(6);
            vy += (ay + T(2) * ay + T(2) * ay + ay) * dt / T(6);
            vz += (az + T(2) * az + T(2) * az + az) * dt / T(6);
        }
        else
        {
            x += vx * dt + T(0.5) * ax * dt * dt;
            y += vy * dt + T(0.5) * ay * dt * dt;
            z += vz * dt + T(0.5) * az * dt * dt;
            vx += ax * dt;
            vy += ay * dt;
            vz += az * dt;
        }
        if(!in_box(vx, vmin, vmax) || !in_box(vy, vmin, vmax) || !in_box(vz, vmin, vmax))
        {
            fail = i + 1;
            break;
        }
        const T p1[3] = {x, y, z};
        if(!step_checks_ws_and_collision<T>(e, p0, p1))
        {
            fail = i + 1;
            break;
        }
    }
    x1[0] = x, x1[1] = y, x1[2] = z, x1[3] = vx, x1[4] = vy, x1[5] = vz;
    return fail;
}

// Dubins airplane, statePropagatorSpatialHash.cu:139-229.  The controls are
// (a, yawRate, pitchRate) in the reference's draw order (:141-143).
template<class T>
int propagate_dubins(const Env& e, const T* x0, const T* u, int steps, T* x1)
{
    const kpax_config& c = *e.cfg;
    const T dt = (T)c.dt, h = T(0.5) * dt, d6 = dt / T(6);
    T x = x0[0], y = x0[1], z = x0[2], yaw = x0[3], pitch = x0[4], v = x0[5];
    const T a = u[0], yr = u[1], pr = u[2];
    int fail = 0;
    for(int i = 0; i < steps; ++i)
    {
        const T p0[3] = {x, y, z};
        T cp, sp, cy, sy, cpm, spm, cym, sym, cpe, spe, cye, sye;
        Math<T>::sincos(pitch, &sp, &cp);
        Math<T>::sincos(yaw, &sy, &cy);
        Math<T>::sincos(pitch + h * pr, &spm, &cpm);
        Math<T>::sincos(yaw + h * yr, &sym, &cym);
        Math<T>::sincos(pitch + dt * pr, &spe, &cpe);
        Math<T>::sincos(yaw + dt * yr, &sye, &cye);
        const T A  = v + h * a;
        const T B  = v + dt * a;
        const T Tx = A * cpm * cym;
        const T Ty = A * cpm * sym;
        const T Tz = A * spm;
        x += d6 * (v * cp * cy + T(2) * (Tx + Tx) + B * cpe * cye);
        y += d6 * (v * cp * sy + T(2) * (Ty + Ty) + B * cpe * sye);
        z += d6 * (v * sp + T(2) * (Tz + Tz) + B * spe);
        yaw += dt * yr;
        pitch += dt * pr;
        v += d6 * (a + T(2) * (a + a) + a);
        if(!in_box(v, (T)c.vel_min, (T)c.vel_max) || !in_box(pitch, (T)c.pitch_min, (T)c.pitch_max))
        {
            fail = i + 1;
            break;
        }
        const T p1[3] = {x, y, z};
        if(!step_checks_ws_and_collision<T>(e, p0, p1))
        {
            fail = i + 1;
            break;
        }
    }
    x1[0] = x, x1[1] = y, x1[2] = z, x1[3] = yaw, x1[4] = pitch, x1[5] = v;
    return fail;
}

template<class T> struct QuadK {
    T negnu, negmu, g, mi, k1x, k1y, k1z, ixi, iyi, izi;
};

// Coefficients exactly as the reference's macros fold them in float
// (config.h:300-308: (IY - IZ) / IX, 1 / IX, 1.0f / MASS, -NU, -MU).
template<class T>
QuadK<T> quad_constants(const kpax_config& c)
{
    const float ix = c.inertia[0], iy = c.inertia[1], iz = c.inertia[2];
    QuadK<T> k;
    k.negnu = (T)(-c.nu);
    k.negmu = (T)(-c.mu);
    k.g     = (T)c.gravity;
    k.mi    = (T)(1.0f / c.mass);
    k.k1x   = (T)((iy - iz) / ix);
    k.k1y   = (T)((iz - ix) / iy);
    k.k1z   = (T)((ix - iy) / iz);
    k.ixi   = (T)(1.0f / ix);
    k.iyi   = (T)(1.0f / iy);
    k.izi   = (T)(1.0f / iz);
    return k;
}

// ode(), src/statePropagator/statePropagator.cu:325-370: the reference's
// terms, each sum of products as nested fused multiply-adds (one rounding per
// fma), the same nesting as the kernel's quad_ode.
// sn/cn: sin and cos of the stage's roll, pitch and yaw (theta is the pitch,
// for the double path's tan).
template<class T>
void quad_ode(T* out, const T* x, const T* hh, T delta, const T* u, const QuadK<T>& k, const T* sn, const T* cn,
              T theta)
{
    auto fma      = [](T a, T b, T c) { return Math<T>::fma(a, b, c); };
    const T uu    = fma(delta, hh[6], x[6]), v = fma(delta, hh[7], x[7]), w = fma(delta, hh[8], x[8]);
    const T p     = fma(delta, hh[9], x[9]), q = fma(delta, hh[10], x[10]), r = fma(delta, hh[11], x[11]);
    const T sf = sn[0], cf = cn[0], st = sn[1], ct = cn[1], ss = sn[2], cs = cn[2];
    const T sec  = T(1) / ct;  // the kernel's single division (exact in the double path's sense)
    const T tt   = Math<T>::tan_from(st, sec, theta);
    const T sfst = sf * st, cfst = cf * st;
    out[0]       = fma(ct * cs, uu, fma(fma(sfst, cs, -(cf * ss)), v, fma(cfst, cs, sf * ss) * w));
    out[1]       = fma(ct * ss, uu, fma(fma(sfst, ss, cf * cs), v, fma(cfst, ss, -(sf * cs)) * w));
    out[2]       = fma(-st, uu, fma(sf * ct, v, cf * ct * w));
    const T qr   = fma(q, sf, r * cf);
    out[3]       = fma(qr, tt, p);
    out[4]       = fma(q, cf, -(r * sf));
    out[5]       = qr * sec;
    const T mX   = k.mi * (k.negnu * Math<T>::sqrt(fma(uu, uu, fma(v, v, w * w))));
    const T gct  = k.g * ct;
    out[6]       = fma(mX, uu, fma(-k.g, st, fma(r, v, -(q * w))));
    out[7]       = fma(mX, v, fma(gct, sf, fma(p, w, -(r * uu))));
    out[8]       = fma(k.mi, u[0], fma(mX, w, fma(gct, cf, fma(q, uu, -(p * v)))));
    const T LMN  = k.negmu * Math<T>::sqrt(fma(p, p, fma(q, q, r * r)));
    out[9]       = fma(k.ixi, u[1], fma(k.ixi * LMN, p, k.k1x * q * r));
    out[10]      = fma(k.iyi, u[2], fma(k.iyi * LMN, q, k.k1y * p * r));
    out[11]      = fma(k.izi, u[3], fma(k.izi * LMN, r, k.k1z * p * q));
}

// Quadcopter RK4, statePropagatorSpatialHash.cu:234-308.
template<class T>
int propagate_quad(const Env& e, const T* x0, const T* u, int steps, T* x1)
{
    const kpax_config& c = *e.cfg;
    const QuadK<T> k = quad_constants<T>(c);
    const T dt = (T)c.dt, half = T(0.5) * dt, d6 = dt / T(6);
    const T vmin = (T)c.vel_min, vmax = (T)c.vel_max;
    T h0[12], k1[12], k2[12], k3[12], k4[12];
    const T zero[12] = {};
    for(int j = 0; j < 12; ++j) h0[j] = x0[j];
    int fail = 0;
    for(int i = 0; i < steps; ++i)
    {
        const T p0[3] = {h0[0], h0[1], h0[2]};
        // stage angles: the state's own, then shifted by delta * rate
        T sb[3], cb[3], sn[3], cn[3];
        for(int j = 0; j < 3; ++j) Math<T>::sincos(h0[3 + j], &sb[j], &cb[j]);
        quad_ode<T>(k1, h0, zero, T(1), u, k, sb, cb, h0[4]);
        for(int j = 0; j < 3; ++j) Math<T>::sincos_shift(sb[j], cb[j], h0[3 + j], half * k1[3 + j], &sn[j], &cn[j]);
        quad_ode<T>(k2, h0, k1, half, u, k, sn, cn, h0[4] + half * k1[4]);
        for(int j = 0; j < 3; ++j) Math<T>::sincos_shift(sb[j], cb[j], h0[3 + j], half * k2[3 + j], &sn[j], &cn[j]);
        quad_ode<T>(k3, h0, k2, half, u, k, sn, cn, h0[4] + half * k2[4]);
        for(int j = 0; j < 3; ++j) Math<T>::sincos_shift(sb[j], cb[j], h0[3 + j], dt * k3[3 + j], &sn[j], &cn[j]);
        quad_ode<T>(k4, h0, k3, dt, u, k, sn, cn, h0[4] + dt * k3[4]);
        for(int j = 0; j < 12; ++j)
            h0[j] = Math<T>::fma(d6, Math<T>::fma(T(2), k3[j], Math<T>::fma(T(2), k2[j], k1[j])) + k4[j], h0[j]);
        if(!in_box(h0[6], vmin, vmax) || !in_box(h0[7], vmin, vmax) || !in_box(h0[8], vmin, vmax))
        {
            fail = i + 1;
            break;
        }
        const T p1[3] = {h0[0], h0[1], h0[2]};
        if(!step_checks_ws_and_collision<T>(e, p0, p1))
        {
            fail = i + 1;
            break;
        }
    }
    for(int j = 0; j < 12; ++j) x1[j] = h0[j];
    return fail;
}

// In float the per-step constants fold as the reference's macros do
// (0.5f * STEP_SIZE, STEP_SIZE / 6.0f); in double they are exact in the
// float step dt.
template<class T>
int propagate(const Env& e, const T* x0, const T* u, int steps, T* x1)
{
    switch(e.cfg->model)
    {
        case KPAX_MODEL_DOUBLE_INTEGRATOR: return propagate_di<T>(e, x0, u, steps, x1);
        case KPAX_MODEL_DUBINS_AIRPLANE: return propagate_dubins<T>(e, x0, u, steps, x1);
        case KPAX_MODEL_QUADCOPTER: return propagate_quad<T>(e, x0, u, steps, x1);
        default: return -1;
    }
}

int state_dim(int model) { return model == KPAX_MODEL_QUADCOPTER ? 12 : 6; }
int control_dim(int model) { return model == KPAX_MODEL_QUADCOPTER ? 4 : 3; }

// ---------------------------------------------------------------------------
// Regions: getRegion / getSubRegion (Graph.cu:99-205) with the R1 min corner
// of initializeRegions_kernel (Graph.cu:58-70), per workspace axis.
// ---------------------------------------------------------------------------
struct Regions {
    int L[3], L2[3];
    float mn[3], ext[3], size[3];
    int r1_count, r2_per;
    explicit Regions(const kpax_config& c)
    {
        for(int d = 0; d < 3; ++d)
        {
            L[d]    = c.r1_len[d];
            L2[d]   = c.r2_len[d];
            mn[d]   = c.ws_min[d];
            ext[d]  = c.ws_max[d] - c.ws_min[d];
            size[d] = ext[d] / (float)L[d];
        }
        r1_count = L[0] * L[1] * L[2];
        r2_per   = L2[0] * L2[1] * L2[2];
    }
    void locate(const float* p, int* r1, int* r2) const
    {
        int idx[3], jdx[3];
        for(int d = 0; d < 3; ++d) idx[d] = clamp_index((float)L[d] * (p[d] - mn[d]) / ext[d], L[d]);
        for(int d = 0; d < 3; ++d)
        {
            const float lo = mn[d] + (float)idx[d] * size[d];
            jdx[d]         = clamp_index((float)L2[d] * (p[d] - lo) / size[d], L2[d]);
        }
        *r1 = (idx[0] * L[1] + idx[1]) * L[2] + idx[2];
        *r2 = *r1 * r2_per + (jdx[0] * L2[1] + jdx[1]) * L2[2] + jdx[2];
    }
};

// Syclop-style region score (Graph.cu:243-255) and its normalisation
// (Graph.cu:297-304).  The total is summed in a fixed order: 1024 strided
// partial sums, then a pairwise tree halving 512, 256, ..., 1.
void scores(const kpax_config& c, const Regions& R, const int32_t* valid, const int32_t* total,
            const int32_t* covered, float* out, float* total_out)
{
    const float vol = R.size[0] * R.size[1] * R.size[2];
    std::vector<float> raw(R.r1_count, 0.0f);
    for(int r = 0; r < R.r1_count; ++r)
    {
        if(valid[r] > 0)
        {
            const float cov = (float)covered[r] / (float)R.r2_per;
            const float ev  = c.eps + (float)valid[r];
            const float fv  = ev / (ev + (float)(total[r] - valid[r])) * vol;
            const float n   = (float)total[r];
            raw[r]          = (fv * fv) * (fv * fv) / ((1.0f + cov) * (1.0f + n * n));
        }
    }
    float part[1024];
    for(int t = 0; t < 1024; ++t)
    {
        float s = 0.0f;
        for(int r = t; r < R.r1_count; r += 1024) s += raw[r];
        part[t] = s;
    }
    for(int off = 512; off >= 1; off >>= 1)
        for(int t = 0; t < off; ++t) part[t] = part[t] + part[t + off];
    const float tot = part[0];
    for(int r = 0; r < R.r1_count; ++r) out[r] = valid[r] == 0 ? 1.0f : c.eps + raw[r] / tot;
    if(total_out) *total_out = tot;
}

// fAccept of updateFrontier (KPAX.cu:527-528): the host computes
// (h_itr_ * EPSILON) in float and pow(treeAddSize, 5) in double.
float f_accept(const kpax_config& c, uint32_t iter, uint32_t tree_plus_new)
{
    const float  tas = 1.0f - ((float)tree_plus_new / (float)c.max_tree);
    const double t   = (double)tas;
    const double p5  = t * t * t * t * t;
    return (float)((double)((float)iter * c.eps) * p5);
}

inline float goal_distance(const float* p, const float* g)
{
    const float dx = p[0] - g[0], dy = p[1] - g[1], dz = p[2] - g[2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// ---------------------------------------------------------------------------
// The iteration's decisions, each as the reference states it.
// ---------------------------------------------------------------------------

// Children per repeat-list entry and children of the iteration
// (propagateFrontier, KPAX.cu:323-333): the block size (fan-out 32) while
// W * 32 fits the remaining capacity, else min(int(float(remaining) /
// float(W)), 32), and 1 when that is 0.  M is capped at the remaining
// capacity (the reference's kernel2 can overflow the tree there).
inline void fan_out(const kpax_config& c, uint64_t W, uint32_t remaining, uint32_t* fan_out, uint64_t* M_out)
{
    uint32_t fan = c.fanout;
    uint64_t M   = W * fan;
    if(M > remaining)
    {
        fan = (uint32_t)(int)((float)remaining / (float)W);
        if(fan > c.fanout) fan = c.fanout;
        if(fan == 0) fan = 1;
        M = std::min<uint64_t>(W * fan, remaining);
    }
    *fan_out = fan;
    *M_out   = M;
}

// A valid child joins the next frontier on a score draw strictly below its
// region's score of the previous iteration (KPAX.cu:393, 433:
// curand_uniform(&randSeed) < vertexScores[x1Vertex]) ...
inline bool admit_by_score(uint32_t word, float score) { return u01(word) < score; }

// ... or when it is the first to reach its R2 cell (!activeSubVertices, KPAX.cu:393-395).

// An old node re-enters the frontier on a draw at or below its region's
// score plus fAccept (KPAX.cu:512: curand_uniform(&seed) <= vertexScores + fAccept).
inline bool readmit(uint32_t word, float score, float fa) { return u01(word) <= score + fa; }

// Repeat count of a new node: 15 while its region holds fewer than 10 valid
// children, else 1 (KPAX.cu:476-479).
inline uint32_t new_node_repeat(const kpax_config& c, int32_t region_valid)
{
    return region_valid < c.sparse_valid_below ? c.sparse_repeat : c.dense_repeat;
}

struct PlanOut {
    int32_t status = KPAX_NOT_FOUND;
    uint32_t iterations = 0, tree_size = 0;
    int32_t goal = -1;
    uint64_t expansions = 0;
};

// ---------------------------------------------------------------------------
// Learned control proposal: no counterpart in the reference (its controls are
// uniform draws, statePropagatorSpatialHash.cu:236-241).  Restated from the
// build's own definition: features of a node, the occupancy lattice they read
// and the Gaussian sampler of a child's control.  Plain float operations in
// source order (fma where the definition says fma), so the results are
// bit-identical to the GPU's.
// ---------------------------------------------------------------------------
namespace prop {

constexpr int kOcc = 16;

struct Feat {
    float mid[3], half[3], mn[3], cell[3], inv_pi, inv_vel, rate;
    int S;
};

Feat feat_params(const kpax_config& c)
{
    Feat F;
    for(int d = 0; d < 3; ++d)
    {
        F.mid[d]  = 0.5f * (c.ws_min[d] + c.ws_max[d]);
        F.half[d] = 0.5f * (c.ws_max[d] - c.ws_min[d]);
        F.mn[d]   = c.ws_min[d];
        F.cell[d] = (c.ws_max[d] - c.ws_min[d]) / (float)kOcc;
    }
    F.inv_pi      = 0.31830988618379067f;
    const float v = std::max(std::fabs(c.vel_min), std::fabs(c.vel_max));
    F.inv_vel     = v > 0.0f ? 1.0f / v : 1.0f;
    F.rate        = 0.25f;
    F.S           = c.model == KPAX_MODEL_QUADCOPTER ? 12 : 6;
    return F;
}

// cell (a, b, c) set when an obstacle overlaps its open box
void occupancy(const Feat& F, const float* obs, uint32_t n, uint32_t* bits)
{
    for(int i = 0; i < kOcc * kOcc * kOcc / 32; ++i) bits[i] = 0;
    for(int a = 0; a < kOcc; ++a)
        for(int b = 0; b < kOcc; ++b)
            for(int c = 0; c < kOcc; ++c)
            {
                const float lo[3] = {F.mn[0] + (float)a * F.cell[0], F.mn[1] + (float)b * F.cell[1],
                                     F.mn[2] + (float)c * F.cell[2]};
                const float hi[3] = {F.mn[0] + (float)(a + 1) * F.cell[0], F.mn[1] + (float)(b + 1) * F.cell[1],
                                     F.mn[2] + (float)(c + 1) * F.cell[2]};
                bool hit = false;
                for(uint32_t o = 0; o < n && !hit; ++o)
                {
                    const float* q = obs + 6 * (size_t)o;
                    hit = q[0] < hi[0] && q[3] > lo[0] && q[1] < hi[1] && q[4] > lo[1] && q[2] < hi[2] && q[5] > lo[2];
                }
                const int i = (a * kOcc + b) * kOcc + c;
                if(hit) bits[i >> 5] |= 1u << (i & 31);
            }
}

int cell_index(float x, float mn, float cell)
{
    float v = (x - mn) / cell + 0.5f;
    if(!(v >= -8.0f)) v = -8.0f;
    if(v > 24.0f) v = 24.0f;
    return (int)std::floor(v) - 2;
}

float bit(const uint32_t* occ, int a, int b, int c)
{
    if(a < 0 || b < 0 || c < 0 || a >= kOcc || b >= kOcc || c >= kOcc) return 1.0f;
    const int i = (a * kOcc + b) * kOcc + c;
    return (occ[i >> 5] >> (i & 31)) & 1u ? 1.0f : 0.0f;
}

void features(const Feat& F, const uint32_t* occ, const float* x, const float* g, float* f)
{
    for(int d = 0; d < 3; ++d) f[d] = (x[d] - F.mid[d]) / F.half[d];
    for(int d = 0; d < 3; ++d) f[3 + d] = x[3 + d] * F.inv_pi;
    for(int d = 0; d < 3; ++d) f[6 + d] = x[6 + d] * F.inv_vel;
    for(int d = 0; d < 3; ++d) f[9 + d] = x[9 + d] * F.rate;
    for(int d = 0; d < 3; ++d) f[12 + d] = (g[d] - x[d]) / F.half[d];
    const int a0 = cell_index(x[0], F.mn[0], F.cell[0]), b0 = cell_index(x[1], F.mn[1], F.cell[1]),
              c0 = cell_index(x[2], F.mn[2], F.cell[2]);
    for(int i = 0; i < 4; ++i)
        for(int j = 0; j < 4; ++j)
            for(int k = 0; k < 4; ++k) f[15 + 16 * i + 4 * j + k] = bit(occ, a0 + i, b0 + j, c0 + k);
    f[79] = 0.0f;
}

// ln x for x in (0, 1]: exponent * ln 2 + 2 atanh((m-1)/(m+1)), odd series to s^11
float ln_f32(float x)
{
    uint32_t bits;
    std::memcpy(&bits, &x, 4);
    int e               = (int)(bits >> 23) - 127;
    const uint32_t mb   = (bits & 0x007fffffu) | 0x3f800000u;
    float m;
    std::memcpy(&m, &mb, 4);
    if(m > 1.41421356f)
    {
        m = m * 0.5f;
        e = e + 1;
    }
    const float s  = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p        = std::fma(s2, 0.09090909f, 0.11111111f);
    p              = std::fma(s2, p, 0.14285714f);
    p              = std::fma(s2, p, 0.2f);
    p              = std::fma(s2, p, 0.33333333f);
    const float lm = 2.0f * std::fma(s * s2, p, s);
    return std::fma((float)e, 0.69314718f, lm);
}

void box_muller(uint32_t w0, uint32_t w1, float* z0, float* z1)
{
    const float u1 = 1.0f - u01(w0);
    const float u2 = u01(w1);
    const float r  = std::sqrt(-2.0f * ln_f32(u1));
    float sn, cs;
    sincos_f32(6.28318531f * u2, &sn, &cs);
    *z0 = r * cs;
    *z1 = r * sn;
}

void propose(const float* mean, const float* sigma, const uint32_t* w, const kpax_config& c, float* u)
{
    float z[4];
    box_muller(w[0], w[1], &z[0], &z[1]);
    box_muller(w[2], w[3], &z[2], &z[3]);
    for(int j = 0; j < 4; ++j)
    {
        float un = mean[j] + sigma[j] * z[j];
        un       = un < -1.0f ? -1.0f : (un > 1.0f ? 1.0f : un);
        const float rng = c.ctrl_max[j] - c.ctrl_min[j];
        u[j]     = c.ctrl_min[j] + (un + 1.0f) * 0.5f * rng;
    }
}

}  // namespace prop

// Per-iteration record of a plan (kTraceWords words per iteration).
constexpr int kTraceWords = 8;
enum : int { TR_W = 0, TR_FAN, TR_M, TR_ADMITTED, TR_TS, TR_FA_BITS, TR_GOAL, TR_NEXT_W };

struct Planner {
    const kpax_config& c;
    Env env;
    Regions R;
    int S, C;
    uint32_t N;
    // tree, AoS per node
    std::vector<float> st, ct;
    std::vector<int32_t> steps, parent, reg;
    std::vector<int32_t> valid, total, covered;
    std::vector<int32_t> r2_first;
    std::vector<float> score;
    std::vector<uint32_t> weight;  // repeat count of every node in the last repeat list built
    float score_total = 0.0f;

    Grid grid;

    Planner(const kpax_config& cfg, const float* obs, uint32_t n_obs)
        : c(cfg), env{&cfg, obs, n_obs}, R(cfg), S(state_dim(cfg.model)), C(control_dim(cfg.model)),
          N(cfg.max_tree), grid(cfg, obs, n_obs)
    {
        env.grid = &grid;
        st.assign((size_t)N * S, 0.0f);
        ct.assign((size_t)N * C, 0.0f);
        steps.assign(N, 0);
        parent.assign(N, -1);
        reg.assign(N, 0);
        valid.assign(R.r1_count, 0);
        total.assign(R.r1_count, 0);
        covered.assign(R.r1_count, 0);
        r2_first.assign((size_t)R.r1_count * R.r2_per, std::numeric_limits<int32_t>::max());
        score.assign(R.r1_count, 0.0f);
        weight.assign(N, 0u);
    }

    // Learned control proposal (the build's own addition): with probability
    // `mix` a child draws its control from its parent's Gaussian, given here
    // per tree node as 8 x max_tree floats (4 means, 4 sigmas; the GPU's
    // network outputs, injected), instead of uniformly.
    const float* prop_table = nullptr;
    float mix               = 0.0f;

    // trace: kTraceWords words per iteration for up to max_iter iterations (optional)
    PlanOut plan(const float* start, const float* goal, uint64_t seed, uint32_t query, uint32_t* trace = nullptr)
    {
        PlanOut out;
        for(int d = 0; d < S; ++d) st[d] = start[d];
        parent[0] = -1;
        int r1, r2;
        R.locate(start, &r1, &r2);
        reg[0]      = r1;
        uint32_t ts = 1;
        // frontier repeat list: the root, root_repeat times (KPAX.cu:75-77)
        std::vector<uint32_t> rep(c.root_repeat, 0u);
        uint64_t W = c.root_repeat;
        weight[0]  = c.root_repeat;
        std::vector<float> cs((size_t)N * S), cc((size_t)N * C);
        std::vector<int32_t> csteps(N), cpar(N), creg(N), cr2(N);
        std::vector<uint8_t> adm(N), cvalid(N);
        std::vector<uint32_t> cword(N);
        for(uint32_t t = 1;; ++t)
        {
            if(t > c.max_iter)
            {
                out.status = KPAX_NOT_FOUND;
                break;
            }
            const uint32_t remaining = N - ts;
            if(remaining == 0)
            {
                out.status = KPAX_BUDGET_EXHAUSTED;
                break;
            }
            out.iterations = t;
            uint32_t fan;
            uint64_t M;
            fan_out(c, W, remaining, &fan, &M);
            out.expansions += M;
            // expansion (propagateFrontier_kernel1/2, KPAX.cu:358-439)
#pragma omp parallel for schedule(dynamic, 256)
            for(int64_t ci = 0; ci < (int64_t)M; ++ci)
            {
                const uint32_t cix = (uint32_t)ci;
                const uint32_t p   = rep[cix / fan];
                const U4 dc        = draw(seed, cix, t, query, PURPOSE_CONTROL);
                const U4 ds        = draw(seed, cix, t, query, PURPOSE_STEPS_ADMIT);
                float u[4];
                for(int i = 0; i < C; ++i) u[i] = c.ctrl_min[i] + u01(dc.v[i]) * (c.ctrl_max[i] - c.ctrl_min[i]);
                if(prop_table && mix > 0.0f && u01(ds.v[2]) < mix)
                {
                    const U4 dp = draw(seed, cix, t, query, 3u);  // the proposal's purpose word
                    float mean[4], sg[4];
                    for(int j = 0; j < 4; ++j)
                    {
                        mean[j] = prop_table[(size_t)j * N + p];
                        sg[j]   = prop_table[(size_t)(4 + j) * N + p];
                    }
                    prop::propose(mean, sg, dp.v, c, u);
                }
                const int span = c.steps_max - c.steps_min + 1;
                int k          = (int)(u01(ds.v[0]) * (float)span);
                if(k > span - 1) k = span - 1;
                const int nsteps = c.steps_min + k;
                float* x1        = &cs[(size_t)cix * S];
                const int fail   = propagate<float>(env, &st[(size_t)p * S], u, nsteps, x1);
                for(int i = 0; i < C; ++i) cc[(size_t)cix * C + i] = u[i];
                csteps[cix] = nsteps;
                cpar[cix]   = (int32_t)p;
                int a1, a2;
                R.locate(x1, &a1, &a2);
                creg[cix] = a1;
                cr2[cix]  = a2;
                __atomic_fetch_add(&total[a1], 1, __ATOMIC_RELAXED);
                if(fail == 0) __atomic_fetch_add(&valid[a1], 1, __ATOMIC_RELAXED);
                cvalid[cix] = fail == 0 ? 1 : 0;
                cword[cix]  = ds.v[1];
            }
            // Admission.  Sub-region novelty is decided as the reference's
            // kernel decides it when its threads run one at a time in slot
            // order (KPAX.cu:394-395): the first valid child to reach an
            // uncovered R2 cell sees it empty.  The score draw compares with
            // the scores of the previous iteration (updateVertices runs after
            // the expansion, KPAX.cu:96-97).
            for(uint32_t ci = 0; ci < M; ++ci)
            {
                bool fresh = false;
                if(cvalid[ci] && r2_first[cr2[ci]] == std::numeric_limits<int32_t>::max())
                {
                    r2_first[cr2[ci]] = (int32_t)t;
                    covered[creg[ci]]++;
                    fresh = true;
                }
                const bool by_score = cvalid[ci] && admit_by_score(cword[ci], score[creg[ci]]);
                adm[ci]             = (by_score || fresh) ? 1 : 0;
                if(adm[ci] && c.prune)
                {
                    // goal-progress pruning (prune_updateFrontier_kernel, PruneKPAX.cu:472-489);
                    // decided before the append, so the kept children leave no holes
                    const float* x1 = &cs[(size_t)ci * S];
                    const float* x0 = &st[(size_t)cpar[ci] * S];
                    const float cx = x1[0] - goal[0], cy = x1[1] - goal[1], cz = x1[2] - goal[2];
                    const float px = x0[0] - goal[0], py = x0[1] - goal[1], pz = x0[2] - goal[2];
                    const float cd = std::sqrt(cx * cx + cy * cy + cz * cz);
                    const float pd = std::sqrt(px * px + py * py + pz * pz);
                    const float progress = pd - cd;
                    float np = progress / c.prune_progress_scale + 0.5f;
                    np       = np < 0.0f ? 0.0f : (np > 1.0f ? 1.0f : np);
                    float pk = c.prune_exploration + c.prune_goal_bias * np;
                    pk       = pk < c.eps ? c.eps : (pk > 1.0f ? 1.0f : pk);
                    const U4 dp = draw(seed, ci, t, query, 4u);
                    adm[ci]     = u01(dp.v[0]) < pk ? 1 : 0;
                }
            }
            // scores (Graph::updateVertices, Graph.cu:207-227)
            scores(c, R, valid.data(), total.data(), covered.data(), score.data(), &score_total);
            // append admitted children in slot order (updateFrontier_kernel, KPAX.cu:458-502)
            const uint32_t ts_old = ts;
            int32_t goal_idx      = -1;
            for(uint32_t ci = 0; ci < M; ++ci)
            {
                if(!adm[ci]) continue;
                const uint32_t idx = ts++;
                for(int d = 0; d < S; ++d) st[(size_t)idx * S + d] = cs[(size_t)ci * S + d];
                for(int d = 0; d < C; ++d) ct[(size_t)idx * C + d] = cc[(size_t)ci * C + d];
                steps[idx]  = csteps[ci];
                parent[idx] = cpar[ci];
                reg[idx]    = creg[ci];
                if(goal_idx < 0 && goal_distance(&cs[(size_t)ci * S], goal) < c.goal_radius) goal_idx = (int32_t)idx;
            }
            // fAccept of this iteration, from the tree size after the append (KPAX.cu:527-528)
            const float fa = f_accept(c, t, ts);
            uint32_t* tr   = trace ? trace + (size_t)(t - 1) * kTraceWords : nullptr;
            if(tr)
            {
                tr[TR_W]        = (uint32_t)W;
                tr[TR_FAN]      = fan;
                tr[TR_M]        = (uint32_t)M;
                tr[TR_ADMITTED] = ts - ts_old;
                tr[TR_TS]       = ts;
                std::memcpy(&tr[TR_FA_BITS], &fa, 4);
                tr[TR_GOAL]   = (uint32_t)goal_idx;
                tr[TR_NEXT_W] = 0;
            }
            if(std::getenv("KPAX_ORACLE_TRACE"))
                std::fprintf(stderr, "t=%u W=%llu fan=%u M=%llu admitted=%u ts=%u fa=%g\n", t, (unsigned long long)W,
                             fan, (unsigned long long)M, ts - ts_old, ts, fa);
            if(goal_idx >= 0)
            {
                out.status = KPAX_OK;
                out.goal   = goal_idx;
                break;
            }
            // next frontier: re-admitted old nodes and every new node (KPAX.cu:471-517)
            // The list holds at most N entries (no launch reads past M <= N);
            // W counts every entry.
            rep.clear();
            W = 0;
            for(uint32_t j = 0; j < ts; ++j)
            {
                uint32_t w;
                if(j < ts_old)
                {
                    const U4 dr = draw(seed, j, t, query, PURPOSE_READMIT);
                    w           = readmit(dr.v[0], score[reg[j]], fa) ? 1u : 0u;
                }
                else
                    w = new_node_repeat(c, valid[reg[j]]);
                weight[j] = w;
                for(uint32_t k = 0; k < w && rep.size() < N; ++k) rep.push_back(j);
                W += w;
            }
            if(tr) tr[TR_NEXT_W] = (uint32_t)W;
        }
        out.tree_size = ts;
        return out;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// C interface for the Python test harness (ctypes).
// ---------------------------------------------------------------------------
extern "C" {

void oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out)
{
    U4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
    U4 r = philox4x32_10(c, key[0], key[1]);
    for(int i = 0; i < 4; ++i) out[i] = r.v[i];
}

void oracle_sincos_shift(const float* a, const float* d, uint32_t n, float* s, float* c)
{
    for(uint32_t i = 0; i < n; ++i)
    {
        float sa, ca;
        sincos_f32(a[i], &sa, &ca);
        Math<float>::sincos_shift(sa, ca, a[i], d[i], &s[i], &c[i]);
    }
}

void oracle_sincos_f32(const float* x, uint32_t n, float* s, float* c)
{
    for(uint32_t i = 0; i < n; ++i) sincos_f32(x[i], &s[i], &c[i]);
}

// Propagate n edges with injected controls/steps (SoA in and out).
// use_double = 0: float path (parity with the GPU); 1: double path (truth).
// out_x1 is float* or double* accordingly.
int oracle_propagate(const kpax_config* cfg, const float* obs, uint32_t n_obs, const float* x0_soa,
                     const float* u_soa, const int32_t* steps, uint32_t n, int32_t use_double, void* out_x1,
                     uint8_t* valid, int32_t* fail_step)
{
    const Env e{cfg, obs, n_obs};
    const int S = state_dim(cfg->model), C = control_dim(cfg->model);
#pragma omp parallel for schedule(dynamic, 64)
    for(int64_t ii = 0; ii < (int64_t)n; ++ii)
    {
        const uint32_t i = (uint32_t)ii;
        int fail;
        if(use_double)
        {
            double x0[12], u[4], x1[12];
            for(int d = 0; d < S; ++d) x0[d] = x0_soa[(size_t)d * n + i];
            for(int d = 0; d < C; ++d) u[d] = u_soa[(size_t)d * n + i];
            fail = propagate<double>(e, x0, u, steps[i], x1);
            for(int d = 0; d < S; ++d) ((double*)out_x1)[(size_t)d * n + i] = x1[d];
        }
        else
        {
            float x0[12], u[4], x1[12];
            for(int d = 0; d < S; ++d) x0[d] = x0_soa[(size_t)d * n + i];
            for(int d = 0; d < C; ++d) u[d] = u_soa[(size_t)d * n + i];
            fail = propagate<float>(e, x0, u, steps[i], x1);
            for(int d = 0; d < S; ++d) ((float*)out_x1)[(size_t)d * n + i] = x1[d];
        }
        valid[i]     = fail == 0 ? 1 : 0;
        fail_step[i] = fail;
    }
    return 0;
}

// Brute-force verdicts for n segment boxes (SoA 3 x n).
void oracle_collide(const float* obs, uint32_t n_obs, const float* bbmin_soa, const float* bbmax_soa, uint32_t n,
                    uint8_t* free_out)
{
#pragma omp parallel for schedule(static)
    for(int64_t ii = 0; ii < (int64_t)n; ++ii)
    {
        const uint32_t i = (uint32_t)ii;
        float lo[3], hi[3];
        for(int d = 0; d < 3; ++d)
        {
            lo[d] = bbmin_soa[(size_t)d * n + i];
            hi[d] = bbmax_soa[(size_t)d * n + i];
        }
        free_out[i] = segment_free<float>(lo, hi, obs, n_obs) ? 1 : 0;
    }
}

// Verdicts of n segment boxes through the oracle's grid (the planner's
// collision check), for comparison with the brute force above.
void oracle_collide_grid(const kpax_config* cfg, const float* obs, uint32_t n_obs, const float* bbmin_soa,
                         const float* bbmax_soa, uint32_t n, uint8_t* free_out)
{
    const Grid g(*cfg, obs, n_obs);
    for(uint32_t i = 0; i < n; ++i)
    {
        float lo[3], hi[3];
        for(int d = 0; d < 3; ++d)
        {
            lo[d] = bbmin_soa[(size_t)d * n + i];
            hi[d] = bbmax_soa[(size_t)d * n + i];
        }
        free_out[i] = g.free(lo, hi, obs) ? 1 : 0;
    }
}

// Uniform grid restated from spatialHash.cu:9-54 (count per cell; the CSR
// start offsets are the exclusive scan of the counts).  Cells are indexed
// x fastest (gridToIndex, spatialHash.cuh:57-60).  The cell of a coordinate
// is (x - min) times cells per metre, the kernels' form (the reference
// divides by the cell size); any monotone map keeps the verdicts equal to
// brute force.
void oracle_grid_counts(const kpax_config* cfg, const float* obs, uint32_t n_obs, int32_t* counts)
{
    const int* D = cfg->grid_dims;
    float inv[3];  // cells per metre, as the kernels' grid
    for(int d = 0; d < 3; ++d) inv[d] = (float)D[d] / (cfg->ws_max[d] - cfg->ws_min[d]);
    std::fill(counts, counts + (size_t)D[0] * D[1] * D[2], 0);
    for(uint32_t i = 0; i < n_obs; ++i)
    {
        int lo[3], hi[3];
        for(int d = 0; d < 3; ++d)
        {
            lo[d] = clamp_index((obs[6 * i + d] - cfg->ws_min[d]) * inv[d], D[d]);
            hi[d] = clamp_index((obs[6 * i + 3 + d] - cfg->ws_min[d]) * inv[d], D[d]);
        }
        for(int z = lo[2]; z <= hi[2]; ++z)
            for(int y = lo[1]; y <= hi[1]; ++y)
                for(int x = lo[0]; x <= hi[0]; ++x) counts[x + D[0] * (y + D[1] * z)]++;
    }
}

void oracle_regions(const kpax_config* cfg, const float* states_soa, uint32_t n, int32_t* r1, int32_t* r2)
{
    const Regions R(*cfg);
    for(uint32_t i = 0; i < n; ++i)
    {
        const float p[3] = {states_soa[i], states_soa[n + i], states_soa[2 * (size_t)n + i]};
        R.locate(p, &r1[i], &r2[i]);
    }
}

// Counters of n children in one iteration (the atomics of KPAX.cu:386-396):
// accumulates into total/valid_count/covered; r2_first (R1*R2per entries,
// INT32_MAX = never reached) carries coverage between calls.  is_new[i] = 1
// for the valid child that first reaches an uncovered R2 cell in slot order.
void oracle_histogram(const kpax_config* cfg, const float* states_soa, const uint8_t* valid_in, uint32_t n,
                      int32_t iter, int32_t* total, int32_t* valid_count, int32_t* covered, int32_t* r2_first,
                      uint8_t* is_new)
{
    const Regions R(*cfg);
    for(uint32_t i = 0; i < n; ++i)
    {
        const float p[3] = {states_soa[i], states_soa[n + i], states_soa[2 * (size_t)n + i]};
        int a1, a2;
        R.locate(p, &a1, &a2);
        total[a1]++;
        uint8_t nw = 0;
        if(valid_in[i])
        {
            valid_count[a1]++;
            if(r2_first[a2] == std::numeric_limits<int32_t>::max())
            {
                r2_first[a2] = iter;
                covered[a1]++;
                nw = 1;
            }
        }
        if(is_new) is_new[i] = nw;
    }
}

void oracle_scores(const kpax_config* cfg, const int32_t* valid, const int32_t* total, const int32_t* covered,
                   float* out, float* total_out)
{
    const Regions R(*cfg);
    scores(*cfg, R, valid, total, covered, out, total_out);
}

float oracle_f_accept(const kpax_config* cfg, uint32_t iter, uint32_t tree_plus_new)
{
    return f_accept(*cfg, iter, tree_plus_new);
}

// findInd + repeatInd (Planner.cu:52-80): index i repeated weights[i] times.
uint64_t oracle_repeat_expand(const uint32_t* weights, uint32_t n, uint32_t* out, uint64_t cap)
{
    uint64_t k = 0;
    for(uint32_t i = 0; i < n; ++i)
        for(uint32_t j = 0; j < weights[i]; ++j, ++k)
            if(k < cap) out[k] = i;
    return k;
}

// Whole planner.  Outputs are optional (NULL to skip); SoA outputs use
// max_tree as column stride.  Returns the status; *info receives
// {iterations, tree_size, goal_index, expansions_lo, expansions_hi}.
// oracle_plan_ex adds: weights_out (max_tree entries: every node's repeat
// count in the last repeat list built) and trace (8 words per iteration for up
// to max_iter iterations: W, fan, M, admitted, tree size after the append,
// fAccept's float bits, goal index or -1, next W).
// prop_table (8 x max_tree floats, optional) with mix > 0: the learned
// proposal's per-node Gaussians, see Planner::prop_table.
int oracle_plan_ex(const kpax_config* cfg, const float* obs, uint32_t n_obs, const float* start, const float* goal,
                   uint64_t seed, int32_t threads, uint32_t* info, float* states_soa, float* controls_soa,
                   int32_t* steps_out, int32_t* parents_out, int32_t* regions_out, int32_t* valid_out,
                   int32_t* total_out, int32_t* covered_out, float* scores_out, uint32_t* weights_out,
                   uint32_t* trace, const float* prop_table, float mix)
{
#ifdef _OPENMP
    if(threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    Planner P(*cfg, obs, n_obs);
    P.prop_table = prop_table;
    P.mix        = mix;
    const PlanOut o = P.plan(start, goal, seed, 0, trace);
    if(weights_out) std::memcpy(weights_out, P.weight.data(), sizeof(uint32_t) * cfg->max_tree);
    const uint32_t N = cfg->max_tree;
    if(info)
    {
        info[0] = o.iterations;
        info[1] = o.tree_size;
        info[2] = (uint32_t)o.goal;
        info[3] = (uint32_t)o.expansions;
        info[4] = (uint32_t)(o.expansions >> 32);
    }
    for(uint32_t i = 0; i < o.tree_size; ++i)
    {
        if(states_soa)
            for(int d = 0; d < P.S; ++d) states_soa[(size_t)d * N + i] = P.st[(size_t)i * P.S + d];
        if(controls_soa)
            for(int d = 0; d < P.C; ++d) controls_soa[(size_t)d * N + i] = P.ct[(size_t)i * P.C + d];
        if(steps_out) steps_out[i] = P.steps[i];
        if(parents_out) parents_out[i] = P.parent[i];
        if(regions_out) regions_out[i] = P.reg[i];
    }
    const int r1 = P.R.r1_count;
    if(valid_out) std::memcpy(valid_out, P.valid.data(), sizeof(int32_t) * r1);
    if(total_out) std::memcpy(total_out, P.total.data(), sizeof(int32_t) * r1);
    if(covered_out) std::memcpy(covered_out, P.covered.data(), sizeof(int32_t) * r1);
    if(scores_out) std::memcpy(scores_out, P.score.data(), sizeof(float) * r1);
    return o.status;
}

int oracle_plan(const kpax_config* cfg, const float* obs, uint32_t n_obs, const float* start, const float* goal,
                uint64_t seed, int32_t threads, uint32_t* info, float* states_soa, float* controls_soa,
                int32_t* steps_out, int32_t* parents_out, int32_t* regions_out, int32_t* valid_out,
                int32_t* total_out, int32_t* covered_out, float* scores_out)
{
    return oracle_plan_ex(cfg, obs, n_obs, start, goal, seed, threads, info, states_soa, controls_soa, steps_out,
                          parents_out, regions_out, valid_out, total_out, covered_out, scores_out, nullptr, nullptr,
                          nullptr, 0.0f);
}

// The quadcopter's derivative (ode, statePropagator.cu:325-370) at one state
// and control, in the float form the planner integrates (use_float = 1, with
// the project's sin/cos) or in double (0).
void oracle_quad_ode(const kpax_config* cfg, const double* x, const double* u, double* out, int32_t use_float)
{
    if(use_float)
    {
        const QuadK<float> k = quad_constants<float>(*cfg);
        float xf[12], uf[4], o[12], sn[3], cn[3];
        const float zero[12] = {};
        for(int j = 0; j < 12; ++j) xf[j] = (float)x[j];
        for(int j = 0; j < 4; ++j) uf[j] = (float)u[j];
        for(int j = 0; j < 3; ++j) sincos_f32(xf[3 + j], &sn[j], &cn[j]);
        quad_ode<float>(o, xf, zero, 1.0f, uf, k, sn, cn, xf[4]);
        for(int j = 0; j < 12; ++j) out[j] = o[j];
    }
    else
    {
        const QuadK<double> k = quad_constants<double>(*cfg);
        double sn[3], cn[3];
        const double zero[12] = {};
        for(int j = 0; j < 3; ++j) Math<double>::sincos(x[3 + j], &sn[j], &cn[j]);
        quad_ode<double>(out, x, zero, 1.0, u, k, sn, cn, x[4]);
    }
}

// The iteration's decision predicates, for boundary tests.
void oracle_fan_out(const kpax_config* cfg, uint64_t W, uint32_t remaining, uint32_t* fan, uint64_t* M)
{
    fan_out(*cfg, W, remaining, fan, M);
}
int32_t oracle_admit_by_score(uint32_t word, float score) { return admit_by_score(word, score) ? 1 : 0; }
int32_t oracle_readmit(uint32_t word, float score, float fa) { return readmit(word, score, fa) ? 1 : 0; }
uint32_t oracle_new_node_repeat(const kpax_config* cfg, int32_t region_valid)
{
    return new_node_repeat(*cfg, region_valid);
}


void oracle_occupancy(const kpax_config* cfg, const float* obs, uint32_t n_obs, uint32_t* bits)
{
    prop::occupancy(prop::feat_params(*cfg), obs, n_obs, bits);
}

void oracle_mlp_features(const kpax_config* cfg, const float* obs, uint32_t n_obs, const float* states_soa,
                         const float* goals_soa, uint32_t n, float* out)
{
    const prop::Feat F = prop::feat_params(*cfg);
    std::vector<uint32_t> occ(prop::kOcc * prop::kOcc * prop::kOcc / 32);
    prop::occupancy(F, obs, n_obs, occ.data());
    for(uint32_t i = 0; i < n; ++i)
    {
        float x[12] = {}, g[3];
        for(int d = 0; d < F.S; ++d) x[d] = states_soa[(size_t)d * n + i];
        for(int d = 0; d < 3; ++d) g[d] = goals_soa[(size_t)d * n + i];
        prop::features(F, occ.data(), x, g, out + 80 * (size_t)i);
    }
}

void oracle_propose_sample(const kpax_config* cfg, const float* mean_sigma, const uint32_t* words, uint32_t n,
                           float* out)
{
    for(uint32_t i = 0; i < n; ++i)
        prop::propose(mean_sigma + 8 * (size_t)i, mean_sigma + 8 * (size_t)i + 4, words + 4 * (size_t)i, *cfg,
                      out + 4 * (size_t)i);
}

float oracle_ln_f32(float x) { return prop::ln_f32(x); }

}  // extern "C"
