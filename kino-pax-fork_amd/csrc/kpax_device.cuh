// kpax_device.cuh — device building blocks of the MI355X planner:
// counter-based RNG, float sin/cos, dynamics with per-step checks, collision
// through the uniform grid, region location, wave/block scans and the
// decoupled look-back used by every compaction.
//
// Floating-point expressions keep the reference's operation order and the
// library is compiled with -ffp-contract=off, so a child's state, its region
// and its verdict are reproducible bit for bit on the host.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "kpax_internal.h"

namespace kpax {
namespace dev {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Counter = (slot,
// iteration, query, purpose), key = the 64-bit plan seed.  Zero bytes of
// generator state in HBM (the reference keeps a 48-byte curandState per tree
// slot, read and written by every expansion: Planner.cu:18, KPAX.cu:384,398).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                        uint32_t k1)
{
#pragma unroll
    for(int r = 0; r < 10; ++r)
    {
        if(r)
        {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // one 32 x 32 -> 64-bit multiply per product (v_mad_u64_u32) gives
        // both halves, instead of separate high and low multiplies
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    return make_uint4(c0, c1, c2, c3);
}

enum : uint32_t {
    PURPOSE_CONTROL     = 0,
    PURPOSE_STEPS_ADMIT = 1,
    PURPOSE_READMIT     = 2,
    PURPOSE_PROPOSAL    = 3,
    PURPOSE_PRUNE       = 4
};

__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

// sin and cos of x: Cody-Waite reduction by pi/2 (three float parts applied
// with fma), then Taylor polynomials (degree 13 / 12) on [-pi/4, pi/4].
__device__ __forceinline__ void sincos_f32(float x, float* s, float* c)
{
    const float k  = __builtin_rintf(x * 0x1.45f306p-1f);
    const int q    = (int)k;
    float r        = __builtin_fmaf(-k, 0x1.921fb6p+0f, x);
    r              = __builtin_fmaf(-k, -0x1.777a5cp-25f, r);
    r              = __builtin_fmaf(-k, -0x1.ee59dap-50f, r);
    const float r2 = r * r;
    float ps       = __builtin_fmaf(r2, 0x1.612462p-33f, -0x1.ae6456p-26f);
    ps             = __builtin_fmaf(r2, ps, 0x1.71de3ap-19f);
    ps             = __builtin_fmaf(r2, ps, -0x1.a01a02p-13f);
    ps             = __builtin_fmaf(r2, ps, 0x1.111112p-7f);
    ps             = __builtin_fmaf(r2, ps, -0x1.555556p-3f);
    const float sp = __builtin_fmaf(r * r2, ps, r);
    float pc       = __builtin_fmaf(r2, 0x1.1eed8ep-29f, -0x1.27e4fcp-22f);
    pc             = __builtin_fmaf(r2, pc, 0x1.a01a02p-16f);
    pc             = __builtin_fmaf(r2, pc, -0x1.6c16c2p-10f);
    pc             = __builtin_fmaf(r2, pc, 0x1.555556p-5f);
    pc             = __builtin_fmaf(r2, pc, -0.5f);
    const float cp = __builtin_fmaf(r2, pc, 1.0f);
    const int qq   = q & 3;
    const float sa = (qq & 1) ? cp : sp;
    const float ca = (qq & 1) ? sp : cp;
    *s             = (qq & 2) ? -sa : sa;
    *c             = ((qq + 1) & 2) ? -ca : ca;
}

// Truncation toward zero clamped to [0, len-1], NaN -> 0; identical to the
// reference's (int) cast + clamp for finite inputs (Graph.cu:107-109).
// Written without branches: max(f, 0) is 0 for NaN, negatives and -0, and
// min(., len - 1) truncates to len - 1 for every f >= len - 1, so this is
// `f >= 0 ? (f >= len ? len - 1 : (int)f) : 0` for every float f.
__host__ __device__ __forceinline__ int clamp_index(float f, int len)
{
    return (int)fminf(fmaxf(f, 0.0f), (float)(len - 1));
}

__device__ __forceinline__ bool inb(float v, float lo, float hi) { return v >= lo && v <= hi; }

// ---------------------------------------------------------------------------
// Uniform grid in CSR form: start[G+1], items[K] (obstacle ids ascending per
// cell), obstacle boxes as two float4 arrays.  Built once per environment.
// ---------------------------------------------------------------------------
// The CSR items hold the obstacles' boxes themselves (two float4 per item,
// copied from the obstacle list after the build), so testing an item is one
// pair of independent loads instead of an index load and then the box.
// An environment's collision grid in a form a work-group stages into LDS:
// a coarse uniform grid of the workspace (dimensions D, the same kind of
// monotone cell map as the CSR grid) whose cells list 16-bit obstacle ids,
// and every obstacle's box.  Image layout in bytes: box i at 24 i as
// (min x, min y), (min z, max x), (max y, max z); cell starts (uint16, one per
// cell plus the end) at off_start; ids (uint16) at off_items; `bytes` a
// multiple of 16.  bytes = 0: the environment has no image (too large for
// the LDS budget) and queries use the CSR grid in global memory.
struct LdsGrid {
    int32_t D[3];
    float mn[3], inv[3];
    uint32_t off_start, off_items, bytes;
    const uint4* img;  // the image in global memory
};

struct GridView {
    GridParams gp;
    const int32_t* start;
    const float4* box;
    const uint32_t* occ;  // occupancy bitmap (learned proposal's features)
    LdsGrid lg;
    const uint8_t* lds;  // the staged image (set by stage_lds_grid inside a kernel), else nullptr
};

// Cell of coordinate x: a multiply by cells per metre instead of a division
// by the cell size.  The grid build and every query use this same monotone
// map, so a box and every obstacle it overlaps share a cell whatever the
// rounding.
__host__ __device__ __forceinline__ int cell_of(float x, float mn, float inv, int D)
{
    return clamp_index((x - mn) * inv, D);
}

// Segment box vs all obstacles in the cells it overlaps
// (isMotionValidSpatialHash, spatialHash.cu:145-194; broad phase
// collisionCheck.cu:3-10).  Any obstacle that overlaps the box shares at
// least one cell with it, so the verdict equals the brute-force one.
__device__ __forceinline__ bool box_hits(const GridView& g, int k, const float bmin[3], const float bmax[3])
{
    const float4 a = g.box[2 * (size_t)k];
    const float4 b = g.box[2 * (size_t)k + 1];
    const bool sep = bmax[0] <= a.x || b.x <= bmin[0] || bmax[1] <= a.y || b.y <= bmin[1] || bmax[2] <= a.z ||
                     b.z <= bmin[2];
    return !sep;
}

// The same query through the staged LDS image: every obstacle listed in the
// coarse cells the box overlaps is tested (isMotionValidSpatialHash,
// spatialHash.cu:145-194; broad phase collisionCheck.cu:3-10).  An obstacle
// that overlaps the box shares a coarse cell with it, so the verdict is the
// brute-force one, as with the CSR grid.
__device__ __forceinline__ bool lds_walk(const GridView& g, const int lo[3], const int hi[3], const float bmin[3],
                                         const float bmax[3])
{
    const LdsGrid& L    = g.lg;
    const uint16_t* st = (const uint16_t*)(g.lds + L.off_start);
    const uint16_t* it = (const uint16_t*)(g.lds + L.off_items);
    const float2* bx   = (const float2*)g.lds;
    for(int z = lo[2]; z <= hi[2]; ++z)
        for(int y = lo[1]; y <= hi[1]; ++y)
        {
            // the cells of one x-row are contiguous
            const int c0 = lo[0] + L.D[0] * (y + L.D[1] * z);
            const int e  = st[c0 + 1 + hi[0] - lo[0]];
            for(int k = st[c0]; k < e; ++k)
            {
                const int o    = it[k];
                const float2 a = bx[3 * o], b = bx[3 * o + 1], c = bx[3 * o + 2];
                const bool sep = bmax[0] <= a.x || b.y <= bmin[0] || bmax[1] <= a.y || c.x <= bmin[1] ||
                                 bmax[2] <= b.x || c.y <= bmin[2];
                if(!sep) return false;
            }
        }
    return true;
}

__device__ __forceinline__ bool seg_free_lds(const GridView& g, const float bmin[3], const float bmax[3])
{
    const LdsGrid& L = g.lg;
    int lo[3], hi[3];
#pragma unroll
    for(int d = 0; d < 3; ++d)
    {
        lo[d] = cell_of(bmin[d], L.mn[d], L.inv[d], L.D[d]);
        hi[d] = cell_of(bmax[d], L.mn[d], L.inv[d], L.D[d]);
    }
    return lds_walk(g, lo, hi, bmin, bmax);
}

// Image cells of a point (the first step of a child carries them from its
// start point, every later step from the previous step's end).
__device__ __forceinline__ void lds_cells(const GridView& g, const float* p, int c[3])
{
#pragma unroll
    for(int d = 0; d < 3; ++d) c[d] = cell_of(p[d], g.lg.mn[d], g.lg.inv[d], g.lg.D[d]);
}

// Copies g's LDS image into `lds` with every thread of the work-group and
// points g at it (no-op without an image).  A work-group barrier: call it
// with all threads, before any of them returns.
__device__ __forceinline__ void stage_lds_grid(GridView& g, uint4* lds)
{
    g.lds = nullptr;
    const uint32_t n = g.lg.bytes / 16;
    if(n == 0) return;
    for(uint32_t i = threadIdx.x; i < n; i += blockDim.x) lds[i] = g.lg.img[i];
    __syncthreads();
    g.lds = (const uint8_t*)lds;
}

__device__ __forceinline__ bool csr_walk(const GridView& g, const int lo[3], const int hi[3], const float bmin[3],
                                         const float bmax[3])
{
    if(hi[0] - lo[0] <= 1 && hi[1] - lo[1] <= 1 && hi[2] - lo[2] <= 1)
    {
        // An integration step spans at most 2 cells per axis: the cell-start
        // words of its (up to) four x-rows are loaded together, so the lookup
        // costs one memory latency instead of one per cell.
        const int nx = hi[0] - lo[0];
        int b[4], e[4];
#pragma unroll
        for(int r = 0; r < 4; ++r)
        {
            const int y = lo[1] + (r & 1), z = lo[2] + (r >> 1);
            const bool in = y <= hi[1] && z <= hi[2];
            const int c   = lo[0] + g.gp.D[0] * ((in ? y : lo[1]) + g.gp.D[1] * (in ? z : lo[2]));
            b[r]          = in ? g.start[c] : 0;
            e[r]          = in ? g.start[c + 1 + nx] : 0;
        }
#pragma unroll
        for(int r = 0; r < 4; ++r)
            for(int k = b[r]; k < e[r]; ++k)
                if(box_hits(g, k, bmin, bmax)) return false;
        return true;
    }
    for(int z = lo[2]; z <= hi[2]; ++z)
        for(int y = lo[1]; y <= hi[1]; ++y)
        {
            // the cells of one x-row are contiguous in the CSR
            const int c0 = lo[0] + g.gp.D[0] * (y + g.gp.D[1] * z);
            const int e  = g.start[c0 + 1 + hi[0] - lo[0]];
            for(int k = g.start[c0]; k < e; ++k)
                if(box_hits(g, k, bmin, bmax)) return false;
        }
    return true;
}

__device__ __forceinline__ bool seg_free_csr(const GridView& g, const float bmin[3], const float bmax[3])
{
    int lo[3], hi[3];
#pragma unroll
    for(int d = 0; d < 3; ++d)
    {
        lo[d] = cell_of(bmin[d], g.gp.mn[d], g.gp.inv[d], g.gp.D[d]);
        hi[d] = cell_of(bmax[d], g.gp.mn[d], g.gp.inv[d], g.gp.D[d]);
    }
    return csr_walk(g, lo, hi, bmin, bmax);
}

// CSR grid cells of a point (carried from step to step like lds_cells).
__device__ __forceinline__ void csr_cells(const GridView& g, const float* p, int c[3])
{
#pragma unroll
    for(int d = 0; d < 3; ++d) c[d] = cell_of(p[d], g.gp.mn[d], g.gp.inv[d], g.gp.D[d]);
}

// Grid query path, fixed at compile time where a kernel knows it (so the
// other path's code and registers stay out of it): kGridAny takes the staged
// LDS image when there is one, else the CSR grid.
enum : int { kGridAny = 0, kGridLds = 1, kGridCsr = 2 };

template<int GQ = kGridAny>
__device__ __forceinline__ bool seg_free(const GridView& g, const float bmin[3], const float bmax[3])
{
    if constexpr(GQ == kGridLds)
        return seg_free_lds(g, bmin, bmax);
    else if constexpr(GQ == kGridCsr)
        return seg_free_csr(g, bmin, bmax);
    else
        return g.lds ? seg_free_lds(g, bmin, bmax) : seg_free_csr(g, bmin, bmax);
}

template<int GQ = kGridAny>
__device__ __forceinline__ bool ws_and_collision(const ModelParams& P, const GridView& G, const float p0[3],
                                                 const float p1[3])
{
#pragma unroll
    for(int d = 0; d < 3; ++d)
        if(!inb(p1[d], P.ws_min[d], P.ws_max[d])) return false;
    float bmin[3], bmax[3];
#pragma unroll
    for(int d = 0; d < 3; ++d)
    {
        const bool sw = p0[d] > p1[d];
        bmin[d]       = sw ? p1[d] : p0[d];
        bmax[d]       = sw ? p0[d] : p1[d];
    }
    return seg_free<GQ>(G, bmin, bmax);
}

// ---------------------------------------------------------------------------
// Dynamics.  Each propagator returns the 1-based failing step (0 = valid) and
// leaves in x1 the state after the last step taken, as the reference does.
// ---------------------------------------------------------------------------
template<int MODEL> struct Dims;
template<> struct Dims<KPAX_MODEL_DOUBLE_INTEGRATOR> {
    static constexpr int S = 6, C = 3;
};
template<> struct Dims<KPAX_MODEL_DUBINS_AIRPLANE> {
    static constexpr int S = 6, C = 3;
};
template<> struct Dims<KPAX_MODEL_QUADCOPTER> {
    static constexpr int S = 12, C = 4;
};

// One integration step of each model on the state x (in place) with the
// controls u, followed by the reference's checks; returns false when the step
// is invalid (the state then holds the failing step's result, as the
// reference leaves it).  The expression text is the reference's, so the
// float result does not depend on how steps are scheduled.

// Double integrator: statePropagatorSpatialHash.cu:90-128.
// integrate_*: the model's step and its own bound checks (velocity, pitch);
// step_* adds the workspace and collision checks of the step's segment.
__device__ __forceinline__ bool integrate_di(const ModelParams& P, float* s, const float* u)
{
    const float dt = P.dt;
    float x = s[0], y = s[1], z = s[2], vx = s[3], vy = s[4], vz = s[5];
    const float ax = u[0], ay = u[1], az = u[2];
    if(!P.di_textbook)
    {
        x += (vx + (vx + 2.0f * (vx + ax * dt / 2.0f) + (vx + ax * dt))) * dt / 6.0f;
        y += (vy + (vy + 2.0f * (vy + ay * dt / 2.0f) + (vy + ay * dt))) * dt / 6.0f;
        z += (vz + (vz + 2.0f * (vz + az * dt / 2.0f) + (vz + az * dt))) * dt / 6.0f;
        vx += (ax + 2.0f * ax + 2.0f * ax + ax) * dt / 6.0f;
        vy += (ay + 2.0f * ay + 2.0f * ay + ay) * dt / 6.0f;
        vz += (az + 2.0f * az + 2.0f * az + az) * dt / 6.0f;
    }
    else
    {
        x += vx * dt + 0.5f * ax * dt * dt;
        y += vy * dt + 0.5f * ay * dt * dt;
        z += vz * dt + 0.5f * az * dt * dt;
        vx += ax * dt;
        vy += ay * dt;
        vz += az * dt;
    }
    s[0] = x, s[1] = y, s[2] = z, s[3] = vx, s[4] = vy, s[5] = vz;
    return inb(vx, P.vmin, P.vmax) && inb(vy, P.vmin, P.vmax) && inb(vz, P.vmin, P.vmax);
}

template<int GQ = kGridAny>
__device__ __forceinline__ bool step_di(const ModelParams& P, const GridView& G, float* s, const float* u)
{
    const float p0[3] = {s[0], s[1], s[2]};
    if(!integrate_di(P, s, u)) return false;
    return ws_and_collision<GQ>(P, G, p0, s);
}

// Dubins airplane: statePropagatorSpatialHash.cu:158-214.
__device__ __forceinline__ bool integrate_dubins(const ModelParams& P, float* s, const float* u)
{
    const float dt = P.dt, h = P.half, d6 = P.d6;
    float x = s[0], y = s[1], z = s[2], yaw = s[3], pitch = s[4], v = s[5];
    const float a = u[0], yr = u[1], pr = u[2];
    float cp, sp, cy, sy, cpm, spm, cym, sym, cpe, spe, cye, sye;
    sincos_f32(pitch, &sp, &cp);
    sincos_f32(yaw, &sy, &cy);
    sincos_f32(pitch + h * pr, &spm, &cpm);
    sincos_f32(yaw + h * yr, &sym, &cym);
    sincos_f32(pitch + dt * pr, &spe, &cpe);
    sincos_f32(yaw + dt * yr, &sye, &cye);
    const float A  = v + h * a;
    const float B  = v + dt * a;
    const float Tx = A * cpm * cym;
    const float Ty = A * cpm * sym;
    const float Tz = A * spm;
    x += d6 * (v * cp * cy + 2.0f * (Tx + Tx) + B * cpe * cye);
    y += d6 * (v * cp * sy + 2.0f * (Ty + Ty) + B * cpe * sye);
    z += d6 * (v * sp + 2.0f * (Tz + Tz) + B * spe);
    yaw += dt * yr;
    pitch += dt * pr;
    v += d6 * (a + 2.0f * (a + a) + a);
    s[0] = x, s[1] = y, s[2] = z, s[3] = yaw, s[4] = pitch, s[5] = v;
    return inb(v, P.vmin, P.vmax) && inb(pitch, P.pmin, P.pmax);
}

template<int GQ = kGridAny>
__device__ __forceinline__ bool step_dubins(const ModelParams& P, const GridView& G, float* s, const float* u)
{
    const float p0[3] = {s[0], s[1], s[2]};
    if(!integrate_dubins(P, s, u)) return false;
    return ws_and_collision<GQ>(P, G, p0, s);
}

// 12D quadcopter ode: statePropagator.cu:325-370.  The reference's terms,
// each sum of products evaluated with fused multiply-adds (one rounding per
// fma; the host restatement calls std::fma on the same operands in the same
// nesting, so the two agree bit for bit).
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// sin and cos of a + d from those of a: the angle-sum identities, with sin d
// and cos d from their Taylor series through d^11 and d^10 when |d| <= 1
// (truncation below 3e-9), else sincos_f32 of the rounded sum a + d.  The
// RK4 stages move the Euler angles by d = delta * rate, small except near
// gimbal lock, so the stages after the first need no range reduction.
__device__ __forceinline__ void sincos_shift(float sa, float ca, float a, float d, float* s, float* c)
{
    if(__builtin_fabsf(d) <= 1.0f)
    {
        const float d2 = d * d;
        const float ps =
          fma_(d2, fma_(d2, fma_(d2, fma_(d2, -2.5052108e-8f, 2.7557319e-6f), -1.9841270e-4f), 8.3333333e-3f),
               -1.6666667e-1f);
        const float pc =
          fma_(d2, fma_(d2, fma_(d2, fma_(d2, -2.7557319e-7f, 2.4801587e-5f), -1.3888889e-3f), 4.1666667e-2f), -0.5f);
        const float sd = fma_(d * d2, ps, d);
        const float cd = fma_(d2, pc, 1.0f);
        *s             = fma_(sa, cd, ca * sd);
        *c             = fma_(ca, cd, -(sa * sd));
    }
    else
        sincos_f32(a + d, s, c);
}

// sn/cn: sin and cos of the stage's roll, pitch and yaw.
__device__ __forceinline__ void quad_ode(float* out, const float* x, const float* hh, float delta, const float* u,
                                         const ModelParams& k, const float* sn, const float* cn)
{
    const float uu = fma_(delta, hh[6], x[6]), v = fma_(delta, hh[7], x[7]), w = fma_(delta, hh[8], x[8]);
    const float p = fma_(delta, hh[9], x[9]), q = fma_(delta, hh[10], x[10]), r = fma_(delta, hh[11], x[11]);
    const float sf = sn[0], cf = cn[0], st = sn[1], ct = cn[1], ss = sn[2], cs = cn[2];
    const float sec  = 1.0f / ct;  // one division for tan theta and psi-dot
    const float tt   = st * sec;
    const float sfst = sf * st, cfst = cf * st;
    out[0]           = fma_(ct * cs, uu, fma_(fma_(sfst, cs, -(cf * ss)), v, fma_(cfst, cs, sf * ss) * w));
    out[1]           = fma_(ct * ss, uu, fma_(fma_(sfst, ss, cf * cs), v, fma_(cfst, ss, -(sf * cs)) * w));
    out[2]           = fma_(-st, uu, fma_(sf * ct, v, cf * ct * w));
    const float qr   = fma_(q, sf, r * cf);
    out[3]           = fma_(qr, tt, p);
    out[4]           = fma_(q, cf, -(r * sf));
    out[5]           = qr * sec;
    const float mX   = k.mi * (k.negnu * __builtin_sqrtf(fma_(uu, uu, fma_(v, v, w * w))));
    const float gct  = k.g * ct;
    out[6]           = fma_(mX, uu, fma_(-k.g, st, fma_(r, v, -(q * w))));
    out[7]           = fma_(mX, v, fma_(gct, sf, fma_(p, w, -(r * uu))));
    out[8]           = fma_(k.mi, u[0], fma_(mX, w, fma_(gct, cf, fma_(q, uu, -(p * v)))));
    const float LMN  = k.negmu * __builtin_sqrtf(fma_(p, p, fma_(q, q, r * r)));
    out[9]           = fma_(k.ixi, u[1], fma_(k.ixi * LMN, p, k.k1x * q * r));
    out[10]          = fma_(k.iyi, u[2], fma_(k.iyi * LMN, q, k.k1y * p * r));
    out[11]          = fma_(k.izi, u[3], fma_(k.izi * LMN, r, k.k1z * p * q));
}

// Quadcopter RK4: statePropagatorSpatialHash.cu:256-296.  The stage sum
// k1 + 2 k2 + 2 k3 + k4 is accumulated as each stage finishes (the same
// operations in the same order), so only two stage vectors are live.
__device__ __forceinline__ bool integrate_quad(const ModelParams& P, float* h0, const float* u)
{
    const float dt = P.dt, half = P.half, d6 = P.d6;
    float ka[12], kb[12], acc[12];
    const float zero[12] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float sb[3], cb[3], sn[3], cn[3];
#pragma unroll
    for(int i = 0; i < 3; ++i) sincos_f32(h0[3 + i], &sb[i], &cb[i]);
    quad_ode(ka, h0, zero, 1.0f, u, P, sb, cb);  // k1
#pragma unroll
    for(int i = 0; i < 3; ++i) sincos_shift(sb[i], cb[i], h0[3 + i], half * ka[3 + i], &sn[i], &cn[i]);
    quad_ode(kb, h0, ka, half, u, P, sn, cn);  // k2
#pragma unroll
    for(int j = 0; j < 12; ++j) acc[j] = fma_(2.0f, kb[j], ka[j]);
#pragma unroll
    for(int i = 0; i < 3; ++i) sincos_shift(sb[i], cb[i], h0[3 + i], half * kb[3 + i], &sn[i], &cn[i]);
    quad_ode(ka, h0, kb, half, u, P, sn, cn);  // k3
#pragma unroll
    for(int j = 0; j < 12; ++j) acc[j] = fma_(2.0f, ka[j], acc[j]);
#pragma unroll
    for(int i = 0; i < 3; ++i) sincos_shift(sb[i], cb[i], h0[3 + i], dt * ka[3 + i], &sn[i], &cn[i]);
    quad_ode(kb, h0, ka, dt, u, P, sn, cn);  // k4
#pragma unroll
    for(int j = 0; j < 12; ++j) h0[j] = fma_(d6, acc[j] + kb[j], h0[j]);
    return inb(h0[6], P.vmin, P.vmax) && inb(h0[7], P.vmin, P.vmax) && inb(h0[8], P.vmin, P.vmax);
}

template<int GQ = kGridAny>
__device__ __forceinline__ bool step_quad(const ModelParams& P, const GridView& G, float* h0, const float* u)
{
    const float p0[3] = {h0[0], h0[1], h0[2]};
    if(!integrate_quad(P, h0, u)) return false;
    return ws_and_collision<GQ>(P, G, p0, h0);
}

template<int MODEL>
__device__ __forceinline__ bool integrate(const ModelParams& P, float* s, const float* u)
{
    if constexpr(MODEL == KPAX_MODEL_DOUBLE_INTEGRATOR)
        return integrate_di(P, s, u);
    else if constexpr(MODEL == KPAX_MODEL_DUBINS_AIRPLANE)
        return integrate_dubins(P, s, u);
    else
        return integrate_quad(P, s, u);
}

// A step with the grid cells of its start point known (cells, carried from
// the previous step): the cells of the segment's box are the minimum and
// maximum of its two ends' cells (the cell map is monotone), so only the end
// point's are computed, and handed back for the next step.  GQ: the staged
// LDS image's cells or the CSR grid's.  The same verdict and state as
// step<MODEL, GQ>.
template<int MODEL, int GQ>
__device__ __forceinline__ bool step_carry(const ModelParams& P, const GridView& G, float* s, const float* u,
                                           int cells[3])
{
    static_assert(GQ == kGridLds || GQ == kGridCsr, "a carried step needs a fixed grid path");
    const float p0[3] = {s[0], s[1], s[2]};
    if(!integrate<MODEL>(P, s, u)) return false;
#pragma unroll
    for(int d = 0; d < 3; ++d)
        if(!inb(s[d], P.ws_min[d], P.ws_max[d])) return false;
    float bmin[3], bmax[3];
    int lo[3], hi[3];
#pragma unroll
    for(int d = 0; d < 3; ++d)
    {
        const bool sw = p0[d] > s[d];
        bmin[d]       = sw ? s[d] : p0[d];
        bmax[d]       = sw ? p0[d] : s[d];
        const int c1  = GQ == kGridLds ? cell_of(s[d], G.lg.mn[d], G.lg.inv[d], G.lg.D[d])
                                       : cell_of(s[d], G.gp.mn[d], G.gp.inv[d], G.gp.D[d]);
        lo[d]         = cells[d] < c1 ? cells[d] : c1;
        hi[d]         = cells[d] < c1 ? c1 : cells[d];
        cells[d]      = c1;
    }
    if constexpr(GQ == kGridLds)
        return lds_walk(G, lo, hi, bmin, bmax);
    else
        return csr_walk(G, lo, hi, bmin, bmax);
}

template<int MODEL, int GQ = kGridAny>
__device__ __forceinline__ bool step(const ModelParams& P, const GridView& G, float* s, const float* u)
{
    if constexpr(MODEL == KPAX_MODEL_DOUBLE_INTEGRATOR)
        return step_di<GQ>(P, G, s, u);
    else if constexpr(MODEL == KPAX_MODEL_DUBINS_AIRPLANE)
        return step_dubins<GQ>(P, G, s, u);
    else
        return step_quad<GQ>(P, G, s, u);
}

// A whole edge: `steps` steps, stopping at the first invalid one.  Returns the
// 1-based failing step or 0; x1 holds the last state computed.
template<int MODEL>
__device__ __forceinline__ int propagate(const ModelParams& P, const GridView& G, const float* x0, const float* u,
                                         int steps, float* x1)
{
    constexpr int S = Dims<MODEL>::S;
#pragma unroll
    for(int d = 0; d < S; ++d) x1[d] = x0[d];
    for(int i = 0; i < steps; ++i)
        if(!step<MODEL>(P, G, x1, u)) return i + 1;
    return 0;
}

// ---------------------------------------------------------------------------
// Regions: getRegion / getSubRegion (Graph.cu:99-205) on the workspace axes,
// x most significant, with the R1 min corner min + idx * size
// (initializeRegions_kernel, Graph.cu:67-70) computed inline.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void locate(const RegionParams& R, const float* p, int* r1, int* r2)
{
    int idx[3], jdx[3];
#pragma unroll
    for(int d = 0; d < 3; ++d) idx[d] = clamp_index((float)R.L[d] * (p[d] - R.mn[d]) / R.ext[d], R.L[d]);
#pragma unroll
    for(int d = 0; d < 3; ++d)
    {
        const float lo = R.mn[d] + (float)idx[d] * R.size[d];
        jdx[d]         = clamp_index((float)R.L2[d] * (p[d] - lo) / R.size[d], R.L2[d]);
    }
    *r1 = (idx[0] * R.L[1] + idx[1]) * R.L[2] + idx[2];
    *r2 = *r1 * R.r2_per + (jdx[0] * R.L2[1] + jdx[1]) * R.L2[2] + jdx[2];
}

// Sub-region novelty (KPAX.cu:394-395, a read-then-exchange race) decided as
// the reference's kernel decides it when its children run one at a time in
// slot order: only the smallest slot among the valid children reaching an
// uncovered R2 cell sees it empty.  Each cell keeps
// (first iteration << 32 | smallest slot), lowered with a fire-and-forget
// atomicMin; a cell stamped with an earlier iteration is settled, so most
// children decide with one plain read.  A candidate wins when, after the
// launch, the cell's word equals its own key.
constexpr uint32_t kValid = 1u, kCandidate = 2u, kScoreAdmit = 4u;

__device__ __forceinline__ uint64_t cell_key(uint32_t t, uint32_t slot) { return ((uint64_t)t << 32) | slot; }

__device__ __forceinline__ uint32_t claim_cell(uint64_t* r2_state, int r2, uint32_t t, uint32_t slot)
{
    const uint64_t seen = __hip_atomic_load(&r2_state[r2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if((uint32_t)(seen >> 32) < t) return 0u;
    (void)atomicMin((unsigned long long*)&r2_state[r2], (unsigned long long)cell_key(t, slot));
    return kCandidate;
}

// ---------------------------------------------------------------------------
// Scans.  256-thread blocks (4 waves of 64) scan 1024-item tiles; tiles are
// handed out in order by an atomic ticket and chained with a decoupled
// look-back.  A status word holds the flag in bits 62-63 (1 = aggregate,
// 2 = inclusive prefix), the launch's epoch in bits 32-61 and the value in
// bits 0-31; a word whose epoch differs from the launch's is treated as not
// yet written, so the status arrays never need clearing between launches.  A
// tile only waits on tiles with smaller tickets, which are already resident,
// so every wait ends.
// ---------------------------------------------------------------------------
constexpr uint64_t kFlagAgg    = 1ull << 62;
constexpr uint64_t kFlagPrefix = 2ull << 62;
constexpr uint32_t kEpochMask  = (1u << 30) - 1u;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for(int off = 1; off < 64; off <<= 1)
    {
        const uint32_t y = __shfl_up(x, off, 64);
        if(lane >= off) x += y;
    }
    return x;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x)
{
#pragma unroll
    for(int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

__device__ __forceinline__ float wave_sum_f(float x)
{
#pragma unroll
    for(int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// Exclusive scan of one value per thread over a block of NW waves; *agg gets
// the block total.  lds must hold NW words; contains two barriers.
template<int NW>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* agg, uint32_t* lds)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if(lane == 63) lds[wid] = inc;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for(int w = 0; w < NW; ++w)
    {
        const uint32_t s = lds[w];
        if(w < wid) before += s;
        total += s;
    }
    *agg = total;
    __syncthreads();
    return before + inc - v;
}

// block_excl_scan of v together with the block total of w, in the same two
// barriers; lds must hold 2 NW words.
template<int NW>
__device__ __forceinline__ uint32_t block_excl_scan_sum(uint32_t v, uint32_t w, uint32_t* agg, uint32_t* wtot,
                                                        uint32_t* lds)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    const uint32_t ws  = wave_sum(w);
    if(lane == 63) lds[wid] = inc;
    if(lane == 0) lds[NW + wid] = ws;
    __syncthreads();
    uint32_t before = 0, total = 0, wt = 0;
#pragma unroll
    for(int k = 0; k < NW; ++k)
    {
        const uint32_t s = lds[k];
        if(k < wid) before += s;
        total += s;
        wt += lds[NW + k];
    }
    *agg  = total;
    *wtot = wt;
    __syncthreads();
    return before + inc - v;
}

__device__ __forceinline__ uint32_t block_excl_scan_256(uint32_t v, uint32_t* agg, uint32_t* lds)
{
    return block_excl_scan<4>(v, agg, lds);
}

__device__ __forceinline__ void status_store(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t status_load(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t status_word(uint64_t flag, uint32_t epoch, uint32_t value)
{
    return flag | ((uint64_t)(epoch & kEpochMask) << 32) | value;
}

// Called by all 64 lanes of wave 0; returns the exclusive prefix of `tile`
// within its segment, whose first tile is `first` (the tiles of one query
// form a segment; tiles before `first` belong to other queries and are not
// read).
__device__ __forceinline__ uint32_t lookback(uint64_t* status, uint32_t tile, uint32_t agg, uint32_t first,
                                             uint32_t epoch)
{
    const int lane = threadIdx.x & 63;
    if(tile == first)
    {
        if(lane == 0) status_store(&status[tile], status_word(kFlagPrefix, epoch, agg));
        return 0;
    }
    if(lane == 0) status_store(&status[tile], status_word(kFlagAgg, epoch, agg));
    const uint32_t ep = epoch & kEpochMask;
    uint32_t excl     = 0;
    int64_t base      = (int64_t)tile - 1;
    while(true)
    {
        const int64_t j = base - lane;
        uint64_t s      = kFlagPrefix;  // before the segment: an empty prefix
        if(j >= (int64_t)first)
        {
            s = status_load(&status[j]);
            while((s >> 62) == 0 || (uint32_t)((s >> 32) & kEpochMask) != ep) s = status_load(&status[j]);
        }
        const uint64_t pm = __ballot((s >> 62) == 2);
        const int stop    = pm ? __builtin_ctzll(pm) : 64;
        excl += wave_sum(lane <= stop ? (uint32_t)s : 0u);
        if(pm) break;
        base -= 64;
    }
    if(lane == 0) status_store(&status[tile], status_word(kFlagPrefix, epoch, excl + agg));
    return excl;
}

// The look-back in two halves, so a block can work between publishing its
// aggregate and waiting on its predecessors: tile_publish (one thread) makes
// the tile's aggregate visible (or its prefix, for a segment's first tile);
// tile_prefix_published (block-wide, two barriers) then returns the tile's
// exclusive prefix and publishes its inclusive one.
__device__ __forceinline__ void tile_publish(uint64_t* status, uint32_t tile, uint32_t agg, uint32_t first,
                                             uint32_t epoch)
{
    status_store(&status[tile], status_word(tile == first ? kFlagPrefix : kFlagAgg, epoch, agg));
}

__device__ __forceinline__ uint32_t tile_prefix_published(uint64_t* status, uint32_t tile, uint32_t agg,
                                                          uint32_t first, uint32_t epoch, uint32_t* lds_word)
{
    if(threadIdx.x < 64)
    {
        const int lane = threadIdx.x & 63;
        uint32_t excl  = 0;
        if(tile != first)
        {
            const uint32_t ep = epoch & kEpochMask;
            int64_t base      = (int64_t)tile - 1;
            while(true)
            {
                const int64_t j = base - lane;
                uint64_t s      = kFlagPrefix;  // before the segment: an empty prefix
                if(j >= (int64_t)first)
                {
                    s = status_load(&status[j]);
                    while((s >> 62) == 0 || (uint32_t)((s >> 32) & kEpochMask) != ep) s = status_load(&status[j]);
                }
                const uint64_t pm = __ballot((s >> 62) == 2);
                const int stop    = pm ? __builtin_ctzll(pm) : 64;
                excl += wave_sum(lane <= stop ? (uint32_t)s : 0u);
                if(pm) break;
                base -= 64;
            }
            if(lane == 0) status_store(&status[tile], status_word(kFlagPrefix, epoch, excl + agg));
        }
        if(threadIdx.x == 0) *lds_word = excl;
    }
    __syncthreads();
    const uint32_t e = *lds_word;
    __syncthreads();
    return e;
}

// Block-wide: exclusive prefix of this tile given its aggregate.
__device__ __forceinline__ uint32_t tile_prefix(uint64_t* status, uint32_t tile, uint32_t agg, uint32_t first,
                                                uint32_t epoch, uint32_t* lds_word)
{
    if(threadIdx.x < 64)
    {
        const uint32_t e = lookback(status, tile, agg, first, epoch);
        if(threadIdx.x == 0) *lds_word = e;
    }
    __syncthreads();
    const uint32_t e = *lds_word;
    __syncthreads();
    return e;
}

// Next dynamic tile for the whole block (one atomic per block).
__device__ __forceinline__ uint32_t next_tile(uint32_t* counter, uint32_t* lds_word)
{
    if(threadIdx.x == 0) *lds_word = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t t = *lds_word;
    __syncthreads();
    return t;
}

// fAccept of updateFrontier (KPAX.cu:527-528): float iteration * eps times the
// double fifth power of the remaining tree fraction (cap = the query's tree cap).
__device__ __forceinline__ float f_accept(uint32_t t, uint32_t tree_plus_new, uint32_t cap, float eps)
{
    const float tas = 1.0f - ((float)tree_plus_new / (float)cap);
    const double x  = (double)tas;
    const double p5 = x * x * x * x * x;
    return (float)((double)((float)t * eps) * p5);
}

// Goal-progress pruning (prune_updateFrontier_kernel, PruneKPAX.cu:472-489):
// keep probability from the parent's and the child's goal distances (xyz,
// helper.cuh:138-152), compared with the child's own uniform draw.
__device__ __forceinline__ bool prune_keep(const PlanParams& PP, const float* child, const float* parent,
                                           const float* goal, uint32_t draw)
{
    const float cx = child[0] - goal[0], cy = child[1] - goal[1], cz = child[2] - goal[2];
    const float px = parent[0] - goal[0], py = parent[1] - goal[1], pz = parent[2] - goal[2];
    const float cd       = __builtin_sqrtf(cx * cx + cy * cy + cz * cz);
    const float pd       = __builtin_sqrtf(px * px + py * py + pz * pz);
    const float progress = pd - cd;
    float np             = progress / PP.prune_scale + 0.5f;
    np                   = np < 0.0f ? 0.0f : (np > 1.0f ? 1.0f : np);
    float p              = PP.prune_expl + PP.prune_goal * np;
    p                    = p < PP.eps ? PP.eps : (p > 1.0f ? 1.0f : p);
    return u01(draw) < p;
}

// Integration steps of an edge from its Philox word: steps_min + floor(u *
// span), u in [0, 1).
__device__ __forceinline__ int draw_steps(const ModelParams& P, uint32_t word)
{
    int kk = (int)(u01(word) * (float)P.span);
    if(kk > P.span - 1) kk = P.span - 1;
    return P.steps_min + kk;
}

// Fan-out of the next iteration (KPAX.cu:323-333), capped so the children
// never exceed the remaining tree capacity.
__device__ __forceinline__ Iter make_iter(uint32_t t, uint32_t ts, uint32_t W, uint32_t cap, const PlanParams& PP)
{
    Iter n;
    n.t                      = t;
    n.ts                     = ts;
    n.W                      = W;
    const uint32_t remaining = ts < cap ? cap - ts : 0u;
    uint32_t fan             = PP.fanout;
    uint64_t M               = (uint64_t)W * fan;
    if(M > remaining)
    {
        fan = (uint32_t)(int)((float)remaining / (float)W);
        if(fan > PP.fanout) fan = PP.fanout;
        if(fan == 0) fan = 1;
        M = (uint64_t)W * fan;
        if(M > remaining) M = remaining;
    }
    n.M      = (uint32_t)M;
    n.fan    = fan;
    n.status = t > PP.max_iter ? kDoneNotFound : (remaining == 0 ? kDoneBudget : kRunning);
    return n;
}

// First index i in [0, n) with off[i + 1] > k (off holds n + 1 ascending
// offsets): the segment that owns item k.
__device__ __forceinline__ uint32_t segment_of(const uint32_t* off, uint32_t n, uint32_t k)
{
    uint32_t lo = 0, hi = n;
    while(hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if(off[mid] <= k)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

}  // namespace dev
}  // namespace kpax
