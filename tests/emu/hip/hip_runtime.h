// hip/hip_runtime.h for a CPU build of the kpax kernels (test harness).
//
// The kernel source is compiled unchanged by the host compiler: each
// work-group runs as blockDim.x host threads (work-groups one after another,
// in launch order), __syncthreads is a barrier over the group, and the wave
// intrinsics (__shfl_up, __shfl_xor, __ballot) exchange values through a
// per-wave buffer between barriers over the 64 lanes of that wave.  Device
// memory is host memory from malloc, so AddressSanitizer sees every access a
// kernel makes.  Runtime calls are synchronous.
#pragma once

#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define amdgpu_waves_per_eu(...) unused
#define __shared__ static
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) ((const void*)&(x))
inline char* emu_dyn_shared = nullptr;  // dynamic LDS of the running work-group
#define HIP_DYNAMIC_SHARED(type, var) type* var = reinterpret_cast<type*>(emu_dyn_shared);

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_idx {
    unsigned x, y, z;
};
struct uint4 {
    uint32_t x, y, z, w;
};
struct float4 {
    float x, y, z, w;
};
struct float2 {
    float x, y;
};
struct int4 {
    int32_t x, y, z, w;
};
inline uint4 make_uint4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return {a, b, c, d}; }
inline int4 make_int4(int32_t a, int32_t b, int32_t c, int32_t d) { return {a, b, c, d}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }

inline thread_local emu_idx threadIdx{0, 0, 0};
inline thread_local emu_idx blockIdx{0, 0, 0};
inline emu_idx blockDim{1, 1, 1};
inline emu_idx gridDim{1, 1, 1};

struct emu_group {
    std::unique_ptr<std::barrier<>> bar;
    std::vector<std::unique_ptr<std::barrier<>>> wave_bar;
    std::vector<uint64_t> xbuf;  // 64 slots per wave
    std::vector<uint32_t> mbuf;  // matrix-core operands: 64 lanes x 8 words per wave
};
inline emu_group* emu_g = nullptr;

inline void __syncthreads() { emu_g->bar->arrive_and_wait(); }

inline uint64_t emu_exchange(uint64_t v, unsigned src_lane, bool* ok = nullptr)
{
    const unsigned t = threadIdx.x, w = t / 64, lane = t % 64;
    auto& b          = *emu_g->wave_bar[w];
    b.arrive_and_wait();
    emu_g->xbuf[w * 64 + lane] = v;
    b.arrive_and_wait();
    const uint64_t r = emu_g->xbuf[w * 64 + (src_lane % 64)];
    b.arrive_and_wait();
    (void)ok;
    return r;
}

template<class T> inline T __shfl_up(T x, unsigned off, int = 64)
{
    const unsigned lane = threadIdx.x % 64;
    uint64_t bits       = 0;
    std::memcpy(&bits, &x, sizeof(T));
    const uint64_t r = emu_exchange(bits, lane >= off ? lane - off : lane);
    T y;
    std::memcpy(&y, &r, sizeof(T));
    return y;
}

template<class T> inline T __shfl_xor(T x, unsigned m, int = 64)
{
    const unsigned lane = threadIdx.x % 64;
    uint64_t bits       = 0;
    std::memcpy(&bits, &x, sizeof(T));
    const uint64_t r = emu_exchange(bits, lane ^ m);
    T y;
    std::memcpy(&y, &r, sizeof(T));
    return y;
}

template<class T> inline T __shfl_down(T x, unsigned off, int = 64)
{
    const unsigned lane = threadIdx.x % 64;
    uint64_t bits       = 0;
    std::memcpy(&bits, &x, sizeof(T));
    const uint64_t r = emu_exchange(bits, lane + off < 64 ? lane + off : lane);
    T y;
    std::memcpy(&y, &r, sizeof(T));
    return y;
}

template<class T> inline T __shfl(T x, int src, int = 64)
{
    uint64_t bits = 0;
    std::memcpy(&bits, &x, sizeof(T));
    const uint64_t r = emu_exchange(bits, (unsigned)src);
    T y;
    std::memcpy(&y, &r, sizeof(T));
    return y;
}

inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }

// a wave's issue priority only orders instruction issue on the GPU
inline void __builtin_amdgcn_s_setprio(int) {}

inline uint64_t __ballot(int pred)
{
    const unsigned t = threadIdx.x, w = t / 64, lane = t % 64;
    auto& b          = *emu_g->wave_bar[w];
    b.arrive_and_wait();
    emu_g->xbuf[w * 64 + lane] = pred ? 1 : 0;
    b.arrive_and_wait();
    uint64_t m = 0;
    for(unsigned i = 0; i < 64; ++i)
        if(emu_g->xbuf[w * 64 + i]) m |= 1ull << i;
    b.arrive_and_wait();
    return m;
}

inline uint32_t __umulhi(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

template<class T> inline T atomicAdd(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
template<class T> inline T atomicOr(T* p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
template<class T> inline T atomicMin(T* p, T v)
{
    T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while(v < cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return cur;
}
template<class T> inline T atomicCAS(T* p, T cmp, T v)
{
    __atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return cmp;
}

#define __HIP_MEMORY_SCOPE_AGENT 0
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), (order))
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), (order))

// ---- matrix-core tile (kpax_mlp.cuh declares the lane layout) ----------------
#define KPAX_HOST_EMULATION 1
struct f32x16 {
    float v[16];
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
};

struct f32x4 {
    float v[4];
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
};

inline void emu_wave_barrier() { emu_g->wave_bar[threadIdx.x / 64]->arrive_and_wait(); }

inline float emu_bf16(uint32_t word, unsigned half)
{
    const uint32_t b = half ? (word >> 16) : (word & 0xffffu);
    const uint32_t f = b << 16;
    float x;
    std::memcpy(&x, &f, 4);
    return x;
}

// D(32x32) = A(32x16) B(16x32) + C: lane l supplies A row l%32 and B column
// l%32 for k = 8 (l / 32) .. +7 and receives column l%32, rows
// 8 (i / 4) + 4 (l / 32) + (i % 4).  All 64 lanes of the wave call it.
inline f32x16 mfma_32x32x16(uint4 a, uint4 b, f32x16 c)
{
    const unsigned t = threadIdx.x, w = t / 64, lane = t % 64;
    auto& bar        = *emu_g->wave_bar[w];
    uint32_t* buf    = emu_g->mbuf.data() + (size_t)w * 64 * 8;
    bar.arrive_and_wait();
    const uint32_t in[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    for(int j = 0; j < 8; ++j) buf[lane * 8 + j] = in[j];
    bar.arrive_and_wait();
    f32x16 d;
    const unsigned col = lane % 32;
    for(int i = 0; i < 16; ++i)
    {
        const unsigned row = 8 * (i / 4) + 4 * (lane / 32) + (i % 4);
        float s            = c[i];
        for(unsigned k = 0; k < 16; ++k)
        {
            const unsigned la = row + 32 * (k / 8), lb = col + 32 * (k / 8), e = k % 8;
            s += emu_bf16(buf[la * 8 + e / 2], e % 2) * emu_bf16(buf[lb * 8 + 4 + e / 2], e % 2);
        }
        d[i] = s;
    }
    bar.arrive_and_wait();
    return d;
}

// D(16x16) = A(16x32) B(32x16) + C: lane l = j + 16 g supplies A row j and
// B column j for the k-slots (g, 0..7) and receives column j, rows 4 g + e.
inline f32x4 mfma_16x16x32(uint4 a, uint4 b, f32x4 c)
{
    const unsigned t = threadIdx.x, w = t / 64, lane = t % 64;
    auto& bar        = *emu_g->wave_bar[w];
    uint32_t* buf    = emu_g->mbuf.data() + (size_t)w * 64 * 8;
    bar.arrive_and_wait();
    const uint32_t in[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    for(int j = 0; j < 8; ++j) buf[lane * 8 + j] = in[j];
    bar.arrive_and_wait();
    f32x4 d;
    const unsigned col = lane % 16;
    for(int i = 0; i < 4; ++i)
    {
        const unsigned row = 4 * (lane / 16) + i;
        float s            = c[i];
        for(unsigned k = 0; k < 32; ++k)
        {
            const unsigned la = row + 16 * (k / 8), lb = col + 16 * (k / 8), e = k % 8;
            s += emu_bf16(buf[la * 8 + e / 2], e % 2) * emu_bf16(buf[lb * 8 + 4 + e / 2], e % 2);
        }
        d[i] = s;
    }
    bar.arrive_and_wait();
    return d;
}

// ---- runtime ----------------------------------------------------------------
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorUnknown = 999 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
typedef struct emu_stream_t* hipStream_t;
typedef struct emu_event_t* hipEvent_t;
struct hipDeviceProp_t {
    int multiProcessorCount;
    size_t sharedMemPerBlock;
    size_t maxSharedMemoryPerMultiProcessor;
};
#define hipStreamNonBlocking 1
#define hipEventDisableTiming 2
#define hipHostMallocDefault 0

inline hipError_t hipGetDeviceCount(int* n)
{
    *n = 1;
    return hipSuccess;
}
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDevice(int* d)
{
    *d = 0;
    return hipSuccess;
}
inline hipError_t hipGetSymbolAddress(void** p, const void* symbol)
{
    *p = const_cast<void*>(symbol);
    return hipSuccess;
}
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int)
{
    // KPAX_EMU_CUS: the compute units the emulated device reports (default 1)
    const char* cus                     = std::getenv("KPAX_EMU_CUS");
    p->multiProcessorCount              = cus ? std::max(1, std::atoi(cus)) : 1;
    p->sharedMemPerBlock                = 160 << 10;
    p->maxSharedMemoryPerMultiProcessor = 160 << 10;
    return hipSuccess;
}
inline hipError_t hipMalloc(void** p, size_t n)
{
    *p = std::malloc(n ? n : 1);
    std::memset(*p, 0xA5, n ? n : 1);  // device memory starts as garbage
    return *p ? hipSuccess : hipErrorUnknown;
}
template<class T> inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
inline hipError_t hipFree(void* p)
{
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipHostMalloc(void** p, size_t n, int) { return hipMalloc(p, n); }
inline hipError_t hipHostFree(void* p) { return hipFree(p); }
inline hipError_t hipHostGetDevicePointer(void** d, void* h, unsigned)
{
    *d = h;
    return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind)
{
    std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t)
{
    return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemset(void* d, int v, size_t n)
{
    std::memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { return hipMemset(d, v, n); }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned)
{
    *s = (hipStream_t)0x1;
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e)
{
    *e = (hipEvent_t)0x1;
    return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t)
{
    *ms = 0.0f;
    return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
// graph capture is not emulated: the library then launches the same kernels directly
typedef struct emu_graph_t* hipGraph_t;
typedef struct emu_graph_exec_t* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorUnknown; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return hipErrorUnknown; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) { return hipErrorUnknown; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorUnknown; }
inline const char* hipGetErrorString(hipError_t) { return "emulated"; }

// Persistent workers, one per work-item slot of the largest work-group.
struct emu_pool {
    std::vector<std::thread> th;
    std::mutex m;
    std::condition_variable cv_start, cv_done;
    uint64_t gen       = 0;
    unsigned active    = 0;
    unsigned remaining = 0;
    std::function<void(unsigned)> job;
    explicit emu_pool(unsigned n)
    {
        for(unsigned i = 0; i < n; ++i) th.emplace_back([this, i] { worker(i); });
    }
    void worker(unsigned i)
    {
        uint64_t seen = 0;
        for(;;)
        {
            std::function<void(unsigned)> j;
            {
                std::unique_lock<std::mutex> lk(m);
                cv_start.wait(lk, [&] { return gen != seen; });
                seen = gen;
                if(i >= active) continue;
                j = job;
            }
            j(i);
            std::lock_guard<std::mutex> lk(m);
            if(--remaining == 0) cv_done.notify_one();
        }
    }
    void run(unsigned n, std::function<void(unsigned)> f)
    {
        {
            std::lock_guard<std::mutex> lk(m);
            job       = std::move(f);
            active    = n;
            remaining = n;
            ++gen;
        }
        cv_start.notify_all();
        std::unique_lock<std::mutex> lk(m);
        cv_done.wait(lk, [&] { return remaining == 0; });
    }
};

inline emu_pool& emu_workers()
{
    static emu_pool* p = new emu_pool(1024);  // never joined: lives until exit
    return *p;
}

template<class F> inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, F, int, size_t)
{
    *n = 1;
    return hipSuccess;
}

template<class... KArgs, class... Args>
void emu_launch(void (*k)(KArgs...), dim3 grid, dim3 block, size_t shmem, Args... args)
{
    std::vector<char> dyn(shmem ? shmem : 1);
    emu_dyn_shared = dyn.data();
    gridDim  = {grid.x, grid.y, 1};
    blockDim = {block.x, 1, 1};
    const unsigned nw = (block.x + 63) / 64;
    for(unsigned by = 0; by < grid.y; ++by)
    for(unsigned b = 0; b < grid.x; ++b)
    {
        emu_group g;
        g.bar = std::make_unique<std::barrier<>>(block.x);
        for(unsigned w = 0; w < nw; ++w)
            g.wave_bar.push_back(std::make_unique<std::barrier<>>(std::min(64u, block.x - 64 * w)));
        g.xbuf.assign(64 * nw, 0);
        g.mbuf.assign(64 * 8 * nw, 0);
        emu_g = &g;
        emu_workers().run(block.x, [&, b](unsigned t) {
            threadIdx = {t, 0, 0};
            blockIdx  = {b, by, 0};
            k(args...);
            // a work-item that returns early no longer takes part in barriers
            g.wave_bar[t / 64]->arrive_and_drop();
            g.bar->arrive_and_drop();
        });
        emu_g = nullptr;
    }
}

#define hipLaunchKernelGGL(k, grid, block, shmem, stream, ...) emu_launch(k, grid, block, (size_t)(shmem), __VA_ARGS__)
