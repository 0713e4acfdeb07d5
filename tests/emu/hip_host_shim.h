// hip_host_shim.h -- runs the HIP sources of this repository on the CPU.
//
// Test infrastructure: lets the unmodified kernels (ps4_kernels.hip) and the
// host library (ps4_lib.cpp) be compiled by g++ with AddressSanitizer, so
// every global-memory and LDS access is bounds-checked on the host before the
// code runs on a GPU.  Each workgroup is emulated by blockDim.x OS threads (one
// per lane); workgroups of a launch run one after another.  Barriers,
// wave shuffles, readfirstlane and atomics follow the GPU semantics the kernels
// rely on (all lanes of a wave active at every cross-lane operation).
#pragma once
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <math.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <mutex>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define PS4_KEEP(...) ((void)0)
#define PS4_OPAQUE(x) ((void)0)
#define PS4_LSHL3_ADD(a, b) (((uint32_t)(a) << 3) + (uint32_t)(b))
#define PS4_BCNT_ADD(x, a) ((uint32_t)__builtin_popcount((uint32_t)(x)) + (uint32_t)(a))
#define PS4_MAX_U16X2(a, b)                                                                       \
    ((std::max((uint32_t)(a) & 0xFFFFu, (uint32_t)(b) & 0xFFFFu)) | (std::max((uint32_t)(a) >> 16, (uint32_t)(b) >> 16) << 16))
#define PS4_SAD_U16(a, b, c)                                                                           \
    ((uint32_t)(c) + (uint32_t)std::abs((int)((uint32_t)(a) & 0xFFFFu) - (int)((uint32_t)(b) & 0xFFFFu)) + \
     (uint32_t)std::abs((int)((uint32_t)(a) >> 16) - (int)((uint32_t)(b) >> 16)))
#define PS4_MAD_U24(a, b, c) (((uint32_t)(a) & 0xFFFFFFu) * ((uint32_t)(b) & 0xFFFFFFu) + (uint32_t)(c))
#define PS4_MAD_I24(a, b, c)                                                                           \
    ((uint32_t)((int64_t)((int32_t)((uint32_t)(a) << 8) >> 8) * (int64_t)((int32_t)((uint32_t)(b) << 8) >> 8) + \
                (int64_t)(uint32_t)(c)))
// LDS: every __shared__ array lives in one section that is filled with garbage
// at each workgroup launch (real LDS starts with arbitrary contents).
#define __shared__ static __attribute__((section("emu_lds")))
extern "C" char __start_emu_lds[], __stop_emu_lds[];

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint4 { unsigned x, y, z, w; };
static inline uint4 make_uint4(unsigned a, unsigned b, unsigned c, unsigned d) { return uint4{a, b, c, d}; }

namespace emu {
struct Barrier {
    std::mutex m;
    std::condition_variable cv;
    unsigned n = 0, count = 0, gen = 0;
    void init(unsigned k) { n = k; count = 0; gen = 0; }
    void wait() {
        std::unique_lock<std::mutex> lk(m);
        const unsigned g = gen;
        if (++count == n) { count = 0; ++gen; cv.notify_all(); return; }
        cv.wait(lk, [&] { return gen != g; });
    }
};
struct Block {
    unsigned nthreads = 0;
    Barrier bar;
    std::vector<Barrier> wave_bar;
    std::vector<uint64_t> xchg;        // one slot per thread for cross-lane exchange
    std::vector<int> red;              // block reduction scratch
};
extern Block *g_block;
extern thread_local dim3 t_tid, t_bid, t_bdim, t_gdim;
}  // namespace emu

#define threadIdx (emu::t_tid)
#define blockIdx (emu::t_bid)
#define blockDim (emu::t_bdim)
#define gridDim (emu::t_gdim)

static inline void __syncthreads() { emu::g_block->bar.wait(); }
// a wave barrier; only kernels whose workgroup is one wave use it, so it is
// the workgroup barrier here
static inline void __builtin_amdgcn_wave_barrier() { emu::g_block->bar.wait(); }

static inline int emu_block_reduce(int v, bool is_or)
{
    emu::Block *b = emu::g_block;
    b->red[threadIdx.x] = v;
    b->bar.wait();
    int r = is_or ? 0 : 1;
    for (unsigned i = 0; i < b->nthreads; ++i) r = is_or ? (r | (b->red[i] != 0)) : (r & (b->red[i] != 0));
    b->bar.wait();
    return r;
}
static inline int __syncthreads_or(int p) { return emu_block_reduce(p, true); }
static inline int __syncthreads_and(int p) { return emu_block_reduce(p, false); }

static inline uint64_t emu_wave_exchange(uint64_t v, unsigned src_lane)
{
    emu::Block *b = emu::g_block;
    const unsigned t = threadIdx.x, w = t / 64, base = w * 64;
    b->xchg[t] = v;
    b->wave_bar[w].wait();
    const uint64_t r = b->xchg[base + (src_lane & 63)];
    b->wave_bar[w].wait();
    return r;
}
static inline unsigned __shfl_up(unsigned v, unsigned d, int width = 64)
{
    (void)width;
    const unsigned lane = threadIdx.x & 63;
    const unsigned src = lane >= d ? lane - d : lane;
    return (unsigned)emu_wave_exchange(v, src);
}
static inline unsigned __shfl_xor(unsigned v, int m, int width = 64)
{
    (void)width;
    const unsigned lane = threadIdx.x & 63;
    return (unsigned)emu_wave_exchange(v, lane ^ (unsigned)m);
}
static inline int __shfl(int v, int src, int width = 64)
{
    (void)width;
    return (int)(uint32_t)emu_wave_exchange((uint32_t)v, (unsigned)src);
}
static inline unsigned long long __ballot(int p)
{
    emu::Block *b = emu::g_block;
    const unsigned t = threadIdx.x, w = t / 64, base = w * 64;
    const unsigned n = std::min(64u, b->nthreads - base);
    b->xchg[t] = p ? 1u : 0u;
    b->wave_bar[w].wait();
    unsigned long long m = 0;
    for (unsigned i = 0; i < n; ++i)
        if (b->xchg[base + i]) m |= 1ull << i;
    b->wave_bar[w].wait();
    return m;
}
template <typename T> static inline T min(T a, T b) { return a < b ? a : b; }
template <typename T> static inline T max(T a, T b) { return a < b ? b : a; }
static inline unsigned __float_as_uint(float f) { unsigned u; memcpy(&u, &f, 4); return u; }
static inline float __uint_as_float(unsigned u) { float f; memcpy(&f, &u, 4); return f; }
static inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
static inline int __clzll(long long x) { return x ? __builtin_clzll((unsigned long long)x) : 64; }
// DPP moves used by the kernels: row_shr:1..15 (0x111..0x11F) inside 16-lane
// rows, row_bcast:15 (0x142) and row_bcast:31 (0x143); rows outside row_mask and
// lanes without a source return `old` (the bank mask is always 0xF here).
static inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl)
{
    (void)bank_mask; (void)bound_ctrl;
    const unsigned lane = threadIdx.x & 63, row = lane >> 4;
    int from = -1;
    if (ctrl >= 0x111 && ctrl <= 0x11F) {
        const unsigned n = (unsigned)ctrl - 0x110;
        if ((lane & 15) >= n) from = (int)(lane - n);
    } else if (ctrl == 0x142) {
        if (row >= 1) from = (int)(row * 16 - 1);
    } else if (ctrl == 0x143) {
        if (row >= 2) from = 31;
    } else {
        abort();
    }
    const int v = (int)(uint32_t)emu_wave_exchange((uint32_t)src, from < 0 ? lane : (unsigned)from);
    if (!((row_mask >> row) & 1) || from < 0) return old;
    return v;
}
static inline int __builtin_amdgcn_readlane(int v, int l) { return (int)(uint32_t)emu_wave_exchange((uint32_t)v, (unsigned)l); }
static inline int __builtin_amdgcn_readfirstlane(int v) { return (int)emu_wave_exchange((uint32_t)v, 0); }
static inline void __builtin_amdgcn_sched_barrier(int) {}   // an instruction-scheduling fence only
// set bits of a 64-lane mask below this lane: low word, then high word
static inline unsigned __builtin_amdgcn_mbcnt_lo(unsigned m, unsigned c)
{
    const unsigned lane = threadIdx.x & 63;
    return c + (unsigned)__builtin_popcount(lane >= 32 ? m : (m & ((1u << lane) - 1u)));
}
static inline unsigned __builtin_amdgcn_mbcnt_hi(unsigned m, unsigned c)
{
    const unsigned lane = threadIdx.x & 63;
    return c + (lane < 32 ? 0u : (unsigned)__builtin_popcount(m & ((1u << (lane - 32)) - 1u)));
}
static inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
// the device reciprocal is an approximation (refined by its callers); the
// emulation perturbs the exact one by a few parts in 2^40 so that the
// refinement is exercised
static inline double __builtin_amdgcn_rcp(double x) { return (1.0 / x) * (1.0 + 0x1p-38); }
// the reciprocal square root likewise, a few parts in 2^30 off
static inline double __builtin_amdgcn_rsq(double x) { return (1.0 / sqrt(x)) * (1.0 - 0x1p-30); }
static inline unsigned __umulhi(unsigned a, unsigned b) { return (unsigned)(((unsigned long long)a * b) >> 32); }
static inline unsigned __umul24(unsigned a, unsigned b) { return (a & 0xFFFFFFu) * (b & 0xFFFFFFu); }

template <typename T, typename V> static inline T atomicAdd(T *p, V v) { return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST); }
template <typename T, typename V> static inline T atomicOr(T *p, V v) { return __atomic_fetch_or(p, (T)v, __ATOMIC_SEQ_CST); }
template <typename T, typename V> static inline T atomicMax(T *p, V vv)
{
    const T v = (T)vv;
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return old;
}

// ---- runtime API subset -------------------------------------------------------------
typedef int hipError_t;
#define hipSuccess 0
#define hipErrorOutOfMemory 2
typedef void *hipStream_t;
typedef double *hipEvent_t;
#define hipStreamNonBlocking 1
#define hipMemcpyHostToDevice 1
#define hipMemcpyDeviceToHost 2
#define HIP_SYMBOL(x) (&(x))
struct hipDeviceProp_t { int multiProcessorCount; char name[256]; };
static inline const char *hipGetErrorString(hipError_t e) { return e ? "emulated error" : "no error"; }
static inline hipError_t hipGetLastError() { return hipSuccess; }
static inline hipError_t hipSetDevice(int) { return hipSuccess; }
static inline hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int)
{
    p->multiProcessorCount = 2;
    strcpy(p->name, "host-emulation");
    return hipSuccess;
}
static inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = (void *)1; return hipSuccess; }
static inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
static inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
static inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
static inline double emu_now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
static inline hipError_t hipEventCreate(hipEvent_t *e) { *e = new double(0); return hipSuccess; }
static inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
static inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { *e = emu_now_ms(); return hipSuccess; }
static inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) { *ms = (float)(*b - *a); return hipSuccess; }
static inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }   // launches run to completion
template <typename T> static inline hipError_t hipMalloc(T **p, size_t n)
{
    *p = (T *)malloc(n ? n : 1);
    if (!*p) return hipErrorOutOfMemory;
    memset(*p, 0xA5, n < ((size_t)64 << 20) ? n : ((size_t)64 << 20));   // device memory is not zero-initialised
    return hipSuccess;
}
template <typename T> static inline hipError_t hipHostMalloc(T **p, size_t n) { return hipMalloc(p, n); }
static inline hipError_t hipFree(void *p) { free(p); return hipSuccess; }
static inline hipError_t hipHostFree(void *p) { free(p); return hipSuccess; }
static inline hipError_t hipMemcpy(void *d, const void *s, size_t n, int) { memcpy(d, s, n); return hipSuccess; }
static inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, int, hipStream_t) { memcpy(d, s, n); return hipSuccess; }
static inline hipError_t hipMemGetInfo(size_t *f, size_t *t) { *f = (size_t)8 << 30; *t = (size_t)16 << 30; return hipSuccess; }
static inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *, int, size_t) { *n = 1; return hipSuccess; }
// device symbols are host objects here
#define HIP_SYMBOL(x) (&(x))
static inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = const_cast<void *>(sym); return hipSuccess; }
enum { hipErrorInvalidDevice = 101, hipErrorInvalidValue = 1 };
// graph capture is not emulated: asking for it fails like an unsupported call
typedef void *hipGraph_t;
typedef void *hipGraphExec_t;
typedef void *hipGraphNode_t;
typedef int hipGraphExecUpdateResult;
enum { hipStreamCaptureModeThreadLocal = 1, hipGraphExecUpdateSuccess = 0 };
static inline hipError_t hipStreamBeginCapture(hipStream_t, int) { return 801; }
static inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) { *g = nullptr; return 801; }
static inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
static inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
static inline hipError_t hipGraphExecUpdate(hipGraphExec_t, hipGraph_t, hipGraphNode_t *, hipGraphExecUpdateResult *) { return 801; }
static inline hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, void *, void *, size_t) { return 801; }
static inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return 801; }

namespace emu {
template <typename F, typename... Args>
void launch(F kernel, dim3 grid, dim3 block, Args... args)
{
    const unsigned nt = block.x;
    for (unsigned bx = 0; bx < grid.x; ++bx) {
        Block b;
        b.nthreads = nt;
        b.bar.init(nt);
        b.wave_bar = std::vector<Barrier>((nt + 63) / 64);
        for (unsigned w = 0; w < b.wave_bar.size(); ++w) b.wave_bar[w].init(std::min(64u, nt - w * 64));
        b.xchg.assign(nt, 0);
        b.red.assign(nt, 0);
        g_block = &b;
        memset(__start_emu_lds, 0x5A ^ (int)(bx & 0xFF), (size_t)(__stop_emu_lds - __start_emu_lds));
        std::vector<std::thread> th;
        th.reserve(nt);
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                t_tid = dim3(t, 0, 0);
                t_bid = dim3(bx, 0, 0);
                t_bdim = block;
                t_gdim = grid;
                kernel(args...);
            });
        for (auto &x : th) x.join();
        g_block = nullptr;
    }
}
}  // namespace emu

#define hipLaunchKernelGGL(k, grid, block, shmem, stream, ...) emu::launch(k, dim3(grid), dim3(block), __VA_ARGS__)
