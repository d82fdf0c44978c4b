// common.h -- shared definitions for the gfx950 RBSC library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

typedef uint32_t u32;
typedef uint64_t u64;

#define RBSC_WAVE 64

#if defined(__HIP_DEVICE_COMPILE__)
// v * 5 and v * 9 (mod 2^64) as one full-rate v_lshl_add_u64 each
// ((v << 2) + v, (v << 3) + v); the compiler otherwise multiplies by the
// small constants with two quarter-rate 32-bit multiply-adds.
__device__ __forceinline__ u64 rbsc_mul5(u64 v) {
    u64 r;
    asm("v_lshl_add_u64 %0, %1, 2, %1" : "=v"(r) : "v"(v));
    return r;
}
__device__ __forceinline__ u64 rbsc_mul9(u64 v) {
    u64 r;
    asm("v_lshl_add_u64 %0, %1, 3, %1" : "=v"(r) : "v"(v));
    return r;
}
// rotate left by 7 as two funnel shifts of the 32-bit halves
__device__ __forceinline__ u64 rbsc_rotl7(u64 v) {
    const u32 lo = (u32)v, hi = (u32)(v >> 32);
    const u32 nhi = __builtin_amdgcn_alignbit(hi, lo, 25);
    const u32 nlo = __builtin_amdgcn_alignbit(lo, hi, 25);
    return ((u64)nhi << 32) | nlo;
}
#endif

// The signature hash of rbsc.cu:65-71 (also rsc.cpp:63-69): the SplitMix64
// finalizer followed by the xoshiro** scrambler.  Every step is a bijection on
// Z/2^64, so distinct colours never collide under it.
__host__ __device__ __forceinline__ u64 rbsc_randomize(u64 v) {
    v = v + 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
#if defined(__HIP_DEVICE_COMPILE__)
    return rbsc_mul9(rbsc_rotl7(rbsc_mul5(v ^ (v >> 31))));
#else
    v = (v ^ (v >> 31)) * 5ULL;
    return ((v << 7) | (v >> 57)) * 9ULL;
#endif
}

// Plain SplitMix64 step, used by the seeded generators.
__host__ __device__ __forceinline__ u64 rbsc_splitmix64(u64 x) {
    u64 z = x + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Look-back status words (radix sort and scans): state in bits 63:62,
// epoch in bits 61:40, count in bits 39:0.  The epoch makes stale words from
// earlier launches invalid, so status arrays never need clearing.
#define LB_STATE_SHIFT 62
#define LB_EPOCH_SHIFT 40
#define LB_COUNT_MASK ((1ULL << 40) - 1)
#define LB_EPOCH_MASK ((1ULL << 22) - 1)
#define LB_AGG 1ULL
#define LB_PREFIX 2ULL

__device__ __forceinline__ u64 lb_pack(u64 state, u32 epoch, u64 count) {
    return (state << LB_STATE_SHIFT) | ((u64)(epoch & LB_EPOCH_MASK) << LB_EPOCH_SHIFT) |
           (count & LB_COUNT_MASK);
}

__device__ __forceinline__ void lb_store(u64* p, u64 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 lb_load(u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded spin for look-back waits: predecessors always hold their tickets
// and finish, so the bound is only a guard against a logic error turning
// into a GPU hang.  On overflow the kernel raises `err` and continues.
#define LB_SPIN_LIMIT (1u << 24)

// Error flag bits raised by kernels (read back by the host).
#define KERR_SPIN 1u
#define KERR_PROBE 2u  // a hash-table insert found no free slot

// Host-side error check that records the failure and returns from an int
// function.
#define HIP_TRY(expr)                                                           \
    do {                                                                        \
        hipError_t _e = (expr);                                                 \
        if (_e != hipSuccess) {                                                 \
            rbsc_set_error("%s failed at %s:%d: %s", #expr, __FILE__, __LINE__, \
                           hipGetErrorString(_e));                              \
            return RBSC_ERR_HIP;                                                \
        }                                                                       \
    } while (0)

extern "C" void rbsc_set_error(const char* fmt, ...);

// Device memory that a context takes and gives back while other contexts may
// have work in flight (its workspace and that workspace's growth, per-launch
// buffers).  hipFree and hipHostFree wait for the work of every stream on the
// device, so one serving worker growing its workspace would stall every other
// refinement; stream-ordered allocation from the device's memory pool orders
// the free after this stream's work only.  The pool keeps freed memory for
// reuse (its release threshold is raised when a context is created).  Memory
// from ctx_alloc is given back with ctx_free only.  When the device is out of
// memory, the pool's cached blocks (which may all be of other sizes) are
// handed back once all frees have completed, and the allocation is retried.
template <class T>
inline hipError_t ctx_alloc(T** p, size_t bytes, hipStream_t s) {
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(p), bytes, s);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        int dev = 0;
        hipMemPool_t pool = nullptr;
        if (hipDeviceSynchronize() == hipSuccess && hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess &&
            hipMemPoolTrimTo(pool, 0) == hipSuccess)
            e = hipMallocAsync(reinterpret_cast<void**>(p), bytes, s);
        if (e != hipSuccess) (void)hipGetLastError();
    }
    return e;
}
inline void ctx_free(void* p, hipStream_t s) {
    if (p) (void)hipFreeAsync(p, s);
}
