// Internal helpers shared by the HIP sources of libg500.so: the library
// context (device, stream, events, tuning knobs), error handling and small
// device utilities for wave64 gfx950 code.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "g500_ext.h"
#include "graph_generation.h"
#include "traversal.h"
#include "worker.h"

namespace g500 {

constexpr int kWave = 64;

struct Context {
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    int num_cus = 0;
    int error_mode = G500_ERRORS_ABORT;
    int status = 0;
    char msg[1024] = {0};
    double last_ms = 0.0;
    const Mrg_Transition_Matrix *d_mrg_table = nullptr;  // [24][256] on the device
    // tuning knobs
    double bfs_alpha = 15.0;      // top-down -> bottom-up when m_frontier > m_unexplored / alpha
    double bfs_beta = 18.0;       // bottom-up -> top-down when n_frontier < n / beta
    int bfs_block = 512;          // threads per block of the persistent BFS kernel
    int bfs_hub_degree = 2048;    // frontier vertices above this degree are split grid-wide
    int sssp_block = 512;
    int gen_blocks = 0;           // 0: one block per 256-edge tile (capped)
};

Context &ctx();

// Records a failure. In abort mode prints and exits; in return mode keeps the
// message and returns so the caller can unwind.
void fail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
inline bool failed() { return ctx().status != 0; }

#define G500_HIP(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            ::g500::fail("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,              \
                         hipGetErrorString(e_));                                         \
    } while (0)

// Device allocation from the worker's arena (256-byte aligned); fails loudly.
void *arena_alloc(Worker_State *ws, u64 bytes, const char *what);

// Timing bracket on the library stream (hipEvents).
void timer_start();
void timer_stop();   // synchronizes, stores ctx().last_ms

inline void sync() { G500_HIP(hipStreamSynchronize(ctx().stream)); }

// Grid size for a cooperative (persistent) kernel: every block co-resident.
int cooperative_blocks(const void *kernel, int block, size_t dyn_lds);

inline unsigned int grid_for(u64 work, int block, u64 cap = 1u << 20) {
    u64 b = (work + block - 1) / block;
    if (b == 0) b = 1;
    if (b > cap) b = cap;
    return (unsigned int)b;
}

// ---------------------------------------------------------------- device side
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Prefix count of set bits below this lane in a 64-bit wave mask.
__device__ __forceinline__ u32 mask_rank(u64 mask) {
    return __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
}

__device__ __forceinline__ u64 ballot(bool p) { return __ballot(p); }

// Wave-aggregated append: every lane with `want` gets a distinct slot of
// *counter; one atomic per wave.
__device__ __forceinline__ u32 wave_append(u32 *counter, bool want) {
    u64 m = ballot(want);
    if (m == 0) return 0;
    int leader = __ffsll((unsigned long long)m) - 1;
    u32 base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, (u32)__popcll(m));
    base = __shfl(base, leader);
    return base + mask_rank(m);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ u64 shfl_xor64(u64 v, int m) {
    u32 lo = __shfl_xor((u32)v, m), hi = __shfl_xor((u32)(v >> 32), m);
    return ((u64)hi << 32) | lo;
}

__device__ __forceinline__ u64 ld_acquire(const u64 *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 ld_relaxed(const u32 *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 ld_relaxed(const u64 *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Control words shared by all XCDs are only written with device-scope atomics
// (never plain stores), so no XCD holds a dirty copy of a line that others
// update atomically.
__device__ __forceinline__ void st_relaxed(u32 *p, u32 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(u64 *p, u64 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(i32 *p, i32 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier for cooperative launches: generation counter + arrival count.
// Every wait is bounded; on timeout it raises *error and returns so that all
// waves still leave the kernel.
struct alignas(256) GridBarrier {
    u32 count;
    u32 gen;
    u32 error;   // set on timeout or by kernels that detect an inconsistency
};

// Per-block trace of the last barrier passed (count, phase, two tags), for
// diagnosing a barrier timeout.
constexpr int kMaxTraceBlocks = 4096;
struct BlockTrace {
    u32 barriers, phase, a, b;
};

__device__ __forceinline__ void grid_sync(GridBarrier *b, u32 *error) {
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 g = __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        u32 arrived = __hip_atomic_fetch_add(&b->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == gridDim.x - 1) {
            __hip_atomic_store(&b->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&b->gen, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const u64 t0 = wall_clock64();
            while (__hip_atomic_load(&b->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(2);
                // ~10 s at the 100 MHz constant clock: only reached if blocks are not
                // co-resident, which the cooperative launch rules out.
                if (wall_clock64() - t0 > 1000000000ull) {
                    __hip_atomic_store(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

}  // namespace g500
