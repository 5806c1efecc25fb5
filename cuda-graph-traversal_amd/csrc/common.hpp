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
constexpr int kSsspBatchRoots = 32;  // roots (and label stride) of a batched SSSP launch

struct Context {
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    int num_cus = 0;
    double wall_clock_khz = 100000.0;  // rate of the device constant clock (wall_clock64)
    int error_mode = G500_ERRORS_ABORT;
    int status = 0;
    char msg[1024] = {0};
    double last_ms = 0.0;
    const Mrg_Transition_Matrix *d_mrg_table = nullptr;  // [24][256] on the device
    u64 *d_counters = nullptr;    // 64 bytes of device counters for the validation entry points
    // tuning knobs
    double bfs_alpha = 60.0;      // top-down -> bottom-up when m_frontier > m_unexplored / alpha
    double bfs_beta = 60.0;       // bottom-up -> top-down when n_frontier < n / beta
    int bfs_block = 512;          // threads per block of the persistent BFS kernel
    int bfs_hub_degree = 2048;    // frontier vertices above this degree are split grid-wide
    double batch_alpha = 15.0;    // the same two thresholds for the batched multi-root BFS
    double batch_beta = 18.0;
    int batch_heavy_degree = 1 << 20; // batched bottom-up: vertices above this degree are split grid-wide
    int sssp_block = 512;
    int batch_bfs_block = 512;    // threads per block of the batched BFS and batched SSSP kernels
    int batch_sssp_block = 512;
    int coop_launch = 1;          // 1: cooperative launch first; 0: plain launch of the co-resident grid
    int gen_blocks = 0;           // 0: one block per 256-edge tile (capped)
    int relabel = 1;              // 1: CSR builds add the degree-ordered traversal copy
    int sssp_prune = 1;           // 1: and the SSSP copy without two-hop-dominated entries
    int sssp_screen = 0;          // 1: single-root SSSP relaxations check the settled bitmap before the probe
    double batch_pull = 0.2;       // batched SSSP pulls its far-class flush when the window's items >= this share
                                   // of the vertices with entries (0: never)
    double sssp_pull = 0.05;       // single-root SSSP pulls a bucket's heavy relaxations when the bucket's light
                                  // items reach this share of the unsettled vertices (0: always push)
};

Context &ctx();

// Error state of a library-internal thread (the serving slots and copier):
// while a thread has a sink, its failures are recorded there instead of in
// the context, so they never leak into the caller's status and never abort.
struct ErrorSink {
    int status = 0;
    char msg[512] = {0};
};
extern thread_local ErrorSink *t_error_sink;

// Records a failure. In abort mode prints and exits; in return mode keeps the
// message and returns so the caller can unwind. Thread-safe.
void fail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
inline bool failed() { return t_error_sink ? t_error_sink->status != 0 : ctx().status != 0; }

// Frees the library's own device resources (stream, events, skip table,
// counters); worker_init creates them again.
void release_library_resources();

// Depth of every vertex from a parent array (bfs_compute_dist_from_pred's
// semantics) into dist[n] on the library stream; `changed` is a device u32
// the level loop uses. Blocking.
void depth_from_pred(const Oned_Graph *g, const i64 *pred, i64 *dist, u64 root, u32 *changed);

// Distances (f32, -1.0f unreached, original ids) from the graph's highest-degree
// vertex into dist_dev[n], by the single-root SSSP on temporary arena memory.
// Blocking; false after a failure.
bool hub_distances(Worker_State *ws, const Oned_Graph *g, f32 *dist_dev);

// Resolves, for the current device, the device address of the static scratch
// array of the no-allocation validation entry point (once per device).
void resolve_noalloc_scratch();

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
// The runtime's blocks-per-CU figure is capped by the kernel's LDS (static +
// dyn_lds) rounded up to 2 KiB per block, so that a block count that only fits
// the CU's LDS without allocation rounding is never taken.
int cooperative_blocks(const void *kernel, int block, size_t dyn_lds);

// Launches a persistent kernel whose grid fits on the device at once. With
// `cooperative`, uses a cooperative launch; if the runtime refuses it (the
// cooperative queue is held by another process on the same device), or
// without `cooperative` (several persistent kernels sharing the device, each
// with its share of the co-resident grid), launches the grid normally. The
// kernels' grid barriers are bounded, so a grid that is not fully resident
// ends with an error instead of waiting forever. stream 0: the library stream.
void launch_persistent(const void *kernel, int grid, int block, void **args, hipStream_t stream = nullptr,
                       bool cooperative = true, size_t dyn_lds = 0);

// Traversal view of a graph: the degree-ordered copy when it exists, else the
// original arrays. Kernels run in traversal ids and report original ids
// through orig_id (identity without the copy).
struct TraversalView {
    const u64 *rs;
    const u32 *col;
    const f32 *w;
    const u32 *perm, *iperm;
};
inline TraversalView traversal_view(const Oned_Graph *g) {
    if (g->perm) return {g->rowstarts_t, g->column_t, g->weights_t, g->perm, g->iperm};
    return {g->rowstarts, g->column, g->weights, nullptr, nullptr};
}
// SSSP view: the pruned SSSP copy when it exists, else the traversal view.
inline TraversalView sssp_view(const Oned_Graph *g) {
    if (g->rowstarts_s) return {g->rowstarts_s, g->column_s, g->weights_s, g->perm, g->iperm};
    return traversal_view(g);
}
// Distances at or above this bound void the pruned copy's guarantee (error 32).
constexpr f32 kPruneDistBound = 2048.0f;
__device__ __forceinline__ u32 orig_id(const u32 *iperm, u32 v) { return iperm ? iperm[v] : v; }
__device__ __forceinline__ u32 trav_id(const u32 *perm, u64 v) { return perm ? perm[v] : (u32)v; }

// Asynchronous batched launches on a caller's stream with a caller's grid (0:
// the full co-resident grid, cooperative, on the library stream with timing).
// They validate like the blocking entry points and return false on failure.
bool bfs_batch_enqueue(const Oned_Graph *g, Bfs_Batch_State *b, const u64 *roots, i32 nroots, hipStream_t stream,
                       int grid);
void bfs_batch_depth_enqueue(const Oned_Graph *g, const Bfs_Batch_State *b, i32 r, i64 *depth, hipStream_t stream);
u32 bfs_batch_error(const Bfs_Batch_State *b, hipStream_t stream);  // synchronizes the stream
int bfs_batch_grid();
// offsets: null, or per root a distance offset (say its distance from a hub
// vertex; negative: none) from which the roots' bucket lags are set
bool sssp_batch_enqueue(const Oned_Graph *g, Sssp_Batch_State *s, const u64 *roots, const f32 *offsets, i32 nroots,
                        hipStream_t stream, int grid);
u32 sssp_batch_error(const Sssp_Batch_State *s, hipStream_t stream);  // synchronizes the stream
// Computes s's light/mid/long/far row splits and row records for its delta and
// class weights on g (only the split arrays and row_info of s are used),
// unless they already hold them; returns the clamped class weights.
void sssp_batch_prepare(const Oned_Graph *g, Sssp_Batch_State *s, hipStream_t st, f32 *wmid, f32 *wlong, f32 *wfar);
int sssp_batch_grid();
// Per-root summaries of the last batch: out[r] = reached vertices, out[R + r] =
// depth sum (BFS, R = 64) or distance sum in 2^-24 fixed point (SSSP, R = 32).
void bfs_batch_summary_enqueue(const Bfs_Batch_State *b, u64 *out, hipStream_t stream);
void sssp_batch_summary_enqueue(const Oned_Graph *g, const Sssp_Batch_State *s, u64 *out, hipStream_t stream);

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

__device__ __forceinline__ u32 wave_min(u32 v) {
    for (int o = 32; o > 0; o >>= 1) {
        const u32 y = __shfl_xor(v, o);
        v = y < v ? y : v;
    }
    return v;
}

__device__ __forceinline__ u64 wave_min64(u64 v) {
    for (int o = 32; o > 0; o >>= 1) {
        const u32 lo = __shfl_xor((u32)v, o), hi = __shfl_xor((u32)(v >> 32), o);
        const u64 y = ((u64)hi << 32) | lo;
        v = y < v ? y : v;
    }
    return v;
}

__device__ __forceinline__ u64 shfl64(u64 v, int src) {
    u32 lo = __shfl((u32)v, src), hi = __shfl((u32)(v >> 32), src);
    return ((u64)hi << 32) | lo;
}

__device__ __forceinline__ u64 shfl_xor64(u64 v, int m) {
    u32 lo = __shfl_xor((u32)v, m), hi = __shfl_xor((u32)(v >> 32), m);
    return ((u64)hi << 32) | lo;
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        T y = __shfl_up(v, o);
        if (lane_id() >= o) v += y;
    }
    return v;
}

// Segmented walk over the concatenated work ranges of one wave's 64 items:
// lane j's item owns positions [off, off + deg) of the concatenation. For the
// 64-position chunk starting at c, returns the item whose range holds
// position c + lane (valid when c + lane < total), and sets `inter`/`last`
// for the lane's own item: whether its range meets the chunk and the lane
// that holds its last position in the chunk. `slot` is the wave's 64 LDS
// words; LDS operations of one wave complete in order.
__device__ __forceinline__ int chunk_owner(volatile int *slot, u32 off, u32 deg, u32 c, bool &inter, int &last) {
    const u32 lane = lane_id();
    const u32 lo = off > c ? off : c;
    const u32 hi = off + deg < c + kWave ? off + deg : c + kWave;
    inter = lo < hi;
    last = inter ? (int)(hi - 1 - c) : (int)lane;
    slot[lane] = -1;
    if (inter) slot[lo - c] = (int)lane;
    int j = slot[lane];
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(j, o);
        if ((int)lane >= o && y > j) j = y;
    }
    return j;
}

// ------------------------------------------------------------ hub work lists
// A hub list holds high-degree vertices together with the offset of each
// one's edge range in a flattened work space where every range is padded to a
// multiple of kHubChunk. One packed u64 counter (count << 38 | padded edge
// total) reserves the slot and the offset with a single atomic, so offsets
// increase with the slot index and a chunk of the work space maps to exactly
// one hub by a monotone search.
constexpr u64 kHubChunk = 256;
constexpr int kHubCountShift = 38;
constexpr u64 kHubTotalMask = (1ull << kHubCountShift) - 1;

__device__ __forceinline__ u64 hub_count(u64 packed) { return packed >> kHubCountShift; }
__device__ __forceinline__ u64 hub_total(u64 packed) { return packed & kHubTotalMask; }

// Wave-converged append of (v, len) for lanes with `want`; returns the lane's
// slot (~0 when it appended nothing).
__device__ __forceinline__ u64 hub_append(u64 *counter, u32 *hv, u64 *hoff, u32 *hlen, bool want, u32 v,
                                          u32 len, u64 cap, u32 *error) {
    const u64 m = ballot(want);
    if (!m) return ~0ull;
    const u64 padded = want ? (u64)(len + kHubChunk - 1) / kHubChunk * kHubChunk : 0;
    const u64 incl = wave_inclusive_scan(padded);
    const u64 total = __shfl(incl, kWave - 1);
    const int leader = __ffsll((unsigned long long)m) - 1;
    u64 old = 0;
    if (lane_id() == leader)
        old = atomicAdd((unsigned long long *)counter, ((u64)__popcll(m) << kHubCountShift) | total);
    old = __shfl(old, leader);
    if (want) {
        const u64 slot = hub_count(old) + mask_rank(m);
        if (slot >= cap) {
            atomicOr(error, 2u);
            return ~0ull;
        }
        hv[slot] = v;
        hoff[slot] = hub_total(old) + incl - padded;
        hlen[slot] = len;
        return slot;
    }
    return ~0ull;
}

// Largest i in [lo, n) with hoff[i] <= p, searching forward from lo
// (galloping, then bisection). Requires hoff[lo] <= p.
__device__ __forceinline__ u64 hub_locate(const u64 *hoff, u64 lo, u64 n, u64 p) {
    u64 step = 1;
    while (lo + step < n && hoff[lo + step] <= p) {
        lo += step;
        step <<= 1;
    }
    u64 hi = lo + step < n ? lo + step : n;
    while (hi - lo > 1) {
        const u64 mid = lo + (hi - lo) / 2;
        if (hoff[mid] <= p) lo = mid;
        else hi = mid;
    }
    return lo;
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

// A counter alone in a 256-byte line: atomics on one line are serialized by
// the L2 channel that owns it, so counters that many waves hit at once each
// get a line.
template <typename T>
struct alignas(256) CounterLine {
    T v;
};

// Grid barrier for cooperative launches: generation counter + arrival count.
// Every wait is bounded; on timeout it raises *error and returns so that all
// waves still leave the kernel.
// Arrivals are counted in kBarrierGroups group counters, each in its own
// 256-byte line (block b counts in group b % groups); the last arrival of a
// group counts in the top counter, and the last group arrival opens the next
// generation. One counter for the whole grid would serialize every block's
// atomic on one address.
constexpr u32 kBarrierGroups = 16;
struct alignas(256) BarrierLine {
    u32 count;
};
struct alignas(256) GridBarrier {
    u32 count;   // groups arrived
    u32 gen;
    u32 error;   // set on timeout or by kernels that detect an inconsistency
    BarrierLine group[kBarrierGroups];
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
        const u32 g = __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const u32 ngroups = gridDim.x < kBarrierGroups ? gridDim.x : kBarrierGroups;
        const u32 grp = blockIdx.x % ngroups;
        const u32 members = (gridDim.x - grp + ngroups - 1) / ngroups;
        bool opened = false;
        const u32 arrived =
            __hip_atomic_fetch_add(&b->group[grp].count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == members - 1) {
            __hip_atomic_store(&b->group[grp].count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const u32 top = __hip_atomic_fetch_add(&b->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (top == ngroups - 1) {
                __hip_atomic_store(&b->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&b->gen, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                opened = true;
            }
        }
        if (!opened) {
            // poll with relaxed loads (an acquire load would invalidate this XCD's L2
            // on every poll); one acquire fence follows the wait
            const u64 t0 = wall_clock64();
            while (__hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(1);
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
