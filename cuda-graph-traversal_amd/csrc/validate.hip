// Graph500 BFS validation, the TEPS edge count, root selection and the
// runner's statistics.
// Reference: examples/graph500_runner.cu:14-74 (statistics, roots),
// :84-169 (validation kernels), :171-201 (edge count), :203-255 (driver).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace g500 {

// Confirmations are one bit per vertex (u32 words), set with atomic ORs; the
// entry points clear the n / 32 words before the init kernel.
__device__ __forceinline__ void confirm(u32 *bits, u64 v) { atomicOr(&bits[v >> 5], 1u << (v & 31)); }
__device__ __forceinline__ bool is_confirmed(const u32 *bits, u64 v) { return (bits[v >> 5] >> (v & 31)) & 1u; }

// Clears the confirmation words and the error counter.
__global__ __launch_bounds__(256) void validate_clear_kernel(u32 *__restrict__ confirmed, u64 nwords,
                                                             u64 *__restrict__ errors) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (u64)gridDim.x * blockDim.x)
        confirmed[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *errors = 0;
}

// Rules 1 and 2 (graph500_runner.cu:84-109).
__global__ __launch_bounds__(256) void validate_init_kernel(const i64 *__restrict__ pred, const i64 *__restrict__ dist,
                                                            u32 *__restrict__ confirmed, u64 *__restrict__ errors,
                                                            u64 n, u64 root) {
    u64 err = 0;
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x) {
        const i64 p = pred[v];
        if (p != -1 && (p < 0 || (u64)p >= n)) ++err;
        if (v == root) {
            if (p != (i64)root || dist[v] != 0) ++err;
            else confirm(confirmed, v);
        }
    }
    err = wave_sum(err);
    if (lane_id() == 0 && err) atomicAdd((unsigned long long *)errors, (unsigned long long)err);
}

__device__ __forceinline__ i64 wrap_inc(i64 x) { return (i64)((u64)x + 1u); }

// Rules 3-6 over every CSR entry, one wave per source row
// (graph500_runner.cu:111-153).
__global__ __launch_bounds__(256) void validate_edges_kernel(const u64 *__restrict__ rs, const u32 *__restrict__ col,
                                                             const i64 *__restrict__ pred, const i64 *__restrict__ dist,
                                                             u32 *__restrict__ confirmed, u64 *__restrict__ errors, u64 n) {
    const u32 lane = lane_id();
    const u64 wave = ((u64)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const u64 nw = (u64)gridDim.x * blockDim.x / kWave;
    u64 err = 0;
    for (u64 s = wave; s < n; s += nw) {
        const i64 sp = pred[s], sd = dist[s];
        const u64 b = rs[s], e = rs[s + 1];
        bool self_ok = false;
        for (u64 k = b + lane; k < e; k += kWave) {
            const u64 d = col[k];
            if (d == s) {
                ++err;
                continue;
            }
            const i64 dp = pred[d], dd = dist[d];
            if ((sd == -1 && dd != -1) || (sd != -1 && dd == -1)) ++err;
            if (sd >= 0 && dd >= 0 && (wrap_inc(sd) < dd || wrap_inc(dd) < sd)) ++err;
            if (sp == (i64)d) self_ok = true;
            if (dp == (i64)s) confirm(confirmed, d);
        }
        if (ballot(self_ok) && lane == 0) confirm(confirmed, s);
    }
    err = wave_sum(err);
    if (lane == 0 && err) atomicAdd((unsigned long long *)errors, (unsigned long long)err);
}

// Rule 7 (graph500_runner.cu:155-169).
__global__ __launch_bounds__(256) void validate_confirmed_kernel(const i64 *__restrict__ pred,
                                                                 const i64 *__restrict__ dist,
                                                                 const u32 *__restrict__ confirmed,
                                                                 u64 *__restrict__ errors, u64 n, u64 root) {
    u64 err = 0;
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x) {
        if (pred[v] == -1) continue;
        if (dist[v] == INT64_MAX || (v != root && !is_confirmed(confirmed, v))) ++err;
    }
    err = wave_sum(err);
    if (lane_id() == 0 && err) atomicAdd((unsigned long long *)errors, (unsigned long long)err);
}

// graph500_runner.cu:171-182: entries u <= v of every reached v.
__global__ __launch_bounds__(256) void edge_count_kernel(const u64 *__restrict__ rs, const u32 *__restrict__ col,
                                                         const i64 *__restrict__ pred, u64 n, u64 *__restrict__ out) {
    const u32 lane = lane_id();
    const u64 wave = ((u64)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const u64 nw = (u64)gridDim.x * blockDim.x / kWave;
    u64 c = 0;
    for (u64 v = wave; v < n; v += nw) {
        if (pred[v] == -1) continue;
        const u64 b = rs[v], e = rs[v + 1];
        for (u64 k = b + lane; k < e; k += kWave) c += col[k] <= v;
    }
    c = wave_sum(c);
    if (lane == 0 && c) atomicAdd((unsigned long long *)out, (unsigned long long)c);
}

// Small device scratch for counters shared by the validation entry points.
static u64 *counter_scratch() {
    Context &c = ctx();
    if (!c.d_counters) G500_HIP(hipMalloc(&c.d_counters, 64));
    return c.d_counters;
}

// Rules 1-7 on (pred, dist) with the confirmation words and error counter
// given; returns the error count (blocking).
static u64 validate_enqueue(const Oned_Graph *g, const i64 *pred, const i64 *dist, u64 root, u32 *confirmed,
                            u64 *errors) {
    Context &c = ctx();
    const u64 n = g->nlocalverts, nwords = (n + 31) / 32;
    const hipStream_t s = c.stream;
    const unsigned gv = grid_for(n, 256, (u64)c.num_cus * 32), ge = grid_for(n * kWave, 256, (u64)c.num_cus * 32);
    validate_clear_kernel<<<grid_for(nwords, 256, (u64)c.num_cus * 8), 256, 0, s>>>(confirmed, nwords, errors);
    validate_init_kernel<<<gv, 256, 0, s>>>(pred, dist, confirmed, errors, n, root);
    validate_edges_kernel<<<ge, 256, 0, s>>>(g->rowstarts, g->column, pred, dist, confirmed, errors, n);
    validate_confirmed_kernel<<<gv, 256, 0, s>>>(pred, dist, confirmed, errors, n, root);
    G500_HIP(hipGetLastError());
    u64 h = 0;
    g500_copy_to_host(&h, errors, sizeof(u64));
    return h;
}

// ------------------------------------------- scratch of the no-allocation path
// oned_graph_bfs_validate_noalloc keeps its scratch in this file-scope array,
// one copy per device, held until the process ends: 256 bytes of counters
// (the error count at 0, the depth loop's flag at 8) and then the
// confirmation bits of up to kNoAllocMaxVerts vertices.
constexpr u64 kNoAllocMaxVerts = 1ull << 27;
constexpr u64 kNoAllocCounters = 256;
constexpr u64 kNoAllocBytes = kNoAllocCounters + kNoAllocMaxVerts / 8;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) u8 scratch[kNoAllocBytes];

static std::mutex &scratch_mutex() {
    static std::mutex m;
    return m;
}
static u8 *g_scratch[kMaxDevices];

// The device address of `scratch` on the current device (resolved once per
// device; the host-side name is not a device address). nullptr after fail().
static u8 *noalloc_scratch() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        fail("no-allocation validation: no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(scratch_mutex());
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fail("no-allocation validation: hipGetSymbolAddress of the scratch array failed on device %d: %s", dev,
                 hipGetErrorString(e));
            return nullptr;
        }
        g_scratch[dev] = (u8 *)p;
    }
    return g_scratch[dev];
}

void resolve_noalloc_scratch() { (void)noalloc_scratch(); }

}  // namespace g500

using namespace g500;

extern "C" {

u64 oned_graph_edge_count(Worker_State *ws, const Oned_Graph *g, const i64 *pred) {
    (void)ws;
    Context &c = ctx();
    u64 *out = counter_scratch();
    if (failed()) return 0;
    const u64 n = g->nlocalverts;
    G500_HIP(hipMemsetAsync(out, 0, sizeof(u64), c.stream));
    edge_count_kernel<<<grid_for(n * kWave, 256, (u64)c.num_cus * 32), 256, 0, c.stream>>>(g->rowstarts, g->column,
                                                                                           pred, n, out);
    G500_HIP(hipGetLastError());
    u64 h = 0;
    g500_copy_to_host(&h, out, sizeof(u64));
    return h;
}

b32 oned_graph_bfs_validate(Worker_State *ws, const Oned_Graph *g, Bfs_State *bfs, u64 root, u64 *errors_out) {
    const i64 *dist = bfs_compute_dist_from_pred(ws, g, bfs, root);
    if (!dist || failed()) return 0;
    const u64 n = g->nlocalverts;
    const u64 mark = ws->gpu_arena.used;
    u32 *confirmed = (u32 *)arena_alloc(ws, (n + 31) / 32 * sizeof(u32), "validation confirmations");
    u64 *errors = counter_scratch();
    if (failed()) return 0;
    const u64 h = validate_enqueue(g, bfs->pred, dist, root, confirmed, errors);
    ws->gpu_arena.used = mark;
    if (failed()) return 0;
    if (errors_out) *errors_out = h;
    if (h) fprintf(stderr, "Validation Error: BFS validation found %llu error(s)\n", (unsigned long long)h);
    return h == 0;
}

b32 oned_graph_bfs_validate_noalloc(const Oned_Graph *g, const i64 *pred, u64 root, i64 *depth, u64 *errors_out) {
    Context &c = ctx();
    if (!c.stream) {
        fail("oned_graph_bfs_validate_noalloc: call worker_init first");
        return 0;
    }
    const u64 n = g->nlocalverts;
    if (n > kNoAllocMaxVerts) {
        fail("oned_graph_bfs_validate_noalloc: %llu vertices exceed the no-allocation maximum of %llu "
             "(use oned_graph_bfs_validate)",
             (unsigned long long)n, (unsigned long long)kNoAllocMaxVerts);
        return 0;
    }
    if (root >= g->nglobalverts) {
        fail("BFS root %llu is outside graph vertex range [0, %llu)", (unsigned long long)root,
             (unsigned long long)g->nglobalverts);
        return 0;
    }
    if (!pred || !depth) {
        fail("oned_graph_bfs_validate_noalloc: pred and depth must be device arrays of %llu entries",
             (unsigned long long)n);
        return 0;
    }
    u8 *base = noalloc_scratch();
    if (!base || failed()) return 0;
    u64 *errors = (u64 *)base;
    u32 *changed = (u32 *)(base + 8);
    u32 *confirmed = (u32 *)(base + kNoAllocCounters);
    depth_from_pred(g, pred, depth, root, changed);
    if (failed()) return 0;
    const u64 h = validate_enqueue(g, pred, depth, root, confirmed, errors);
    if (failed()) return 0;
    if (errors_out) *errors_out = h;
    if (h) fprintf(stderr, "Validation Error: BFS validation found %llu error(s)\n", (unsigned long long)h);
    return h == 0;
}

i32 generate_bfs_roots(Worker_State *ws, const Oned_Graph *g, u64 seed1, u64 seed2, u64 *roots, i32 max_roots) {
    const u64 n = g->nglobalverts;
    u64 counter = 0;
    i32 count = 0;
    for (; count < max_roots; ++count) {
        u64 root = 0;
        bool found = false;
        for (;;) {
            f64 d[2];
            make_random_numbers(2, seed1, seed2, (i64)counter, d);
            root = (u64)((d[0] + d[1]) * (f64)n) % n;
            counter += 2;
            if (counter > 2 * n) break;
            bool dup = false;
            for (i32 i = 0; i < count; ++i)
                if (roots[i] == root) {
                    dup = true;
                    break;
                }
            if (dup) continue;
            if (oned_graph_is_vertex_isolated(ws, g, root)) continue;
            if (failed()) return count;
            found = true;
            break;
        }
        if (!found) break;
        roots[count] = root;
    }
    return count;
}

void get_statistics(const f64 *x, i32 n, f64 r[7]) {
    f64 mean = 0.0;
    for (i32 i = 0; i < n; ++i) mean += x[i];
    mean /= n;
    f64 var = 0.0;
    for (i32 i = 0; i < n; ++i) var += (x[i] - mean) * (x[i] - mean);
    var /= n - 1;
    std::vector<f64> s(x, x + n);
    std::sort(s.begin(), s.end());
    r[0] = s[0];
    r[1] = (s[(n - 1) / 4] + s[n / 4]) * 0.5;
    r[2] = (s[(n - 1) / 2] + s[n / 2]) * 0.5;
    r[3] = (s[n - 1 - (n - 1) / 4] + s[n - 1 - n / 4]) * 0.5;
    r[4] = s[n - 1];
    r[5] = mean;
    r[6] = std::sqrt(var);
}

}  // extern "C"
