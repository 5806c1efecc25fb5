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

// ------------------------------------------------------------ certificates
// Depth certificate, one wave per vertex: d is the BFS depth array of root
// iff d[root] = 0, every reached v != root has a neighbour at d[v] - 1, and
// every entry of a reached u joins a reached vertex at most one level away
// (values below -1 count as errors too).
__global__ __launch_bounds__(256) void depth_check_kernel(const u64 *__restrict__ rs, const u32 *__restrict__ col,
                                                          const i64 *__restrict__ depth, u64 n, u64 root,
                                                          u64 *__restrict__ errors) {
    const u32 lane = lane_id();
    const u64 wave = ((u64)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const u64 nw = (u64)gridDim.x * blockDim.x / kWave;
    u64 err = 0;
    for (u64 u = wave; u < n; u += nw) {
        const i64 du = depth[u];
        if (du < -1) {
            err += lane == 0;
            continue;
        }
        if (du == -1) continue;
        bool parent = false;
        for (u64 k = rs[u] + lane; k < rs[u + 1]; k += kWave) {
            const i64 dv = depth[col[k]];
            if (dv < 0 || dv > du + 1 || dv < du - 1) ++err;
            parent |= dv == du - 1;
        }
        const bool any_parent = ballot(parent) != 0;  // u is the same in every lane
        if (lane == 0 && u != root && !any_parent) ++err;
    }
    err = wave_sum(err);
    if (lane == 0 && err) atomicAdd((unsigned long long *)errors, (unsigned long long)err);
}

// SSSP certificate over the full rows, one wave per vertex: the root has
// distance 0 and parent root; an unreached vertex has distance -1.0f and
// parent -1; every entry (u, v, w) of a reached u has v reached with
// d[v] <= fl(d[u] + w); every reached u != root has an in-range reached parent
// p and an entry (u, p, w) with fl(d[p] + w) == d[u]. flat[0] counts the
// tight parent entries that do not lower the distance (d[p] == d[u]): only
// those can close a parent cycle.
__global__ __launch_bounds__(256) void sssp_check_kernel(const u64 *__restrict__ rs, const u32 *__restrict__ col,
                                                         const f32 *__restrict__ w, const f32 *__restrict__ dist,
                                                         const i64 *__restrict__ pred, u64 n, u64 root,
                                                         u64 *__restrict__ errors, u64 *__restrict__ flat) {
    const u32 lane = lane_id();
    const u64 wave = ((u64)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const u64 nw = (u64)gridDim.x * blockDim.x / kWave;
    u64 err = 0, nflat = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0 && (dist[root] != 0.0f || pred[root] != (i64)root)) ++err;
    for (u64 u = wave; u < n; u += nw) {
        const f32 du = dist[u];
        const i64 p = pred[u];
        if (!(du >= 0.0f)) {
            if (lane == 0 && (__float_as_uint(du) != __float_as_uint(-1.0f) || p != -1)) ++err;
            continue;
        }
        const bool has_p = u != root && p >= 0 && (u64)p < n;
        const f32 dp = has_p ? dist[p] : -1.0f;
        const bool p_ok = has_p && dp >= 0.0f;
        bool tight = false;
        for (u64 k = rs[u] + lane; k < rs[u + 1]; k += kWave) {
            const u32 v = col[k];
            const f32 wk = w[k];
            const f32 dv = dist[v];
            const f32 cand = __fadd_rn(du, wk);
            if (!(dv >= 0.0f) || dv > cand) ++err;
            if (p_ok && v == (u32)p && __fadd_rn(dp, wk) == du) tight = true;
        }
        const bool any_tight = ballot(tight) != 0;
        if (lane == 0 && u != root) {
            if (!p_ok || !any_tight) ++err;
            else if (dp == du) ++nflat;
        }
    }
    err = wave_sum(err);
    nflat = wave_sum(nflat);
    if (lane == 0 && err) atomicAdd((unsigned long long *)errors, (unsigned long long)err);
    if (lane == 0 && nflat) atomicAdd((unsigned long long *)flat, (unsigned long long)nflat);
}

// Parent chains by pointer jumping: anc[v] starts at pred[v] (root: root; no
// valid parent: the sentinel n) and is replaced by anc[anc[v]] until every
// chain of up to n steps has been followed; a reached vertex whose chain does
// not end at the root is one error.
__global__ __launch_bounds__(256) void chain_init_kernel(const i64 *__restrict__ pred, u64 *__restrict__ anc, u64 n,
                                                         u64 root) {
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x) {
        const i64 p = pred[v];
        anc[v] = v == root ? root : (p >= 0 && (u64)p < n ? (u64)p : n);
    }
}

__global__ __launch_bounds__(256) void chain_jump_kernel(const u64 *__restrict__ src, u64 *__restrict__ dst, u64 n) {
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x) {
        const u64 a = src[v];
        dst[v] = a < n ? src[a] : n;
    }
}

__global__ __launch_bounds__(256) void chain_count_kernel(const u64 *__restrict__ anc, const f32 *__restrict__ dist,
                                                          u64 n, u64 root, u64 *__restrict__ errors) {
    u64 err = 0;
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x)
        err += dist[v] >= 0.0f && anc[v] != root;
    err = wave_sum(err);
    if (lane_id() == 0 && err) atomicAdd((unsigned long long *)errors, (unsigned long long)err);
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

u64 oned_graph_bfs_depth_check(Worker_State *ws, const Oned_Graph *g, const i64 *depth, u64 root) {
    (void)ws;
    Context &c = ctx();
    if (root >= g->nglobalverts) {
        fail("BFS root %llu is outside graph vertex range [0, %llu)", (unsigned long long)root,
             (unsigned long long)g->nglobalverts);
        return 0;
    }
    u64 *errors = counter_scratch();
    if (failed()) return 0;
    const u64 n = g->nlocalverts;
    G500_HIP(hipMemsetAsync(errors, 0, sizeof(u64), c.stream));
    depth_check_kernel<<<grid_for(n * kWave, 256, (u64)c.num_cus * 32), 256, 0, c.stream>>>(g->rowstarts, g->column,
                                                                                          depth, n, root, errors);
    G500_HIP(hipGetLastError());
    u64 err = 0;
    g500_copy_to_host(&err, errors, sizeof(u64));
    i64 dr = 0;
    g500_copy_to_host(&dr, depth + root, sizeof(i64));
    return err + (dr != 0);
}

u64 oned_graph_sssp_check(Worker_State *ws, const Oned_Graph *g, const f32 *dist, const i64 *pred, u64 root) {
    Context &c = ctx();
    if (root >= g->nglobalverts) {
        fail("SSSP root %llu is outside graph vertex range [0, %llu)", (unsigned long long)root,
             (unsigned long long)g->nglobalverts);
        return 0;
    }
    u64 *cnt = counter_scratch();
    if (failed()) return 0;
    const u64 n = g->nlocalverts;
    const hipStream_t s = c.stream;
    G500_HIP(hipMemsetAsync(cnt, 0, 2 * sizeof(u64), s));
    sssp_check_kernel<<<grid_for(n * kWave, 256, (u64)c.num_cus * 32), 256, 0, s>>>(
        g->rowstarts, g->column, g->weights, dist, pred, n, root, cnt, cnt + 1);
    G500_HIP(hipGetLastError());
    u64 h[2] = {0, 0};
    g500_copy_to_host(h, cnt, sizeof(h));
    if (failed() || (h[0] == 0 && h[1] == 0)) return h[0];
    // Errors or parent entries that keep the distance: follow the chains.
    // (Without either, distances fall strictly along every parent entry, so
    // every chain ends at the root.)
    const u64 mark = ws->gpu_arena.used;
    u64 *a = (u64 *)arena_alloc(ws, n * sizeof(u64), "chain ancestors");
    u64 *b = (u64 *)arena_alloc(ws, n * sizeof(u64), "chain ancestors");
    if (failed()) return h[0];
    const unsigned gv = grid_for(n, 256, (u64)c.num_cus * 32);
    chain_init_kernel<<<gv, 256, 0, s>>>(pred, a, n, root);
    for (u64 span = 1; span < n + 1; span *= 2) {
        chain_jump_kernel<<<gv, 256, 0, s>>>(a, b, n);
        std::swap(a, b);
    }
    G500_HIP(hipMemsetAsync(cnt, 0, sizeof(u64), s));
    chain_count_kernel<<<gv, 256, 0, s>>>(a, dist, n, root, cnt);
    G500_HIP(hipGetLastError());
    u64 chains = 0;
    g500_copy_to_host(&chains, cnt, sizeof(u64));
    ws->gpu_arena.used = mark;
    return h[0] + chains;
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
