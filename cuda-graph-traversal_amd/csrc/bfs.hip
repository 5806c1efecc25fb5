// Direction-optimizing BFS as one persistent cooperative kernel per root, and
// the depth-from-parent derivation used by validation.
// Reference: src/traversal.cu:75-207 (kernels), :412-568 (state, run, depth).
//
// The reference runs a top-down level loop in one persistent kernel with one
// CAS per examined edge. Here each level is either
//   top-down: frontier queue, one wave per vertex (lanes over its row), and
//             high-degree "hub" vertices in a flattened edge work space whose
//             fixed-size chunks are spread over the whole grid; a plain
//             load of the visited bitmap filters most edges before the atomic
//             OR that claims a vertex (exactly one winner writes pred);
//   bottom-up: every unvisited vertex scans its row until it meets a frontier
//             vertex in the frontier bitmap; a wave owns 64 consecutive
//             vertices (one bitmap word), so no atomics are needed.
// The switch follows Beamer's rule: top-down -> bottom-up when the frontier's
// edges exceed the unexplored edges / alpha, bottom-up -> top-down when the
// frontier is below n / beta and shrinking. All blocks evaluate the rule on
// the same counters after the level barrier, so they agree without an extra
// barrier. Frontier counters rotate over three slots (read / write / clear).
#include "common.hpp"

namespace g500 {

constexpr int kMaxLevelLog = 256;

struct BfsCtl {
    GridBarrier bar;
    // counters that many waves add to sit on lines of their own, so that
    // their atomics do not queue behind each other's
    alignas(256) u32 qlen[3];
    alignas(256) u64 hcnt[3];   // packed hub counters (count << 38 | padded edge total)
    alignas(256) u64 fcount[3];
    alignas(256) u64 fedges[3];
    u32 levels;
    u32 depth_changed;
    alignas(256) i32 mode_log[kMaxLevelLog];
    u64 size_log[kMaxLevelLog];
    u64 tick_log[kMaxLevelLog];   // constant-clock ticks of each level seen by block 0
    alignas(256) BlockTrace trace[kMaxTraceBlocks];
};

__device__ __forceinline__ void bfs_sync(BfsCtl *ctl, u32 phase, u32 level) {
    grid_sync(&ctl->bar, &ctl->bar.error);
    if (threadIdx.x == 0 && blockIdx.x < kMaxTraceBlocks) {
        BlockTrace *t = &ctl->trace[blockIdx.x];
        st_relaxed(&t->barriers, ld_relaxed(&t->barriers) + 1);
        st_relaxed(&t->phase, phase);
        st_relaxed(&t->a, level);
    }
}

struct BfsArgs {
    u64 n, nwords, m_total;
    const u64 *rs;      // traversal view
    const u32 *col;
    const u32 *perm, *iperm;
    i64 *pred;          // original ids
    u64 *visited;
    u64 *fb[2];
    u32 *q[2];
    u32 *hubs[2];
    u64 *hoff[2];
    u32 *hlen[2];
    BfsCtl *ctl;
    u64 root;
    f32 alpha, beta;
    u32 hub_degree;
};

__device__ __forceinline__ u64 ld64(const u64 *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Claims v for parent u when this lane is active; appends winners to the next
// queue (or the hub queue) with one atomic per wave. All lanes must call it.
// A wave's top-down queue appends are staged in LDS (sq, sn entries) and
// written out in runs: one atomic on the queue counter per run instead of one
// per 64 edges with a new vertex.
constexpr u32 kTdStage = 256;
__device__ __forceinline__ void td_flush(u32 *qn, BfsCtl *ctl, int nxt, volatile u32 *sq, u32 &sn, u64 cap) {
    if (!sn) return;
    u32 base = 0;
    if (lane_id() == 0) base = atomicAdd(&ctl->qlen[nxt], sn);
    base = __shfl(base, 0);
    for (u32 i = lane_id(); i < sn; i += kWave)
        if ((u64)base + i < cap) qn[base + i] = sq[i];
    sn = 0;
}

__device__ __forceinline__ void td_visit(const BfsArgs &a, u32 u, u32 v, bool active, u32 *qn, int nb,
                                         BfsCtl *ctl, int nxt, u32 &cnt, u64 &edges, volatile u32 *sq, u32 &sn) {
    bool won = false;
    if (active) {
        const u64 bit = 1ull << (v & 63);
        if (!(a.visited[v >> 6] & bit)) {
            const u64 old = atomicOr((unsigned long long *)&a.visited[v >> 6], (unsigned long long)bit);
            won = !(old & bit);
        }
    }
    u32 deg = 0;
    if (won) {
        a.pred[orig_id(a.iperm, v)] = orig_id(a.iperm, u);
        deg = (u32)(a.rs[v + 1] - a.rs[v]);
        ++cnt;
        edges += deg;
    }
    const bool hub = won && deg > a.hub_degree;
    const u64 m = ballot(won && !hub);
    if (m) {
        const u32 k = (u32)__popcll(m);
        if (sn + k > kTdStage) td_flush(qn, ctl, nxt, sq, sn, a.n);
        if (won && !hub) sq[sn + mask_rank(m)] = v;
        sn += k;
    }
    hub_append(&ctl->hcnt[nxt], a.hubs[nb], a.hoff[nb], a.hlen[nb], hub, v, deg, a.n, &ctl->bar.error);
}

__global__ __launch_bounds__(512) void bfs_kernel(BfsArgs a) {
    BfsCtl *ctl = a.ctl;
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const u64 gthreads = (u64)gridDim.x * blockDim.x;
    const u32 lane = lane_id();
    const u64 wave_gid = gtid / kWave, nwaves = gthreads / kWave;
    __shared__ u32 tq[512 / kWave][kTdStage];
    volatile u32 *sq = tq[threadIdx.x / kWave];
    u32 sn = 0;  // staged queue entries of this wave (wave-uniform)

    if (gtid == 0) {  // level 0: the root
        const u64 r = trav_id(a.perm, a.root);
        a.pred[a.root] = (i64)a.root;
        atomicOr((unsigned long long *)&a.visited[r >> 6], 1ull << (r & 63));
        const u32 deg = (u32)(a.rs[r + 1] - a.rs[r]);
        if (deg > a.hub_degree) {
            a.hubs[0][0] = (u32)r;
            a.hoff[0][0] = 0;
            a.hlen[0][0] = deg;
            st_relaxed(&ctl->hcnt[0], (1ull << kHubCountShift) | ((deg + kHubChunk - 1) / kHubChunk * kHubChunk));
        } else {
            a.q[0][0] = (u32)r;
            st_relaxed(&ctl->qlen[0], 1u);
        }
        st_relaxed(&ctl->fcount[0], 1ull);
        st_relaxed(&ctl->fedges[0], (u64)deg);
    }
    bfs_sync(ctl, 1, 0);

    bool bottom_up = false;
    u64 m_explored = 0, prev_nf = 0;
    u64 t_prev = gtid == 0 ? wall_clock64() : 0;
    for (u32 L = 0;; ++L) {
        const int cur = L % 3, nxt = (L + 1) % 3, clr = (L + 2) % 3;
        const u64 nf = ld64(&ctl->fcount[cur]);
        const u64 mf = ld64(&ctl->fedges[cur]);
        if (nf == 0 || ld_relaxed(&ctl->bar.error) || L >= 0x7FFFFFFFu) break;
        const u64 m_unexplored = a.m_total > m_explored + mf ? a.m_total - m_explored - mf : 0;
        bool next_bu = bottom_up;
        if (!bottom_up) next_bu = m_unexplored > 0 && (f32)mf > (f32)m_unexplored / a.alpha;
        else next_bu = !((f32)nf < (f32)a.n / a.beta && nf < prev_nf);
        m_explored += mf;
        prev_nf = nf;
        if (gtid == 0) {
            st_relaxed(&ctl->qlen[clr], 0u);
            st_relaxed(&ctl->hcnt[clr], 0ull);
            st_relaxed(&ctl->fcount[clr], 0ull);
            st_relaxed(&ctl->fedges[clr], 0ull);
            if (L < kMaxLevelLog) {
                st_relaxed(&ctl->mode_log[L], (i32)next_bu);
                st_relaxed(&ctl->size_log[L], nf);
            }
            st_relaxed(&ctl->levels, L + 1);
        }
        u32 *qc = a.q[L & 1], *qn = a.q[(L + 1) & 1];
        u64 *fbc = a.fb[L & 1], *fbn = a.fb[(L + 1) & 1];

        if (next_bu && !bottom_up) {  // frontier queue -> bitmap
            for (u64 w = gtid; w < a.nwords; w += gthreads) fbc[w] = 0;
            bfs_sync(ctl, 2, L);
            const u32 nq = ld_relaxed(&ctl->qlen[cur]);
            const u64 nh = hub_count(ld_relaxed(&ctl->hcnt[cur]));
            const u32 *hc = a.hubs[L & 1];
            for (u64 i = gtid; i < (u64)nq + nh; i += gthreads) {
                const u32 v = i < nq ? qc[i] : hc[i - nq];
                atomicOr((unsigned long long *)&fbc[v >> 6], 1ull << (v & 63));
            }
            bfs_sync(ctl, 3, L);
        } else if (!next_bu && bottom_up) {  // frontier bitmap -> queues
            for (u64 w0 = wave_gid * kWave; w0 < a.nwords; w0 += nwaves * kWave) {
                const u64 w = w0 + lane;
                u64 bits = w < a.nwords ? fbc[w] : 0;
                // each lane emits the vertices of its word; appends stay wave-converged
                for (;;) {
                    const bool has = bits != 0;
                    if (!ballot(has)) break;
                    u32 v = 0, deg = 0;
                    if (has) {
                        const int b = __ffsll((unsigned long long)bits) - 1;
                        bits &= bits - 1;
                        v = (u32)(w * 64 + b);
                        deg = (u32)(a.rs[v + 1] - a.rs[v]);
                    }
                    const bool hub = has && deg > a.hub_degree;
                    const u32 s1 = wave_append(&ctl->qlen[cur], has && !hub);
                    if (has && !hub) qc[s1] = v;
                    hub_append(&ctl->hcnt[cur], a.hubs[L & 1], a.hoff[L & 1], a.hlen[L & 1], hub, v, deg, a.n,
                               &ctl->bar.error);
                }
            }
            bfs_sync(ctl, 4, L);
        }
        bottom_up = next_bu;

        u32 cnt = 0;
        u64 edges = 0;
        if (!bottom_up) {
            const u32 nq = ld_relaxed(&ctl->qlen[cur]);
            const u64 hp = ld_relaxed(&ctl->hcnt[cur]);
            const u64 nh = hub_count(hp), ht = hub_total(hp);
            const u32 *hc = a.hubs[L & 1];
            const u64 *ho = a.hoff[L & 1];
            const u32 *hl = a.hlen[L & 1];
            u64 hi = 0;
            // hubs: chunks of the flattened hub edge space over the whole grid
            for (u64 c = wave_gid; c * kHubChunk < ht; c += nwaves) {
                const u64 p = c * kHubChunk;
                hi = hub_locate(ho, hi, nh, p);
                const u32 u = hc[hi];
                const u64 b = a.rs[u] + (p - ho[hi]);
                const u64 e = a.rs[u] + hl[hi];
#pragma unroll
                for (u64 j = 0; j < kHubChunk; j += kWave) {
                    const u64 k = b + j + lane;
                    const bool act = k < e;
                    td_visit(a, u, act ? a.col[k] : 0, act, qn, (int)((L + 1) & 1), ctl, nxt, cnt, edges, sq, sn);
                }
            }
            for (u64 i = wave_gid; i < nq; i += nwaves) {  // one wave per frontier vertex
                const u32 u = qc[i];
                const u64 b = a.rs[u], e = a.rs[u + 1];
                for (u64 k0 = b; k0 < e; k0 += kWave) {
                    const u64 k = k0 + lane;
                    const bool act = k < e;
                    td_visit(a, u, act ? a.col[k] : 0, act, qn, (int)((L + 1) & 1), ctl, nxt, cnt, edges, sq, sn);
                }
            }
            td_flush(qn, ctl, nxt, sq, sn, a.n);
        } else {
            for (u64 w = wave_gid; w < a.nwords; w += nwaves) {
                const u64 vis = a.visited[w];
                const u64 v = w * 64 + lane;
                bool found = false;
                if (v < a.n && !((vis >> lane) & 1)) {
                    const u64 b = a.rs[v], e = a.rs[v + 1];
                    for (u64 k = b; k < e; ++k) {
                        const u32 u = a.col[k];
                        if ((fbc[u >> 6] >> (u & 63)) & 1) {
                            a.pred[orig_id(a.iperm, (u32)v)] = orig_id(a.iperm, u);
                            found = true;
                            edges += e - b;
                            ++cnt;
                            break;
                        }
                    }
                }
                const u64 fm = ballot(found);
                if (lane == 0) {
                    fbn[w] = fm;
                    if (fm) a.visited[w] = vis | fm;
                }
            }
        }
        const u32 wc = wave_sum(cnt);
        const u64 we = wave_sum(edges);
        if (lane == 0 && wc) {
            atomicAdd((unsigned long long *)&ctl->fcount[nxt], (unsigned long long)wc);
            atomicAdd((unsigned long long *)&ctl->fedges[nxt], (unsigned long long)we);
        }
        bfs_sync(ctl, bottom_up ? 6 : 5, L);
        if (gtid == 0 && L < kMaxLevelLog) {
            const u64 t = wall_clock64();
            st_relaxed(&ctl->tick_log[L], t - t_prev);
            t_prev = t;
        }
    }
}

// ------------------------------------------------- depth from the parent tree
constexpr i64 kUnresolved = INT64_MAX;      // reached, no depth (reference value)
constexpr i64 kPending = INT64_MAX - 1;     // parent edge exists, depth not yet known

// traversal.cu:75-86 plus the parent-edge test of :97-104: a vertex can only
// get a depth through an entry pred[v] in its row (the CSR is symmetric, so
// this is the entry v in the row of pred[v] that the reference looks for).
__global__ __launch_bounds__(256) void depth_init_kernel(const u64 *__restrict__ rs, const u32 *__restrict__ col,
                                                         const i64 *__restrict__ pred, i64 *__restrict__ dist,
                                                         u64 n, u64 root) {
    const u32 lane = lane_id();
    const u64 wave = ((u64)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const u64 nw = (u64)gridDim.x * blockDim.x / kWave;
    for (u64 v = wave; v < n; v += nw) {
        const i64 p = pred[v];
        i64 d;
        if (v == root) d = 0;
        else if (p == -1) d = -1;
        else if (p < 0 || (u64)p >= n) d = kUnresolved;
        else {
            const u64 b = rs[v], e = rs[v + 1];
            bool hit = false;
            for (u64 k0 = b; k0 < e && !hit; k0 += kWave) {
                const u64 k = k0 + lane;
                hit = ballot(k < e && col[k] == (u32)p) != 0;
            }
            d = hit ? kPending : kUnresolved;
        }
        if (lane == 0) dist[v] = d;
    }
}

__global__ __launch_bounds__(256) void depth_step_kernel(const i64 *__restrict__ pred, i64 *__restrict__ dist,
                                                         u64 n, u32 *__restrict__ changed) {
    u32 any = 0;
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x) {
        if (dist[v] != kPending) continue;
        const i64 dp = dist[pred[v]];
        if (dp >= 0 && dp < kPending) {
            dist[v] = dp + 1;
            any = 1;
        }
    }
    if (__any(any) && lane_id() == 0) atomicOr(changed, 1u);
}

__global__ __launch_bounds__(256) void depth_final_kernel(i64 *__restrict__ dist, u64 n) {
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x)
        if (dist[v] == kPending) dist[v] = kUnresolved;
}

void depth_from_pred(const Oned_Graph *g, const i64 *pred, i64 *dist, u64 root, u32 *changed) {
    Context &c = ctx();
    const u64 n = g->nlocalverts;
    const hipStream_t s = c.stream;
    depth_init_kernel<<<grid_for(n * kWave, 256, (u64)c.num_cus * 32), 256, 0, s>>>(g->rowstarts, g->column, pred,
                                                                                   dist, n, root);
    for (u64 iter = 0; iter <= n; ++iter) {
        G500_HIP(hipMemsetAsync(changed, 0, sizeof(u32), s));
        depth_step_kernel<<<grid_for(n, 256, (u64)c.num_cus * 32), 256, 0, s>>>(pred, dist, n, changed);
        u32 h = 0;
        G500_HIP(hipMemcpyAsync(&h, changed, sizeof(u32), hipMemcpyDeviceToHost, s));
        sync();
        if (!h || failed()) break;
    }
    depth_final_kernel<<<grid_for(n, 256, (u64)c.num_cus * 32), 256, 0, s>>>(dist, n);
    G500_HIP(hipGetLastError());
    sync();
}

}  // namespace g500

using namespace g500;

extern "C" {

Bfs_State oned_graph_bfs_create(Worker_State *ws, const Oned_Graph *g) {
    Bfs_State b;
    memset(&b, 0, sizeof(b));
    const u64 n = g->nlocalverts, nw = (n + 63) / 64;
    b.max_nlocalverts = n;
    b.pred = (i64 *)arena_alloc(ws, n * sizeof(i64), "BFS pred");
    b.visited = (u64 *)arena_alloc(ws, nw * sizeof(u64), "BFS visited");
    for (int i = 0; i < 2; ++i) {
        b.frontier[i] = (u32 *)arena_alloc(ws, n * sizeof(u32), "BFS queue");
        b.frontier_bits[i] = (u64 *)arena_alloc(ws, nw * sizeof(u64), "BFS frontier bitmap");
    }
    b.hubs = (u32 *)arena_alloc(ws, 2 * n * sizeof(u32), "BFS hub queues");
    b.hub_offsets = (u64 *)arena_alloc(ws, 2 * n * sizeof(u64), "BFS hub offsets");
    b.hub_lengths = (u32 *)arena_alloc(ws, 2 * n * sizeof(u32), "BFS hub lengths");
    b.ctl = arena_alloc(ws, sizeof(BfsCtl), "BFS control");
    if (failed()) return b;
    oned_graph_bfs_clear(ws, &b);
    return b;
}

void oned_graph_bfs_clear(Worker_State *ws, Bfs_State *bfs) {
    (void)ws;
    const u64 n = bfs->max_nlocalverts, nw = (n + 63) / 64;
    const hipStream_t s = ctx().stream;
    G500_HIP(hipMemsetAsync(bfs->pred, 0xff, n * sizeof(i64), s));
    if (bfs->dist) G500_HIP(hipMemsetAsync(bfs->dist, 0xff, n * sizeof(i64), s));
    G500_HIP(hipMemsetAsync(bfs->visited, 0, nw * sizeof(u64), s));
    G500_HIP(hipMemsetAsync(bfs->ctl, 0, sizeof(BfsCtl), s));
    sync();
}

void oned_graph_bfs_run(Worker_State *ws, const Oned_Graph *g, Bfs_State *bfs, u64 root) {
    (void)ws;
    Context &c = ctx();
    if (root >= g->nglobalverts) {
        fail("BFS root %llu is outside graph vertex range [0, %llu)", (unsigned long long)root,
             (unsigned long long)g->nglobalverts);
        return;
    }
    BfsArgs a;
    a.n = g->nlocalverts;
    a.nwords = (a.n + 63) / 64;
    a.m_total = g->nlocaledges;
    const TraversalView tv = traversal_view(g);
    a.rs = tv.rs;
    a.col = tv.col;
    a.perm = tv.perm;
    a.iperm = tv.iperm;
    a.pred = bfs->pred;
    a.visited = bfs->visited;
    a.fb[0] = bfs->frontier_bits[0];
    a.fb[1] = bfs->frontier_bits[1];
    a.q[0] = bfs->frontier[0];
    a.q[1] = bfs->frontier[1];
    for (int i = 0; i < 2; ++i) {
        a.hubs[i] = bfs->hubs + i * a.n;
        a.hoff[i] = bfs->hub_offsets + i * a.n;
        a.hlen[i] = bfs->hub_lengths + i * a.n;
    }
    a.ctl = (BfsCtl *)bfs->ctl;
    a.root = root;
    a.alpha = (f32)c.bfs_alpha;
    a.beta = (f32)c.bfs_beta;
    a.hub_degree = (u32)c.bfs_hub_degree;
    const int block = ctx().bfs_block;
    int grid = cooperative_blocks((const void *)bfs_kernel, block, 0);
    if (failed()) return;
    void *args[] = {&a};
    G500_HIP(hipMemsetAsync(bfs->ctl, 0, sizeof(BfsCtl), c.stream));
    timer_start();
    launch_persistent((const void *)bfs_kernel, grid, block, args);
    timer_stop();
    static BfsCtl h;
    g500_copy_to_host(&h, bfs->ctl, sizeof(h));
    if (h.bar.error) {
        u32 lo = 0xFFFFFFFFu, hi = 0;
        int at_lo = 0, at_hi = 0;
        for (int b = 0; b < MIN(grid, kMaxTraceBlocks); ++b) {
            if (h.trace[b].barriers < lo) { lo = h.trace[b].barriers; at_lo = b; }
            if (h.trace[b].barriers > hi) { hi = h.trace[b].barriers; at_hi = b; }
        }
        fail("BFS kernel error %u (1 barrier timeout, 2 hub queue overflow); %d blocks; barriers passed min %u "
             "(block %d phase %u level %u) max %u (block %d phase %u level %u)",
             h.bar.error, grid, lo, at_lo, h.trace[at_lo].phase, h.trace[at_lo].a, hi, at_hi, h.trace[at_hi].phase,
             h.trace[at_hi].a);
    }
}

i64 *bfs_compute_dist_from_pred(Worker_State *ws, const Oned_Graph *g, Bfs_State *bfs, u64 root) {
    if (root >= g->nglobalverts) {
        fail("BFS root %llu is outside graph vertex range [0, %llu)", (unsigned long long)root,
             (unsigned long long)g->nglobalverts);
        return nullptr;
    }
    const u64 n = g->nlocalverts;
    if (!bfs->dist) bfs->dist = (i64 *)arena_alloc(ws, n * sizeof(i64), "BFS depth");
    if (failed()) return nullptr;
    depth_from_pred(g, bfs->pred, bfs->dist, root, &((BfsCtl *)bfs->ctl)->depth_changed);
    return failed() ? nullptr : bfs->dist;
}

void oned_graph_bfs_destroy(Bfs_State *bfs) {
    if (bfs) memset(bfs, 0, sizeof(*bfs));
}

i32 g500_bfs_level_log(const Bfs_State *bfs, i32 *modes, u64 *frontier, f64 *ms, i32 max_levels) {
    static BfsCtl h;
    g500_copy_to_host(&h, bfs->ctl, sizeof(h));
    i32 L = (i32)MIN((u32)h.levels, (u32)kMaxLevelLog);
    for (i32 i = 0; i < L && i < max_levels; ++i) {
        if (modes) modes[i] = h.mode_log[i];
        if (frontier) frontier[i] = h.size_log[i];
        if (ms) ms[i] = (f64)h.tick_log[i] / ctx().wall_clock_khz;
    }
    return L;
}

}  // extern "C"
