// Batched multi-source BFS: up to 64 roots in one persistent cooperative
// kernel. No counterpart in the reference, whose runner traverses one root
// at a time (examples/graph500_runner.cu:306-336); each root's depths equal
// those of the single-root BFS (traversal.cu:139-207 semantics).
//
// State per vertex: seen (bit r = root r reached it). A level's frontier is a
// run of discovery-log records (vertex, mask of roots that reached it first at
// that level). Levels run in one of two directions:
//   top-down: a wave per record (hub records over the whole grid in 256-edge
//     chunks); for each neighbor u, new = mask & ~seen[u] is OR-ed into
//     next[u]; the first OR into next[u] in a level appends u to a touched
//     list; a compaction pass turns touched vertices into the next level's
//     records and folds them into seen;
//   bottom-up: every vertex that still misses roots of the live set ORs the
//     frontier masks of its neighbors (dense array), stopping as soon as all
//     missing live roots are found; each vertex is owned by one lane, so
//     there are no atomics on seen.
// The switch follows Beamer's rule on the records' edge counts. Counters of a
// level rotate over three slots (written by the level's producer, read by its
// consumer, cleared two levels ahead), as in the single-root kernel. Records
// produced for one level are contiguous in the log, so a level's record range
// is the previous end plus that slot's record count. The hub lists of two
// consecutive levels live in the two halves of the hub arrays (level parity).
#include "common.hpp"

namespace g500 {

constexpr int kBatchLevelLog = 256;

struct BatchCtl {
    GridBarrier bar;
    alignas(256) u64 nrec;         // discovery-log records so far
    u64 rcount[3];                 // records produced for a level, per slot
    u32 ntouched[3];               // top-down touched vertices, per slot
    u64 hub[3];                    // packed hub counters (record indices), per slot
    u64 fedges[3];                 // sum of degrees of a level's records, per slot
    u64 alive[3];                  // OR of a level's masks, per slot
    u32 levels;
    u32 pad;
    u64 roots[64];
    alignas(256) i32 mode_log[kBatchLevelLog];
    u64 size_log[kBatchLevelLog];
    u64 tick_log[kBatchLevelLog];
    alignas(256) BlockTrace trace[kMaxTraceBlocks];
};

struct BatchArgs {
    u64 n, m_total;
    const u64 *rs;
    const u32 *col;
    u64 *seen;
    u64 *fm[2];
    u64 *next;
    u32 *touched;
    u32 *lv;          // log vertex
    u64 *lm;          // log mask
    u64 *level_start;
    u64 cap;          // log capacity
    u32 *hubs;
    u64 *hoff;
    u32 *hlen;
    BatchCtl *ctl;
    i32 nroots;
    f32 alpha, beta;
    u32 hub_degree;
};

__device__ __forceinline__ void batch_sync(BatchCtl *ctl, u32 phase, u32 level) {
    grid_sync(&ctl->bar, &ctl->bar.error);
    if (threadIdx.x == 0 && blockIdx.x < kMaxTraceBlocks) {
        BlockTrace *t = &ctl->trace[blockIdx.x];
        st_relaxed(&t->barriers, ld_relaxed(&t->barriers) + 1);
        st_relaxed(&t->phase, phase);
        st_relaxed(&t->a, level);
    }
}

// Wave-converged append of records (v, mask) for lanes with `want`; also the
// level's edge count, live-root mask and hub list in `slot`.
__device__ __forceinline__ void append_record(const BatchArgs &a, bool want, u32 v, u64 mask, int slot, int buf,
                                              u64 &edges, u64 &alive) {
    const u64 m = ballot(want);
    if (!m) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    u64 base = 0;
    if (lane_id() == leader) {
        base = atomicAdd((unsigned long long *)&a.ctl->nrec, (unsigned long long)__popcll(m));
        atomicAdd((unsigned long long *)&a.ctl->rcount[slot], (unsigned long long)__popcll(m));
    }
    base = shfl64(base, leader);
    const u64 idx = base + mask_rank(m);
    u32 deg = 0;
    if (want) {
        if (idx < a.cap) {
            a.lv[idx] = v;
            a.lm[idx] = mask;
        } else {
            atomicOr(&a.ctl->bar.error, 2u);
        }
        deg = (u32)(a.rs[v + 1] - a.rs[v]);
        edges += deg;
        alive |= mask;
    }
    const u64 h = (u64)buf * a.n;
    hub_append(&a.ctl->hub[slot], a.hubs + h, a.hoff + h, a.hlen + h, want && deg > a.hub_degree && idx < a.cap,
               (u32)idx, deg, a.n, &a.ctl->bar.error);
}

// Folds per-lane edge counts and masks into the slot's counters.
__device__ __forceinline__ void flush_level(const BatchArgs &a, int slot, u64 edges, u64 alive) {
    edges = wave_sum(edges);
    for (int o = 32; o > 0; o >>= 1) alive |= shfl_xor64(alive, o);
    if (lane_id() == 0) {
        if (edges) atomicAdd((unsigned long long *)&a.ctl->fedges[slot], (unsigned long long)edges);
        if (alive) atomicOr((unsigned long long *)&a.ctl->alive[slot], (unsigned long long)alive);
    }
}

// Top-down edge: OR the roots new to u into next[u]; the first OR of the level
// appends u to the touched list.
__device__ __forceinline__ void td_edge(const BatchArgs &a, bool act, u32 u, u64 mask, int slot) {
    bool first = false;
    if (act) {
        const u64 nb = mask & ~a.seen[u];
        if (nb) first = atomicOr((unsigned long long *)&a.next[u], (unsigned long long)nb) == 0;
    }
    const u32 s = wave_append(&a.ctl->ntouched[slot], first);
    if (first) a.touched[s] = u;
}

// Touched vertices of `slot` -> records of the next level.
__device__ void compact_touched(const BatchArgs &a, int slot, int buf) {
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u64 count = MIN((u64)ld_relaxed(&a.ctl->ntouched[slot]), a.n);
    const u64 nround = (count + gthreads - 1) / gthreads * gthreads;
    u64 edges = 0, alive = 0;
    for (u64 i = gtid; i < nround; i += gthreads) {
        bool want = false;
        u32 u = 0;
        u64 nb = 0;
        if (i < count) {
            u = a.touched[i];
            nb = a.next[u];
            a.next[u] = 0;
            nb &= ~a.seen[u];
            a.seen[u] |= nb;
            want = nb != 0;
        }
        append_record(a, want, u, nb, slot, buf, edges, alive);
    }
    flush_level(a, slot, edges, alive);
}

__global__ __launch_bounds__(512) void batch_bfs_kernel(BatchArgs a) {
    BatchCtl *ctl = a.ctl;
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u32 lane = lane_id();
    const u64 wave_gid = gtid / kWave, nwaves = gthreads / kWave;

    // level 0: the roots (repeats share one record)
    if (gtid < kWave) {
        bool first = false;
        u32 v = 0;
        if ((i32)gtid < a.nroots) {
            v = (u32)ctl->roots[gtid];
            first = atomicOr((unsigned long long *)&a.next[v], 1ull << gtid) == 0;
        }
        const u32 s = wave_append(&ctl->ntouched[0], first);
        if (first) a.touched[s] = v;
    }
    batch_sync(ctl, 1, 0);
    compact_touched(a, 0, 0);
    batch_sync(ctl, 2, 0);

    u64 lo = 0, hi = ld_relaxed(&ctl->rcount[0]);
    bool bu_prev = false;  // whether the current level was produced by a bottom-up pass
    bool bottom_up = false;
    u64 explored = 0, prev_count = 0;
    u64 t_prev = gtid == 0 ? wall_clock64() : 0;
    for (u32 L = 0;; ++L) {
        const int cur = (int)(L % 3), nxt = (int)((L + 1) % 3), clr = (int)((L + 2) % 3);
        if (hi == lo || ld_relaxed(&ctl->bar.error) || hi > a.cap) {
            if (gtid == 0) {
                st_relaxed(&ctl->levels, L);
                st_relaxed(&a.level_start[L], hi);
            }
            break;
        }
        const u64 count = hi - lo;
        const u64 fedges = ld_relaxed(&ctl->fedges[cur]);
        const u64 alive_cur = ld_relaxed(&ctl->alive[cur]);
        const u64 unexplored = a.m_total > explored + fedges ? a.m_total - explored - fedges : 0;
        if (!bottom_up) bottom_up = unexplored > 0 && (f32)fedges > (f32)unexplored / a.alpha;
        else bottom_up = !((f32)count < (f32)a.n / a.beta && count < prev_count);
        explored += fedges;
        prev_count = count;
        if (gtid == 0) {
            st_relaxed(&ctl->ntouched[clr], 0u);
            st_relaxed(&ctl->hub[clr], 0ull);
            st_relaxed(&ctl->fedges[clr], 0ull);
            st_relaxed(&ctl->alive[clr], 0ull);
            st_relaxed(&ctl->rcount[clr], 0ull);
            st_relaxed(&a.level_start[L], lo);
            if (L < kBatchLevelLog) {
                st_relaxed(&ctl->mode_log[L], (i32)bottom_up);
                st_relaxed(&ctl->size_log[L], count);
            }
        }
        u64 edges = 0, alive = 0;
        if (bottom_up) {
            u64 *fmc = a.fm[L & 1], *fmn = a.fm[(L + 1) & 1];
            if (!bu_prev) {  // the level's records -> dense frontier masks
                for (u64 v = gtid; v < a.n; v += gthreads) fmc[v] = 0;
                batch_sync(ctl, 3, L);
                for (u64 i = lo + gtid; i < hi; i += gthreads) fmc[a.lv[i]] = a.lm[i];
                batch_sync(ctl, 4, L);
            }
            const u64 nround = (a.n + gthreads - 1) / gthreads * gthreads;
            for (u64 u = gtid; u < nround; u += gthreads) {
                u64 nb = 0;
                if (u < a.n) {
                    const u64 s = a.seen[u];
                    const u64 need = alive_cur & ~s;
                    if (need) {
                        u64 acc = 0;
                        for (u64 k = a.rs[u], e = a.rs[u + 1]; k < e; ++k) {
                            acc |= fmc[a.col[k]];
                            if ((acc & need) == need) break;
                        }
                        nb = acc & need;
                        if (nb) a.seen[u] = s | nb;
                    }
                    fmn[u] = nb;
                }
                append_record(a, nb != 0, (u32)u, nb, nxt, (int)((L + 1) & 1), edges, alive);
            }
            flush_level(a, nxt, edges, alive);
            batch_sync(ctl, 5, L);
            bu_prev = true;
        } else {
            // hub records: chunks of the flattened hub edge space over the grid
            const u64 hp = ld_relaxed(&ctl->hub[cur]);
            const u64 nh = hub_count(hp), ht = hub_total(hp);
            const u64 hb = (u64)(L & 1) * a.n;
            const u32 *hubs = a.hubs + hb, *hlen = a.hlen + hb;
            const u64 *hoff = a.hoff + hb;
            u64 hi_idx = 0;
            for (u64 c = wave_gid; c * kHubChunk < ht; c += nwaves) {
                const u64 p = c * kHubChunk;
                hi_idx = hub_locate(hoff, hi_idx, nh, p);
                const u32 rec = hubs[hi_idx];
                const u32 v = a.lv[rec];
                const u64 mask = a.lm[rec];
                const u64 b = a.rs[v] + (p - hoff[hi_idx]), e = a.rs[v] + hlen[hi_idx];
#pragma unroll
                for (u64 j = 0; j < kHubChunk; j += kWave) {
                    const u64 k = b + j + lane;
                    td_edge(a, k < e, k < e ? a.col[k] : 0, mask, nxt);
                }
            }
            for (u64 i = lo + wave_gid; i < hi; i += nwaves) {  // a wave per other record
                const u32 v = a.lv[i];
                const u64 b = a.rs[v], e = a.rs[v + 1];
                if (e - b > a.hub_degree) continue;
                const u64 mask = a.lm[i];
                for (u64 k0 = b; k0 < e; k0 += kWave) {
                    const u64 k = k0 + lane;
                    td_edge(a, k < e, k < e ? a.col[k] : 0, mask, nxt);
                }
            }
            batch_sync(ctl, 6, L);
            compact_touched(a, nxt, (int)((L + 1) & 1));
            batch_sync(ctl, 7, L);
            bu_prev = false;
        }
        lo = hi;
        hi = lo + ld_relaxed(&ctl->rcount[nxt]);
        if (gtid == 0 && L < kBatchLevelLog) {
            const u64 t = wall_clock64();
            st_relaxed(&ctl->tick_log[L], t - t_prev);
            t_prev = t;
        }
    }
}

// Depth of every vertex from root r out of the discovery log.
__global__ __launch_bounds__(256) void batch_depth_fill(i64 *__restrict__ depth, u64 n) {
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x) depth[v] = -1;
}

__global__ __launch_bounds__(256) void batch_depth_scatter(const u32 *__restrict__ lv, const u64 *__restrict__ lm,
                                                           const u64 *__restrict__ level_start, u32 levels, int r,
                                                           i64 *__restrict__ depth) {
    const u64 nrec = level_start[levels];
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nrec; i += (u64)gridDim.x * blockDim.x) {
        if (!((lm[i] >> r) & 1)) continue;
        u32 lo = 0, hi = levels;  // level L with level_start[L] <= i < level_start[L + 1]
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) / 2;
            if (level_start[mid] <= i) lo = mid;
            else hi = mid;
        }
        depth[lv[i]] = lo;
    }
}

}  // namespace g500

using namespace g500;

extern "C" {

Bfs_Batch_State oned_graph_bfs_batch_create(Worker_State *ws, const Oned_Graph *g) {
    Bfs_Batch_State b;
    memset(&b, 0, sizeof(b));
    const u64 n = g->nlocalverts;
    b.max_nlocalverts = n;
    b.log_capacity = 8 * n + 64;
    b.seen = (u64 *)arena_alloc(ws, n * sizeof(u64), "batch seen");
    b.fmask[0] = (u64 *)arena_alloc(ws, n * sizeof(u64), "batch frontier masks");
    b.fmask[1] = (u64 *)arena_alloc(ws, n * sizeof(u64), "batch frontier masks");
    b.next = (u64 *)arena_alloc(ws, n * sizeof(u64), "batch next masks");
    b.touched = (u32 *)arena_alloc(ws, n * sizeof(u32), "batch touched");
    b.log_vertex = (u32 *)arena_alloc(ws, b.log_capacity * sizeof(u32), "batch log vertices");
    b.log_mask = (u64 *)arena_alloc(ws, b.log_capacity * sizeof(u64), "batch log masks");
    b.level_start = (u64 *)arena_alloc(ws, (n + 2) * sizeof(u64), "batch level starts");
    b.hubs = (u32 *)arena_alloc(ws, 2 * n * sizeof(u32), "batch hubs");
    b.hub_offsets = (u64 *)arena_alloc(ws, 2 * n * sizeof(u64), "batch hub offsets");
    b.hub_lengths = (u32 *)arena_alloc(ws, 2 * n * sizeof(u32), "batch hub lengths");
    b.ctl = arena_alloc(ws, sizeof(BatchCtl), "batch control");
    return b;
}

void oned_graph_bfs_batch_run(Worker_State *ws, const Oned_Graph *g, Bfs_Batch_State *b, const u64 *roots,
                              i32 nroots) {
    (void)ws;
    Context &c = ctx();
    if (nroots < 1 || nroots > 64) {
        fail("batched BFS takes 1..64 roots (got %d)", nroots);
        return;
    }
    for (i32 i = 0; i < nroots; ++i)
        if (roots[i] >= g->nglobalverts) {
            fail("BFS root %llu is outside graph vertex range [0, %llu)", (unsigned long long)roots[i],
                 (unsigned long long)g->nglobalverts);
            return;
        }
    const u64 n = g->nlocalverts;
    static BatchCtl h;
    memset(&h, 0, sizeof(h));
    for (i32 i = 0; i < nroots; ++i) h.roots[i] = roots[i];
    const hipStream_t s = c.stream;
    G500_HIP(hipMemcpyAsync(b->ctl, &h, sizeof(h), hipMemcpyHostToDevice, s));
    G500_HIP(hipMemsetAsync(b->seen, 0, n * sizeof(u64), s));
    G500_HIP(hipMemsetAsync(b->next, 0, n * sizeof(u64), s));
    sync();
    BatchArgs a;
    a.n = n;
    a.m_total = g->nlocaledges;
    a.rs = g->rowstarts;
    a.col = g->column;
    a.seen = b->seen;
    a.fm[0] = b->fmask[0];
    a.fm[1] = b->fmask[1];
    a.next = b->next;
    a.touched = b->touched;
    a.lv = b->log_vertex;
    a.lm = b->log_mask;
    a.level_start = b->level_start;
    a.cap = b->log_capacity;
    a.hubs = b->hubs;
    a.hoff = b->hub_offsets;
    a.hlen = b->hub_lengths;
    a.ctl = (BatchCtl *)b->ctl;
    a.nroots = nroots;
    a.alpha = (f32)c.batch_alpha;
    a.beta = (f32)c.batch_beta;
    a.hub_degree = (u32)c.bfs_hub_degree;
    const int block = 512;
    int grid = cooperative_blocks((const void *)batch_bfs_kernel, block, 0);
    if (failed()) return;
    void *args[] = {&a};
    timer_start();
    launch_persistent((const void *)batch_bfs_kernel, grid, block, args);
    timer_stop();
    g500_copy_to_host(&h, b->ctl, sizeof(h));
    b->nroots = nroots;
    b->levels = (i32)h.levels;
    if (h.bar.error)
        fail("batched BFS kernel error %u (1 barrier timeout, 2 log or hub overflow)", h.bar.error);
}

void oned_graph_bfs_batch_depth(Worker_State *ws, const Oned_Graph *g, const Bfs_Batch_State *b, i32 r,
                                i64 *depth) {
    (void)ws;
    Context &c = ctx();
    if (r < 0 || r >= b->nroots) {
        fail("root index %d outside the last batch of %d roots", r, b->nroots);
        return;
    }
    const u64 n = g->nlocalverts;
    batch_depth_fill<<<grid_for(n, 256, (u64)c.num_cus * 32), 256, 0, c.stream>>>(depth, n);
    u64 nrec = 0;
    G500_HIP(hipMemcpyAsync(&nrec, b->level_start + b->levels, sizeof(u64), hipMemcpyDeviceToHost, c.stream));
    sync();
    batch_depth_scatter<<<grid_for(nrec, 256, (u64)c.num_cus * 32), 256, 0, c.stream>>>(
        b->log_vertex, b->log_mask, b->level_start, (u32)b->levels, r, depth);
    G500_HIP(hipGetLastError());
    sync();
}

i32 g500_bfs_batch_level_log(const Bfs_Batch_State *b, i32 *modes, u64 *records, f64 *ms, i32 max_levels) {
    static BatchCtl h;
    g500_copy_to_host(&h, b->ctl, sizeof(h));
    const i32 L = (i32)MIN((u32)h.levels, (u32)kBatchLevelLog);
    for (i32 i = 0; i < L && i < max_levels; ++i) {
        if (modes) modes[i] = h.mode_log[i];
        if (records) records[i] = h.size_log[i];
        if (ms) ms[i] = (f64)h.tick_log[i] / ctx().wall_clock_khz;
    }
    return L;
}

void oned_graph_bfs_batch_destroy(Bfs_Batch_State *b) {
    if (b) memset(b, 0, sizeof(*b));
}

}  // extern "C"
