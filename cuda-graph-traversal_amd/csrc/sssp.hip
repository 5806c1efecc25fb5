// Delta-stepping SSSP as one persistent cooperative kernel per root.
// Reference: src/traversal.cu:209-340 (kernels), :570-716 (state, host bucket
// loop with delta = 0.1f).
//
// Labels: one u64 per vertex, (f32 distance bits << 32) | parent. Distances
// are non-negative, so their bit patterns order like the floats; a relaxation
// lowers a label with a compare-and-swap loop only while the candidate is
// strictly smaller (the reference's CAS plus per-vertex spin lock,
// traversal.cu:211-240, becomes one lock-free 64-bit CAS). Every terminating
// label-correcting schedule reaches the same FP32 fixed point, so the bucket
// schedule below does not change any distance.
//
// Buckets are [k*delta, (k+1)*delta) by the index floor(d / delta). Rows are
// sorted by weight, so the light edges (w < delta) of a row are its first
// split[u] entries and the heavy edges the rest. Per bucket:
//   light iterations: frontier vertices relax their light edges; vertices
//     lowered into the bucket are queued for the next iteration (de-duplicated
//     per iteration) and logged as bucket members;
//   heavy pass: every member not yet processed relaxes its heavy edges once;
//     a heavy candidate that rounds back into the bucket re-opens the light
//     phase for a new member round;
//   scan: one pass over all labels collects bucket k+1 and the smallest later
//     bucket (empty buckets are skipped).
// Work is mapped by row length: a lane per vertex (<= 32 entries), a wave per
// vertex (<= 8192), or the whole grid per vertex (longer).
#include "common.hpp"

namespace g500 {

constexpr u32 kSmall = 32;
constexpr u32 kHub = 8192;
constexpr u64 kInfLabel = (0x7F800000ull << 32) | 0xFFFFFFFFull;

// Error bits: 1 barrier timeout, 2 member log overflow, 4 queue entry out of
// range, 8 queue length above n.
struct SsspCtl {
    GridBarrier bar;
    alignas(256) u32 flen[3][3];   // [slot][tier] light queue lengths
    u32 mlen[3];                   // member log lengths per tier
    u32 nmin[3];                   // smallest bucket index beyond the next, per scan slot
    alignas(64) u64 buckets, light_iters, improvements;
    alignas(256) BlockTrace trace[kMaxTraceBlocks];
};

enum : u32 { kPhaseInit = 1, kPhaseLight = 2, kPhaseHeavy = 3, kPhaseScan = 4 };

__device__ __forceinline__ void sync_traced(SsspCtl *ctl, u32 phase, u32 a, u32 b) {
    grid_sync(&ctl->bar, &ctl->bar.error);
    if (threadIdx.x == 0 && blockIdx.x < kMaxTraceBlocks) {
        BlockTrace *t = &ctl->trace[blockIdx.x];
        st_relaxed(&t->barriers, ld_relaxed(&t->barriers) + 1);
        st_relaxed(&t->phase, phase);
        st_relaxed(&t->a, a);
        st_relaxed(&t->b, b);
    }
}

struct SsspArgs {
    u64 n;
    const u64 *rs;
    const u32 *col;
    const f32 *w;
    const u32 *split;
    u64 *lab;
    u32 *fq[2][3];   // light queues [buffer][tier]
    u32 *mq[3];      // member logs [tier]
    u64 mcap;        // capacity of each member log
    u32 *stamp, *mstamp;
    SsspCtl *ctl;
    i64 *pred;
    f32 *dist;
    u64 root;
    f32 inv_delta;
};

__device__ __forceinline__ f32 label_dist(u64 l) { return __uint_as_float((u32)(l >> 32)); }

__device__ __forceinline__ u32 bucket_of(f32 d, f32 inv_delta) {
    const f32 x = d * inv_delta;
    return x >= 4.0e9f ? 0xFFFFFFF0u : (u32)x;
}

__device__ __forceinline__ int tier_of(u32 len) { return len <= kSmall ? 0 : (len <= kHub ? 1 : 2); }

// Strictly lowers the label of v to (cand, parent); true if this call did it.
__device__ __forceinline__ bool relax(u64 *lab, u32 v, f32 cand, u32 parent) {
    const u32 cb = __float_as_uint(cand);
    const u64 want = ((u64)cb << 32) | parent;
    u64 old = lab[v];
    while ((u32)(old >> 32) > cb) {
        const u64 prev = atomicCAS((unsigned long long *)&lab[v], (unsigned long long)old, (unsigned long long)want);
        if (prev == old) return true;
        old = prev;
    }
    return false;
}

// Wave-converged: queues v (lanes with `want`) into the light queue `slot` of
// buffer `buf` under tag `ltag`, and logs it as a member under `mtag`.
__device__ __forceinline__ void push_vertex(const SsspArgs &a, bool want, u32 v, int buf, int slot, u32 ltag,
                                            bool log_member, u32 mtag) {
    if (!ballot(want)) return;
    bool q = false, m = false;
    u32 lw = 0, hw = 0;
    if (want) {
        q = atomicExch(&a.stamp[v], ltag) != ltag;
        if (log_member) m = atomicExch(&a.mstamp[v], mtag) != mtag;
        if (q || m) {
            const u32 deg = (u32)(a.rs[v + 1] - a.rs[v]);
            lw = a.split[v];
            hw = deg - lw;
        }
    }
    SsspCtl *ctl = a.ctl;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const bool qt = q && tier_of(lw) == t;
        const u32 s = wave_append(&ctl->flen[slot][t], qt);
        if (qt) a.fq[buf][t][s] = v;
        const bool mt = m && tier_of(hw) == t;
        const u32 s2 = wave_append(&ctl->mlen[t], mt);
        if (mt) {
            if (s2 < a.mcap) a.mq[t][s2] = v;
            else atomicOr(&ctl->bar.error, 2u);
        }
    }
}

struct Visit {
    int buf, slot;     // where pushes go
    u32 ltag, mtag;
    bool log_member;
    u32 bucket;        // current bucket: pushes only for candidates landing in it or below
};

// One relaxation per active lane, then the wave-converged pushes.
__device__ __forceinline__ void visit_edge(const SsspArgs &a, const Visit &vs, bool act, u32 u, f32 du, u64 k,
                                           u64 &improved) {
    bool push = false;
    u32 v = 0;
    if (act) {
        v = a.col[k];
        const f32 cand = du + a.w[k];
        if (relax(a.lab, v, cand, u)) {
            ++improved;
            push = bucket_of(cand, a.inv_delta) <= vs.bucket;
        }
    }
    push_vertex(a, push, v, vs.buf, vs.slot, vs.ltag, vs.log_member, vs.mtag);
}

// Processes the vertices list[begin, end) of one work tier; light selects the
// row prefix [0, split) or the suffix [split, deg).
template <bool LIGHT>
__device__ void process_tier(const SsspArgs &a, const u32 *list, u32 begin, u32 end, int tier, const Visit &vs,
                             u64 &improved) {
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u32 lane = lane_id();
    const u64 wave_gid = gtid / kWave, nwaves = gthreads / kWave;
    if (tier == 0) {
        for (u64 i0 = begin + wave_gid * kWave; i0 < end; i0 += nwaves * kWave) {
            const u64 i = i0 + lane;
            u64 b = 0, e = 0;
            u32 u = 0;
            f32 du = 0.f;
            if (i < end && list[i] < a.n) {
                u = list[i];
                const u64 r0 = a.rs[u], s = a.split[u];
                b = LIGHT ? r0 : r0 + s;
                e = LIGHT ? r0 + s : a.rs[u + 1];
                du = label_dist(a.lab[u]);
            }
            u32 len = (u32)(e - b), mx = len;
            for (int o = 32; o > 0; o >>= 1) mx = max(mx, (u32)__shfl_xor((int)mx, o));
            for (u32 j = 0; j < mx; ++j) visit_edge(a, vs, j < len, u, du, b + j, improved);
        }
    } else if (tier == 1) {
        for (u64 i = begin + wave_gid; i < end; i += nwaves) {
            const u32 u = list[i];
            if (u >= a.n) {
                atomicOr(&a.ctl->bar.error, 4u);
                continue;
            }
            const u64 r0 = a.rs[u], s = a.split[u];
            const u64 b = LIGHT ? r0 : r0 + s, e = LIGHT ? r0 + s : a.rs[u + 1];
            const f32 du = label_dist(a.lab[u]);
            for (u64 k0 = b; k0 < e; k0 += kWave) visit_edge(a, vs, k0 + lane < e, u, du, k0 + lane, improved);
        }
    } else {
        for (u64 i = begin; i < end; ++i) {
            const u32 u = list[i];
            if (u >= a.n) {
                atomicOr(&a.ctl->bar.error, 4u);
                continue;
            }
            const u64 r0 = a.rs[u], s = a.split[u];
            const u64 b = LIGHT ? r0 : r0 + s, e = LIGHT ? r0 + s : a.rs[u + 1];
            const f32 du = label_dist(a.lab[u]);
            for (u64 k0 = b + wave_gid * kWave; k0 < e; k0 += nwaves * kWave)
                visit_edge(a, vs, k0 + lane < e, u, du, k0 + lane, improved);
        }
    }
}

// Control flow is a sequence of steps (light iterations, heavy passes, scans).
// Step `st` reads its queue and its decision counters from slot st % 3,
// appends only into slot (st + 1) % 3 (queue buffer (st + 1) & 1, tag
// st + 2) and clears slot (st + 2) % 3, so a block that already acts on the
// decision of step st never changes a counter that a slower block has still to
// read for the same decision.
__global__ __launch_bounds__(512) void sssp_kernel(SsspArgs a) {
    SsspCtl *ctl = a.ctl;
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u32 lane = lane_id();
    u64 improved = 0;

    u32 st = 0;      // step counter
    u32 mt = 1;      // member round tag
    u32 mproc[3] = {0, 0, 0};
    u32 k = 0;       // current bucket

    if (gtid == 0) {
        const u32 r = (u32)a.root;
        st_relaxed(&a.lab[r], (u64)r);  // distance 0.0f, parent = root
#pragma unroll
        for (int i = 0; i < 3; ++i) st_relaxed(&ctl->nmin[i], 0xFFFFFFFFu);
    }
    sync_traced(ctl, kPhaseInit, 0, 0);
    // step 0 queues the root (lane 0 of the first wave) into the list of step 1
    if (gtid < kWave) push_vertex(a, lane == 0, (u32)a.root, 1, 1, 2, true, mt);
    sync_traced(ctl, kPhaseInit, 1, 0);
    st = 1;

    auto slot_total = [&](u32 step, u32 *cnt) {
        u32 total = 0;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            u32 c = ld_relaxed(&ctl->flen[step % 3][t]);
            if (c > a.n) {
                if (gtid == 0) atomicOr(&ctl->bar.error, 8u);
                c = 0;
            }
            if (cnt) cnt[t] = c;
            total += c;
        }
        return total;
    };
    // Every step clears the queue counters of step + 2 and the scan minimum of
    // step + 1: both were last read right after the barrier that closed step - 1.
    auto clear_ahead = [&](u32 step) {
        if (gtid == 0) {
#pragma unroll
            for (int t = 0; t < 3; ++t) st_relaxed(&ctl->flen[(step + 2) % 3][t], 0u);
            st_relaxed(&ctl->nmin[(step + 1) % 3], 0xFFFFFFFFu);
        }
    };

    for (;;) {  // buckets
        for (;;) {  // member rounds of bucket k
            for (;;) {  // light iterations
                u32 cnt[3];
                const u32 total = slot_total(st, cnt);
                clear_ahead(st);
                if (total == 0 || ld_relaxed(&ctl->bar.error)) break;
                if (gtid == 0) atomicAdd((unsigned long long *)&ctl->light_iters, 1ull);
                const Visit vs{(int)((st + 1) & 1), (int)((st + 1) % 3), st + 2, mt, true, k};
#pragma unroll
                for (int t = 0; t < 3; ++t) process_tier<true>(a, a.fq[st & 1][t], 0, cnt[t], t, vs, improved);
                sync_traced(ctl, kPhaseLight, k, st);
                ++st;
            }
            if (ld_relaxed(&ctl->bar.error)) break;
            // heavy pass (same step: its input queue was empty) over the
            // unprocessed member suffix; re-opened vertices go to the next step's
            // queue under a new member round
            u32 mend[3];
#pragma unroll
            for (int t = 0; t < 3; ++t) mend[t] = min(ld_relaxed(&ctl->mlen[t]), (u32)a.mcap);
            const Visit vs{(int)((st + 1) & 1), (int)((st + 1) % 3), st + 2, mt + 1, true, k};
#pragma unroll
            for (int t = 0; t < 3; ++t) process_tier<false>(a, a.mq[t], mproc[t], mend[t], t, vs, improved);
#pragma unroll
            for (int t = 0; t < 3; ++t) mproc[t] = mend[t];
            sync_traced(ctl, kPhaseHeavy, k, st);
            ++st;
            ++mt;
            if (!slot_total(st, nullptr) || ld_relaxed(&ctl->bar.error)) break;
        }
        if (ld_relaxed(&ctl->bar.error)) break;
        if (gtid == 0) atomicAdd((unsigned long long *)&ctl->buckets, 1ull);
        // scan steps until the next non-empty bucket is queued
        bool done = false;
        for (;;) {
            clear_ahead(st);
            const int ss = (int)(st % 3);
            const u64 nround = (a.n + gthreads - 1) / gthreads * gthreads;
            for (u64 v = gtid; v < nround; v += gthreads) {
                bool want = false;
                if (v < a.n) {
                    const u64 l = a.lab[v];
                    if ((u32)(l >> 32) < 0x7F800000u) {
                        const u32 bk = bucket_of(label_dist(l), a.inv_delta);
                        if (bk == k + 1) want = true;
                        else if (bk > k + 1) atomicMin(&ctl->nmin[ss], bk);
                    }
                }
                push_vertex(a, want, (u32)v, (int)((st + 1) & 1), (int)((st + 1) % 3), st + 2, true, mt);
            }
            sync_traced(ctl, kPhaseScan, k, st);
            const u32 nm = ld_relaxed(&ctl->nmin[ss]);
            ++st;
            if (slot_total(st, nullptr)) {
                ++k;
                break;
            }
            if (nm == 0xFFFFFFFFu || ld_relaxed(&ctl->bar.error)) {
                done = true;
                break;
            }
            k = nm - 1;
        }
        if (done) break;
    }
    improved = wave_sum(improved);
    if (lane == 0 && improved) atomicAdd((unsigned long long *)&ctl->improvements, (unsigned long long)improved);
    // export (traversal.cu:333-340): unreached -> -1.0f / -1
    for (u64 v = gtid; v < a.n; v += gthreads) {
        const u64 l = a.lab[v];
        const bool reached = (u32)(l >> 32) < 0x7F800000u;
        a.dist[v] = reached ? label_dist(l) : -1.0f;
        a.pred[v] = reached ? (i64)(u32)l : -1;
    }
}

__global__ __launch_bounds__(256) void sssp_split_kernel(const u64 *__restrict__ rs, const f32 *__restrict__ w,
                                                         u64 n, f32 delta, u32 *__restrict__ split) {
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x) {
        u64 lo = rs[v], hi = rs[v + 1];  // first entry with w >= delta
        const u64 b = lo;
        while (lo < hi) {
            const u64 mid = lo + (hi - lo) / 2;
            if (w[mid] < delta) lo = mid + 1;
            else hi = mid;
        }
        split[v] = (u32)(lo - b);
    }
}

__global__ __launch_bounds__(256) void sssp_clear_kernel(u64 *__restrict__ lab, u32 *__restrict__ stamp,
                                                         u32 *__restrict__ mstamp, u64 n) {
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x) {
        lab[v] = kInfLabel;
        stamp[v] = 0;
        mstamp[v] = 0;
    }
}

}  // namespace g500

using namespace g500;

extern "C" {

Sssp_State oned_graph_sssp_create(Worker_State *ws, const Oned_Graph *g) {
    Sssp_State s;
    memset(&s, 0, sizeof(s));
    const u64 n = g->nlocalverts;
    s.max_nlocalverts = n;
    s.delta = 0.1f;
    s.pred = (i64 *)arena_alloc(ws, n * sizeof(i64), "SSSP pred");
    s.dist = (f32 *)arena_alloc(ws, n * sizeof(f32), "SSSP dist");
    s.dist_pred = (u64 *)arena_alloc(ws, n * sizeof(u64), "SSSP labels");
    s.frontier = (u32 *)arena_alloc(ws, 6 * n * sizeof(u32), "SSSP queues");
    s.members = (u32 *)arena_alloc(ws, 6 * n * sizeof(u32), "SSSP member logs");
    s.split = (u32 *)arena_alloc(ws, n * sizeof(u32), "SSSP light split");
    s.stamp = (u32 *)arena_alloc(ws, n * sizeof(u32), "SSSP stamps");
    s.mstamp = (u32 *)arena_alloc(ws, n * sizeof(u32), "SSSP member stamps");
    s.ctl = arena_alloc(ws, sizeof(SsspCtl), "SSSP control");
    if (failed()) return s;
    oned_graph_sssp_clear(ws, &s);
    return s;
}

void oned_graph_sssp_clear(Worker_State *ws, Sssp_State *s) {
    (void)ws;
    Context &c = ctx();
    const u64 n = s->max_nlocalverts;
    sssp_clear_kernel<<<grid_for(n, 256, (u64)c.num_cus * 32), 256, 0, c.stream>>>(s->dist_pred, s->stamp, s->mstamp, n);
    G500_HIP(hipGetLastError());
    G500_HIP(hipMemsetAsync(s->ctl, 0, sizeof(SsspCtl), c.stream));
    sync();
}

void oned_graph_sssp_run(Worker_State *ws, const Oned_Graph *g, Sssp_State *s, u64 root) {
    (void)ws;
    Context &c = ctx();
    if (root >= g->nglobalverts) {
        fail("SSSP root %llu is outside graph vertex range [0, %llu)", (unsigned long long)root,
             (unsigned long long)g->nglobalverts);
        return;
    }
    if (!(s->delta > 0.0f)) {
        fail("SSSP delta must be positive (got %g)", (double)s->delta);
        return;
    }
    const u64 n = g->nlocalverts;
    if (s->split_delta != s->delta) {  // light/heavy split of every row for this delta
        sssp_split_kernel<<<grid_for(n, 256, (u64)c.num_cus * 32), 256, 0, c.stream>>>(g->rowstarts, g->weights, n,
                                                                                      s->delta, s->split);
        G500_HIP(hipGetLastError());
        s->split_delta = s->delta;
    }
    SsspArgs a;
    a.n = n;
    a.rs = g->rowstarts;
    a.col = g->column;
    a.w = g->weights;
    a.split = s->split;
    a.lab = s->dist_pred;
    for (int b = 0; b < 2; ++b)
        for (int t = 0; t < 3; ++t) a.fq[b][t] = s->frontier + (u64)(b * 3 + t) * n;
    a.mcap = 2 * n;
    for (int t = 0; t < 3; ++t) a.mq[t] = s->members + (u64)t * 2 * n;
    a.stamp = s->stamp;
    a.mstamp = s->mstamp;
    a.ctl = (SsspCtl *)s->ctl;
    a.pred = s->pred;
    a.dist = s->dist;
    a.root = root;
    a.inv_delta = 1.0f / s->delta;
    const int block = 512;
    int grid = cooperative_blocks((const void *)sssp_kernel, block, 0);
    if (failed()) return;
    void *args[] = {&a};
    G500_HIP(hipMemsetAsync(s->ctl, 0, sizeof(SsspCtl), c.stream));
    timer_start();
    G500_HIP(hipLaunchCooperativeKernel((const void *)sssp_kernel, dim3(grid), dim3(block), args, 0, c.stream));
    timer_stop();
    static SsspCtl h;
    g500_copy_to_host(&h, s->ctl, sizeof(h));
    if (h.bar.error) {
        u32 lo = 0xFFFFFFFFu, hi = 0;
        int at_lo = 0, at_hi = 0;
        for (int b = 0; b < MIN(grid, kMaxTraceBlocks); ++b) {
            if (h.trace[b].barriers < lo) { lo = h.trace[b].barriers; at_lo = b; }
            if (h.trace[b].barriers > hi) { hi = h.trace[b].barriers; at_hi = b; }
        }
        const BlockTrace &tl = h.trace[at_lo], &th = h.trace[at_hi];
        fail("SSSP kernel error %u (1 barrier timeout, 2 member log overflow, 4 bad queue entry, 8 queue "
             "overflow); %d blocks; barriers passed min %u (block %d phase %u tags %u/%u) max %u (block %d "
             "phase %u tags %u/%u)",
             h.bar.error, grid, lo, at_lo, tl.phase, tl.a, tl.b, hi, at_hi, th.phase, th.a, th.b);
    }
}

void oned_graph_sssp_destroy(Sssp_State *s) {
    if (s) memset(s, 0, sizeof(*s));
}

void g500_sssp_run_stats(const Sssp_State *s, u64 *buckets, u64 *light_iterations, u64 *relaxations) {
    static SsspCtl h;
    g500_copy_to_host(&h, s->ctl, sizeof(h));
    if (buckets) *buckets = h.buckets;
    if (light_iterations) *light_iterations = h.light_iters;
    if (relaxations) *relaxations = h.improvements;
}

}  // extern "C"
