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
constexpr int kSteps = 4;  // bottom-up chunks in flight per wave
constexpr u32 kRecStage = 256;  // bottom-up records a wave stages in LDS before appending them

struct BatchCtl {
    GridBarrier bar;
    // counters that many waves append to sit on lines of their own, so that
    // their atomics do not queue behind each other's
    alignas(256) u64 nrec;         // discovery-log records so far
    alignas(256) u64 rcount[3];    // records produced for a level, per slot
    alignas(256) u32 ntouched[3];  // top-down touched vertices, per slot
    alignas(256) u64 hub[3];       // packed hub counters (record indices), per slot
    alignas(256) u64 fedges[3];    // sum of degrees of a level's records, per slot
    u64 alive[3];                  // OR of a level's masks, per slot
    u64 bhub;                      // packed counter of the heavy-vertex list (bottom-up)
    u32 levels;
    u32 pad;
    alignas(256) i32 mode_log[kBatchLevelLog];
    u64 size_log[kBatchLevelLog];
    u64 edge_log[kBatchLevelLog];  // sum of the degrees of a level's records
    u64 tick_log[kBatchLevelLog];
    alignas(256) BlockTrace trace[kMaxTraceBlocks];
};

struct BatchArgs {
    u64 n, m_total;
    const u64 *rs;    // traversal view
    const u32 *col;
    const u32 *perm;
    u64 *seen;
    u64 *fm[2];
    u64 *fb[2];       // frontier bitmaps matching fm
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
    u64 roots[64];
    i32 nroots;
    f32 alpha, beta;
    u32 hub_degree;    // top-down: records above this degree go to the grid-wide hub list
    u32 heavy_degree;  // bottom-up: vertices above this degree are split over the grid
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

// Top-down edge of a dense level (see compact_dense): the OR alone, without
// waiting for its old value and without a touched list.
__device__ __forceinline__ void td_edge_dense(const BatchArgs &a, bool act, u32 u, u64 mask) {
    if (act) {
        const u64 nb = mask & ~a.seen[u];
        if (nb) atomicOr((unsigned long long *)&a.next[u], (unsigned long long)nb);
    }
}

// Compaction of a dense top-down level, one whose records' edges reach a
// large share of the graph (at SCALE 26 the level after the roots: 3.9e3
// records, 5.4e7 edges, 8.7e6 vertices reached). Instead of a touched list,
// whose appends would take one atomic on a single counter per 64 edges, every
// vertex's next mask is read in one coalesced pass, and a block makes one
// atomic per pass step for the records of all its waves.
__device__ void compact_dense(const BatchArgs &a, int slot, int buf) {
    __shared__ u32 wcnt[1024 / kWave];
    __shared__ u64 bbase;
    const u32 lane = lane_id();
    const int w = (int)(threadIdx.x / kWave), nw = (int)(blockDim.x / kWave);
    u64 edges = 0, alive = 0;
    for (u64 u0 = (u64)blockIdx.x * blockDim.x; u0 < a.n; u0 += (u64)gridDim.x * blockDim.x) {
        const u64 u = u0 + threadIdx.x;
        u64 nb = 0;
        if (u < a.n) {
            nb = a.next[u];
            if (nb) {
                a.next[u] = 0;
                const u64 sn = a.seen[u];
                nb &= ~sn;
                if (nb) a.seen[u] = sn | nb;
            }
        }
        const bool want = nb != 0;
        const u64 m = ballot(want);
        if (lane == 0) wcnt[w] = (u32)__popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            u32 tot = 0;
            for (int i = 0; i < nw; ++i) {
                const u32 c = wcnt[i];
                wcnt[i] = tot;
                tot += c;
            }
            u64 base = 0;
            if (tot) {
                base = atomicAdd((unsigned long long *)&a.ctl->nrec, (unsigned long long)tot);
                atomicAdd((unsigned long long *)&a.ctl->rcount[slot], (unsigned long long)tot);
            }
            bbase = base;
        }
        __syncthreads();
        const u64 idx = bbase + wcnt[w] + mask_rank(m);
        u32 deg = 0;
        if (want) {
            if (idx < a.cap) {
                a.lv[idx] = (u32)u;
                a.lm[idx] = nb;
            } else {
                atomicOr(&a.ctl->bar.error, 2u);
            }
            deg = (u32)(a.rs[u + 1] - a.rs[u]);
            edges += deg;
            alive |= nb;
        }
        const u64 h = (u64)buf * a.n;
        hub_append(&a.ctl->hub[slot], a.hubs + h, a.hoff + h, a.hlen + h, want && deg > a.hub_degree && idx < a.cap,
                   (u32)idx, deg, a.n, &a.ctl->bar.error);
        __syncthreads();  // wcnt and bbase are reused by the next step
    }
    flush_level(a, slot, edges, alive);
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

__device__ __forceinline__ u64 shfl_up64(u64 v, int o) {
    u32 lo = __shfl_up((u32)v, o), hi = __shfl_up((u32)(v >> 32), o);
    return ((u64)hi << 32) | lo;
}

__device__ __forceinline__ u64 uniform64(u64 v) {
    return ((u64)__builtin_amdgcn_readfirstlane((u32)(v >> 32)) << 32) | __builtin_amdgcn_readfirstlane((u32)v);
}

__device__ __forceinline__ bool frontier_bit(const u64 *fb, u32 v) { return (fb[v >> 6] >> (v & 63)) & 1; }

// Bottom-up pass over the vertices that still miss live roots. A wave takes
// 64 consecutive vertices; the edge lists of those with missing roots and
// degree <= heavy_degree are walked as one concatenated range, 64 edges per
// step (coalesced column reads). The frontier bitmap is tested before the
// 8-byte mask gather; a segmented OR scan hands each vertex's partial result
// to its own lane, and vertices whose missing roots are all found stop.
__device__ void bu_groups(const BatchArgs &a, u32 L, u64 alive_cur, int nxt, u64 &edges, u64 &alive,
                          volatile int *slot) {
    const u64 *fmc = a.fm[L & 1], *fbc = a.fb[L & 1];
    u64 *fmn = a.fm[(L + 1) & 1], *fbn = a.fb[(L + 1) & 1];
    const u32 lane = lane_id();
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u64 wave_gid = gtid / kWave, nwaves = gthreads / kWave;
    const u64 ngroups = (a.n + kWave - 1) / kWave;
    // The wave's records are staged in LDS and appended in runs of up to
    // kRecStage: one pair of counter atomics per run instead of one per
    // 64-vertex group (a large level has records in almost every group).
    __shared__ u32 stv[512 / kWave][kRecStage];
    __shared__ u64 stm[512 / kWave][kRecStage];
    volatile u32 *sv = stv[threadIdx.x / kWave];
    volatile u64 *sm = stm[threadIdx.x / kWave];
    u32 sn = 0;  // staged records (wave-uniform)
    const int hbuf = (int)((L + 1) & 1);
    auto stage_flush = [&]() {
        if (!sn) return;
        u64 rb = 0;
        if (lane == 0) {
            rb = atomicAdd((unsigned long long *)&a.ctl->nrec, (unsigned long long)sn);
            atomicAdd((unsigned long long *)&a.ctl->rcount[nxt], (unsigned long long)sn);
        }
        rb = shfl64(rb, 0);
        const u64 h = (u64)hbuf * a.n;
        for (u32 i0 = 0; i0 < sn; i0 += kWave) {
            const u32 i = i0 + lane;
            const bool want = i < sn;
            const u64 idx = rb + i;
            u32 deg = 0, v = 0;
            if (want) {
                v = sv[i];
                const u64 m = sm[i];
                if (idx < a.cap) {
                    a.lv[idx] = v;
                    a.lm[idx] = m;
                } else {
                    atomicOr(&a.ctl->bar.error, 2u);
                }
                deg = (u32)(a.rs[v + 1] - a.rs[v]);
                edges += deg;
                alive |= m;
            }
            hub_append(&a.ctl->hub[nxt], a.hubs + h, a.hoff + h, a.hlen + h,
                       want && deg > a.hub_degree && idx < a.cap, (u32)idx, deg, a.n, &a.ctl->bar.error);
        }
        sn = 0;
    };
    for (u64 grp = wave_gid; grp < ngroups; grp += nwaves) {
        const u64 u = grp * kWave + lane;
        u64 need = 0, base = 0;
        u32 deg = 0;
        if (u < a.n) {
            need = alive_cur & ~a.seen[u];
            if (need) {
                const u64 r0 = a.rs[u], r1 = a.rs[u + 1];
                if (r1 - r0 > a.heavy_degree) need = 0;  // handled by bu_heavy_chunks
                deg = (u32)(r1 - r0);
                base = r0;
            }
        }
        if (!need) deg = 0;
        const u32 incl = wave_inclusive_scan(deg);
        const u32 off = incl - deg, total = __shfl(incl, kWave - 1);
        const u64 shift = base - (u64)off;  // CSR position of concatenated position e: shift + e
        u64 acc = 0;
        u64 done = ballot(deg == 0);
        // kSteps chunks per round: their column reads, bitmap probes and mask
        // gathers are issued together, so each wave keeps several independent
        // load chains in flight
        for (u32 c = 0; c < total && done != ~0ull; c += kSteps * kWave) {
            int jq[kSteps], lastq[kSteps];
            bool interq[kSteps], okq[kSteps];
            u32 vq[kSteps];
            u64 mq[kSteps];
#pragma unroll
            for (int q = 0; q < kSteps; ++q) {
                jq[q] = chunk_owner(slot, off, deg, c + q * kWave, interq[q], lastq[q]);
                if (jq[q] < 0) jq[q] = 0;  // chunk past the end: no valid lane
            }
#pragma unroll
            for (int q = 0; q < kSteps; ++q) {
                const u32 e = c + q * kWave + lane;
                const u64 sj = shfl64(shift, jq[q]);
                okq[q] = e < total && !((done >> jq[q]) & 1);
                vq[q] = okq[q] ? a.col[sj + e] : 0;
            }
#pragma unroll
            for (int q = 0; q < kSteps; ++q) okq[q] = okq[q] && frontier_bit(fbc, vq[q]);
#pragma unroll
            for (int q = 0; q < kSteps; ++q) {
                const u64 nj = shfl64(need, jq[q]);
                mq[q] = okq[q] ? fmc[vq[q]] & nj : 0;
            }
#pragma unroll
            for (int q = 0; q < kSteps; ++q) {
                if (!ballot(mq[q] != 0)) continue;
                u64 x = mq[q];
                const int j = jq[q];
#pragma unroll
                for (int o = 1; o < kWave; o <<= 1) {
                    const u64 y = shfl_up64(x, o);
                    const int yj = __shfl_up(j, o);
                    if ((int)lane >= o && yj == j) x |= y;
                }
                const u64 got = shfl64(x, lastq[q]);
                if (interq[q]) acc |= got;
            }
            done |= ballot(deg > 0 && (acc & need) == need);
        }
        const u64 nb = acc & need;
        if (nb) {
            a.seen[u] |= nb;
            fmn[u] = nb;
        }
        const u64 word = ballot(nb != 0);
        if (lane == 0) fbn[grp] = word;
        if (word) {
            const u32 k = (u32)__popcll(word);
            if (sn + k > kRecStage) stage_flush();
            if (nb) {
                const u32 pos = sn + mask_rank(word);
                sv[pos] = (u32)u;
                sm[pos] = nb;
            }
            sn += k;
        }
    }
    stage_flush();
}

// Heavy vertices with missing roots: their edge lists in 256-edge chunks over
// the grid; each chunk ORs its finds into next[v].
__device__ void bu_heavy_chunks(const BatchArgs &a, u32 L, u64 alive_cur) {
    const u64 *fmc = a.fm[L & 1], *fbc = a.fb[L & 1];
    const u32 lane = lane_id();
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u64 wave_gid = gtid / kWave, nwaves = gthreads / kWave;
    const u64 hp = ld_relaxed(&a.ctl->bhub);
    const u64 nh = hub_count(hp), ht = hub_total(hp);
    const u64 hb = 2 * a.n;
    const u32 *hv = a.hubs + hb, *hlen = a.hlen + hb;
    const u64 *hoff = a.hoff + hb;
    u64 hi_idx = 0;
    for (u64 c = wave_gid; c * kHubChunk < ht; c += nwaves) {
        const u64 p = c * kHubChunk;
        hi_idx = hub_locate(hoff, hi_idx, nh, p);
        const u32 v = hv[hi_idx];
        const u64 need = uniform64(alive_cur & ~a.seen[v]);
        if (!need) continue;
        if ((uniform64(ld_relaxed(&a.next[v])) & need) == need) continue;
        const u64 b = a.rs[v] + (p - hoff[hi_idx]), e = a.rs[v] + hlen[hi_idx];
        u64 acc = 0;
#pragma unroll
        for (u64 j = 0; j < kHubChunk; j += kWave) {
            const u64 k = b + j + lane;
            if (k < e) {
                const u32 w = a.col[k];
                if (frontier_bit(fbc, w)) acc |= fmc[w] & need;
            }
        }
        for (int o = 32; o > 0; o >>= 1) acc |= shfl_xor64(acc, o);
        if (lane == 0 && acc) atomicOr((unsigned long long *)&a.next[v], (unsigned long long)acc);
    }
}

// Heavy vertices: fold the chunk results (next[v], cleared here) into seen,
// the next frontier and the log.
__device__ void bu_heavy_fixup(const BatchArgs &a, u32 L, u64 alive_cur, int nxt, u64 &edges, u64 &alive) {
    u64 *fmn = a.fm[(L + 1) & 1], *fbn = a.fb[(L + 1) & 1];
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u64 nh = MIN(hub_count(ld_relaxed(&a.ctl->bhub)), a.n);
    const u64 nround = (nh + gthreads - 1) / gthreads * gthreads;
    const u32 *hv = a.hubs + 2 * a.n;
    for (u64 i = gtid; i < nround; i += gthreads) {
        u32 v = 0;
        u64 nb = 0;
        if (i < nh) {
            v = hv[i];
            const u64 nx = a.next[v];
            if (nx) {
                a.next[v] = 0;
                const u64 s = a.seen[v];
                nb = nx & alive_cur & ~s;
                if (nb) {
                    a.seen[v] = s | nb;
                    fmn[v] = nb;
                    atomicOr((unsigned long long *)&fbn[v >> 6], 1ull << (v & 63));
                }
            }
        }
        append_record(a, nb != 0, v, nb, nxt, (int)((L + 1) & 1), edges, alive);
    }
}

__global__ __launch_bounds__(512) void batch_bfs_kernel(BatchArgs a) {
    BatchCtl *ctl = a.ctl;
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u32 lane = lane_id();
    const u64 wave_gid = gtid / kWave, nwaves = gthreads / kWave;

    __shared__ int slots[512 / kWave][kWave];
    volatile int *slot = slots[threadIdx.x / kWave];

    // heavy-vertex list for bottom-up levels
    {
        const u64 nround = (a.n + gthreads - 1) / gthreads * gthreads;
        const u64 hb = 2 * a.n;
        for (u64 v = gtid; v < nround; v += gthreads) {
            const u64 deg = v < a.n ? a.rs[v + 1] - a.rs[v] : 0;
            hub_append(&ctl->bhub, a.hubs + hb, a.hoff + hb, a.hlen + hb, deg > a.heavy_degree, (u32)v, (u32)deg,
                       a.n, &ctl->bar.error);
        }
    }
    // level 0: the roots (repeats share one record)
    if (gtid < kWave) {
        bool first = false;
        u32 v = 0;
        if ((i32)gtid < a.nroots) {
            v = trav_id(a.perm, a.roots[gtid]);
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
    u64 prev_count = 0;
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
        // Beamer's rule adapted to root masks: a level's top-down work is the
        // degree sum of its records; bottom-up work is bounded by the edge total
        if (!bottom_up) bottom_up = (f32)fedges > (f32)a.m_total / a.alpha;
        else bottom_up = !((f32)count < (f32)a.n / a.beta && count < prev_count);
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
                st_relaxed(&ctl->edge_log[L], fedges);
            }
        }
        u64 edges = 0, alive = 0;
        if (bottom_up) {
            if (!bu_prev) {  // the level's records -> frontier bitmap and masks
                u64 *fmc = a.fm[L & 1], *fbc = a.fb[L & 1];
                const u64 nwords = (a.n + 63) / 64;
                for (u64 w = gtid; w < nwords; w += gthreads) fbc[w] = 0;
                batch_sync(ctl, 3, L);
                for (u64 i = lo + gtid; i < hi; i += gthreads) {
                    const u32 v = a.lv[i];
                    fmc[v] = a.lm[i];
                    atomicOr((unsigned long long *)&fbc[v >> 6], 1ull << (v & 63));
                }
                batch_sync(ctl, 4, L);
            }
            const bool heavy = hub_count(ld_relaxed(&ctl->bhub)) > 0;
            if (heavy) bu_heavy_chunks(a, L, alive_cur);
            bu_groups(a, L, alive_cur, nxt, edges, alive, slot);
            if (heavy) {
                batch_sync(ctl, 8, L);
                bu_heavy_fixup(a, L, alive_cur, nxt, edges, alive);
            }
            flush_level(a, nxt, edges, alive);
            batch_sync(ctl, 5, L);
            bu_prev = true;
        } else {
            // a level whose edges reach a large share of the graph is dense (grid-uniform)
            const bool dense = fedges >= a.n / 16;
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
                    if (dense)
                        td_edge_dense(a, k < e, k < e ? a.col[k] : 0, mask);
                    else
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
                    if (dense)
                        td_edge_dense(a, k < e, k < e ? a.col[k] : 0, mask);
                    else
                        td_edge(a, k < e, k < e ? a.col[k] : 0, mask, nxt);
                }
            }
            batch_sync(ctl, 6, L);
            if (dense)
                compact_dense(a, nxt, (int)((L + 1) & 1));
            else
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
                                                           const u64 *__restrict__ level_start,
                                                           const BatchCtl *__restrict__ ctl, int r,
                                                           const u32 *__restrict__ iperm, i64 *__restrict__ depth) {
    const u32 levels = ctl->levels;
    const u64 nrec = level_start[levels];
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nrec; i += (u64)gridDim.x * blockDim.x) {
        if (!((lm[i] >> r) & 1)) continue;
        u32 lo = 0, hi = levels;  // level L with level_start[L] <= i < level_start[L + 1]
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) / 2;
            if (level_start[mid] <= i) lo = mid;
            else hi = mid;
        }
        depth[orig_id(iperm, lv[i])] = lo;
    }
}

bool bfs_batch_enqueue(const Oned_Graph *g, Bfs_Batch_State *b, const u64 *roots, i32 nroots, hipStream_t stream,
                       int grid) {
    Context &c = ctx();
    if (nroots < 1 || nroots > 64) {
        fail("batched BFS takes 1..64 roots (got %d)", nroots);
        return false;
    }
    for (i32 i = 0; i < nroots; ++i)
        if (roots[i] >= g->nglobalverts) {
            fail("BFS root %llu is outside graph vertex range [0, %llu)", (unsigned long long)roots[i],
                 (unsigned long long)g->nglobalverts);
            return false;
        }
    const u64 n = g->nlocalverts;
    if (n > b->max_nlocalverts) {
        fail("batched BFS state holds %llu vertices, graph has %llu", (unsigned long long)b->max_nlocalverts,
             (unsigned long long)n);
        return false;
    }
    const bool library = stream == nullptr;
    const hipStream_t s = library ? c.stream : stream;
    G500_HIP(hipMemsetAsync(b->ctl, 0, sizeof(BatchCtl), s));
    G500_HIP(hipMemsetAsync(b->seen, 0, n * sizeof(u64), s));
    G500_HIP(hipMemsetAsync(b->next, 0, n * sizeof(u64), s));
    BatchArgs a;
    a.n = n;
    a.m_total = g->nlocaledges;
    const TraversalView tv = traversal_view(g);
    a.rs = tv.rs;
    a.col = tv.col;
    a.perm = tv.perm;
    a.seen = b->seen;
    a.fm[0] = b->fmask[0];
    a.fm[1] = b->fmask[1];
    a.fb[0] = b->fbits[0];
    a.fb[1] = b->fbits[1];
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
    memset(a.roots, 0, sizeof(a.roots));
    for (i32 i = 0; i < nroots; ++i) a.roots[i] = roots[i];
    a.nroots = nroots;
    a.alpha = (f32)c.batch_alpha;
    a.beta = (f32)c.batch_beta;
    a.hub_degree = (u32)c.bfs_hub_degree;
    a.heavy_degree = (u32)c.batch_heavy_degree;
    const int block = ctx().batch_bfs_block;
    if (grid <= 0) grid = cooperative_blocks((const void *)batch_bfs_kernel, block, 0);
    if (failed()) return false;
    void *args[] = {&a};
    b->nroots = nroots;
    if (library) {
        timer_start();
        launch_persistent((const void *)batch_bfs_kernel, grid, block, args);
        timer_stop();
    } else {
        launch_persistent((const void *)batch_bfs_kernel, grid, block, args, s, false);
    }
    return !failed();
}

void bfs_batch_depth_enqueue(const Oned_Graph *g, const Bfs_Batch_State *b, i32 r, i64 *depth, hipStream_t stream) {
    Context &c = ctx();
    const hipStream_t s = stream ? stream : c.stream;
    const u64 n = g->nlocalverts;
    batch_depth_fill<<<grid_for(n, 256, (u64)c.num_cus * 32), 256, 0, s>>>(depth, n);
    batch_depth_scatter<<<(unsigned)c.num_cus * 32, 256, 0, s>>>(b->log_vertex, b->log_mask, b->level_start,
                                                                  (const BatchCtl *)b->ctl, r,
                                                                  traversal_view(g).iperm, depth);
    G500_HIP(hipGetLastError());
}

u32 bfs_batch_error(const Bfs_Batch_State *b, hipStream_t stream) {
    u32 e = 0;
    G500_HIP(hipMemcpyAsync(&e, &((const BatchCtl *)b->ctl)->bar.error, sizeof(u32), hipMemcpyDeviceToHost,
                            stream ? stream : ctx().stream));
    G500_HIP(hipStreamSynchronize(stream ? stream : ctx().stream));
    return e;
}

int bfs_batch_grid() { return cooperative_blocks((const void *)batch_bfs_kernel, ctx().batch_bfs_block, 0); }

// Per-root summary of a finished batch: reached vertices and depth sum.
__global__ __launch_bounds__(256) void batch_summary_kernel(const u64 *__restrict__ lm,
                                                            const u64 *__restrict__ level_start,
                                                            const BatchCtl *__restrict__ ctl, u64 *__restrict__ out) {
    __shared__ unsigned long long cnt[64], sum[64];
    if (threadIdx.x < 64) {
        cnt[threadIdx.x] = 0;
        sum[threadIdx.x] = 0;
    }
    __syncthreads();
    const u32 levels = ctl->levels;
    const u64 nrec = level_start[levels];
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nrec; i += (u64)gridDim.x * blockDim.x) {
        u32 lo = 0, hi = levels;
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) / 2;
            if (level_start[mid] <= i) lo = mid;
            else hi = mid;
        }
        for (u64 m = lm[i]; m; m &= m - 1) {
            const int r = __ffsll((unsigned long long)m) - 1;
            atomicAdd(&cnt[r], 1ull);
            atomicAdd(&sum[r], (unsigned long long)lo);
        }
    }
    __syncthreads();
    if (threadIdx.x < 64 && cnt[threadIdx.x]) {
        atomicAdd((unsigned long long *)&out[threadIdx.x], cnt[threadIdx.x]);
        atomicAdd((unsigned long long *)&out[64 + threadIdx.x], sum[threadIdx.x]);
    }
}

void bfs_batch_summary_enqueue(const Bfs_Batch_State *b, u64 *out, hipStream_t stream) {
    Context &c = ctx();
    const hipStream_t s = stream ? stream : c.stream;
    G500_HIP(hipMemsetAsync(out, 0, 128 * sizeof(u64), s));
    batch_summary_kernel<<<(unsigned)c.num_cus * 4, 256, 0, s>>>(b->log_mask, b->level_start,
                                                                 (const BatchCtl *)b->ctl, out);
    G500_HIP(hipGetLastError());
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
    const u64 nwords = (n + 63) / 64 + 1;
    b.fbits[0] = (u64 *)arena_alloc(ws, nwords * sizeof(u64), "batch frontier bitmaps");
    b.fbits[1] = (u64 *)arena_alloc(ws, nwords * sizeof(u64), "batch frontier bitmaps");
    b.next = (u64 *)arena_alloc(ws, n * sizeof(u64), "batch next masks");
    b.touched = (u32 *)arena_alloc(ws, n * sizeof(u32), "batch touched");
    b.log_vertex = (u32 *)arena_alloc(ws, b.log_capacity * sizeof(u32), "batch log vertices");
    b.log_mask = (u64 *)arena_alloc(ws, b.log_capacity * sizeof(u64), "batch log masks");
    b.level_start = (u64 *)arena_alloc(ws, (n + 2) * sizeof(u64), "batch level starts");
    b.hubs = (u32 *)arena_alloc(ws, 3 * n * sizeof(u32), "batch hubs");
    b.hub_offsets = (u64 *)arena_alloc(ws, 3 * n * sizeof(u64), "batch hub offsets");
    b.hub_lengths = (u32 *)arena_alloc(ws, 3 * n * sizeof(u32), "batch hub lengths");
    b.ctl = arena_alloc(ws, sizeof(BatchCtl), "batch control");
    return b;
}

void oned_graph_bfs_batch_run(Worker_State *ws, const Oned_Graph *g, Bfs_Batch_State *b, const u64 *roots,
                              i32 nroots) {
    (void)ws;
    if (!bfs_batch_enqueue(g, b, roots, nroots, nullptr, 0)) return;
    static BatchCtl h;
    g500_copy_to_host(&h, b->ctl, sizeof(h));
    b->levels = (i32)h.levels;
    if (h.bar.error)
        fail("batched BFS kernel error %u (1 barrier timeout, 2 log or hub overflow)", h.bar.error);
}

void oned_graph_bfs_batch_depth(Worker_State *ws, const Oned_Graph *g, const Bfs_Batch_State *b, i32 r,
                                i64 *depth) {
    (void)ws;
    if (r < 0 || r >= b->nroots) {
        fail("root index %d outside the last batch of %d roots", r, b->nroots);
        return;
    }
    bfs_batch_depth_enqueue(g, b, r, depth, nullptr);
    sync();
}

i32 g500_bfs_batch_level_log(const Bfs_Batch_State *b, i32 *modes, u64 *records, u64 *edges, f64 *ms,
                             i32 max_levels) {
    static BatchCtl h;
    g500_copy_to_host(&h, b->ctl, sizeof(h));
    const i32 L = (i32)MIN((u32)h.levels, (u32)kBatchLevelLog);
    for (i32 i = 0; i < L && i < max_levels; ++i) {
        if (modes) modes[i] = h.mode_log[i];
        if (records) records[i] = h.size_log[i];
        if (edges) edges[i] = h.edge_log[i];
        if (ms) ms[i] = (f64)h.tick_log[i] / ctx().wall_clock_khz;
    }
    return L;
}

void oned_graph_bfs_batch_destroy(Bfs_Batch_State *b) {
    if (b) memset(b, 0, sizeof(*b));
}

}  // extern "C"
