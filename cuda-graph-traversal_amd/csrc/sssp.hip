// Delta-stepping SSSP as one persistent cooperative kernel per root.
// Reference: src/traversal.cu:209-340 (kernels), :570-716 (state, host bucket
// loop with delta = 0.1f).
//
// Labels: one u64 per vertex, (f32 distance bits << 32) | parent, and a u32
// probe per vertex with the distance bits alone (the state's dist array, in
// traversal ids while the kernel runs). Distances are non-negative, so their
// bit patterns order like the floats. A relaxation reads the probe, lowers it
// with a 32-bit atomic min, and only when that strictly lowered it commits
// (distance, parent) to the label with a 64-bit atomic min (the reference's
// CAS plus per-vertex spin lock, traversal.cu:211-240). Each strict lowering
// of a probe is by a distinct value, so two label commits never tie and the
// 64-bit min keeps the commit with the smaller distance: the label ends as
// (final distance, a parent that achieved it). Failed relaxations, the large
// majority, touch only the probe array, half the label array's size. Every
// terminating label-correcting schedule reaches the same FP32 fixed point, so
// the schedule below does not change any distance.
//
// Buckets follow the reference's bounds: [lo, hi) with hi = fl(lo + delta)
// (traversal.cu:693-694). Rows are sorted by weight, so the light edges
// (w < delta) of a row are its first split[u] entries. Because rounding is
// monotone, a heavy edge from u (d_u >= lo, w >= delta) yields
// fl(d_u + w) >= fl(lo + delta) = hi: heavy relaxations never land back in
// the current bucket. Per bucket:
//   light steps: queued vertices relax their light edges; vertices lowered
//     below hi are queued for the next step (de-duplicated per queue);
//   one fused dense step: every vertex of the bucket relaxes its heavy edges
//     (short rows inline, longer rows through a wave list and a hub list in a
//     second phase), and every vertex already in the next bucket
//     [hi, fl(hi + delta)) is queued, together with those that the heavy
//     relaxations lower into it;
//   when the next bucket is empty: a min step finds the smallest remaining
//     distance and a collect step queues the bucket that holds it.
// The fused step also marks the bucket's vertices settled in a bitmap; every
// later candidate is >= the current bucket floor > a settled distance, so
// relaxations into settled vertices skip the label (the bitmap is 1/64 of the
// label array and stays in cache).
// Pull form of a fused step, for buckets that hold a large share of the
// unsettled graph (the first bucket of a Kronecker graph reaches millions of
// vertices through light entries of hubs): instead of every vertex of the
// bucket pushing its heavy entries, every unsettled vertex v looks for the
// bucket's vertices among the heavy entries of its own row, lightest first.
// The SSSP rows are symmetric, so these are the same relaxations. The scan of
// v stops at the first entry with fl(lo + w) >= v's best distance so far:
// every later candidate fl(d_u + w') has d_u >= lo and w' >= w, and rounding
// is monotone. Rows up to kPullLane heavy entries are scanned by one lane,
// longer ones by a wave in chunks of 64.
// Work is mapped by row-segment length: a lane per vertex (<= 32 entries), a
// wave per vertex (<= 1024), or chunks of a flattened hub work space spread
// over the whole grid (longer).
//
// Steps and shared counters: step x reads the queue written by step x - 1
// (slot x % 3), writes only slot (x + 1) % 3 and clears slot (x + 2) % 3, so
// a block acting on a decision never changes a counter that a slower block
// still has to read for the same decision.
#include <algorithm>
#include <vector>

#include "sssp_common.hpp"

namespace g500 {

constexpr u32 kSmall = 32;
constexpr u64 kPullLane = 32;
constexpr u32 kHub = 1024;
constexpr u32 kStepLog = 1024;

// Error bits: 1 barrier timeout, 2 hub list overflow, 4 queue entry out of
// range, 8 queue length above n, 16 bucket invariant violated, 32 a distance
// at or above the pruned SSSP copy's bound (the copy may then be inexact).
struct SsspCtl {
    GridBarrier bar;
    // every counter that waves append to has a 256-byte line of its own
    CounterLine<u32> flen[3][2];   // light queue lengths [slot][tier 0-1]
    CounterLine<u64> fhub[3];      // light queue hub counters [slot]
    CounterLine<u32> hlen[3];      // heavy wave-tier list length [slot]
    CounterLine<u64> hhub[3];      // heavy hub-tier counters [slot]
    alignas(256) u32 nmin[3];      // smallest distance bits >= hi, per slot
    u32 nnext[3];                  // 1: a label entered [hi, hi + delta) (the bucket is not empty), per slot
    alignas(256) u64 nset[3];      // vertices settled by a fused step, per slot
    alignas(256) u64 buckets, light_iters, improvements, scans, pulls;
    u64 ticks[4];                  // constant-clock ticks of block 0: light, fused, min+collect, kernel
    // per-step log (block 0): kind (0 light, 1 fused, 2 min, 3 collect), queue
    // entries of tiers 0/1, packed hub counter, constant-clock ticks
    u32 step_kind[kStepLog], step_n0[kStepLog], step_n1[kStepLog];
    u64 step_hub[kStepLog], step_ticks[kStepLog];
    u32 nsteps;
    alignas(256) BlockTrace trace[kMaxTraceBlocks];
};

enum : u32 { kPhaseInit = 1, kPhaseLight = 2, kPhaseFused = 3, kPhaseHeavy = 4, kPhaseMin = 5, kPhaseCollect = 6 };

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

// A row's bounds in one 16-byte record (one memory transaction per item):
// row start, light entries (w < delta), row length.
struct RowRec {
    u64 start;
    u32 light, len;
};
static_assert(sizeof(RowRec) == 16, "one 16-byte record per row");

__device__ __forceinline__ RowRec load_rec(const RowRec *r, u32 v) {
    const uint4 x = *reinterpret_cast<const uint4 *>(r + v);
    return RowRec{((u64)x.y << 32) | x.x, x.z, x.w};
}

struct SsspArgs {
    u64 n;
    f32 dbound;  // distances at or above it set error 32 (+inf without the pruned copy)
    const u64 *rs;
    const u32 *col;
    const f32 *w;
    const u32 *split;
    const RowRec *rows;       // [n] row records for the same split
    const u32 *perm, *iperm;  // traversal <-> original ids (null: identity)
    u64 *lab;
    u32 *dl;         // probe: distance bits per vertex (aliases the dist output)
    u32 *fq[2][3];   // light queues [buffer][tier]; tier 2 holds hub vertices
    u64 *fq_off[2];  // hub-tier work offsets per buffer
    u32 *fq_len[2];  // hub-tier segment lengths per buffer
    u32 *hq1;        // heavy wave-tier list
    u32 *hqh;        // heavy hub-tier list
    u64 *hq_off;
    u32 *hq_len;
    u32 *stamp;
    u64 *settled;    // bit per vertex: distance final (its bucket is done)
    u64 *reached;    // bit per vertex: label finite (set on the first relaxation)
    u64 nact;        // vertices that can have entries: [0, nact) (the degree order puts isolated ones last)
    f32 pull_ratio;  // pull a bucket's heavy relaxations when its light items >= pull_ratio * unsettled (0: never)
    SsspCtl *ctl;
    i64 *pred;
    f32 *dist;
    u64 root;
    f32 delta;
    u32 screen;      // 1: relaxations skip targets marked in `settled` before probing them
};

__device__ __forceinline__ int tier_of(u32 len) { return len <= kSmall ? 0 : (len <= kHub ? 1 : 2); }

// Where the pushes of a step go: queue buffer, counter slot and tag of the
// queue, and the bound below which a lowered vertex is queued.
struct Target {
    int buf, slot;
    u32 tag;
    f32 below;   // queue v when its new distance is < below
    f32 floor;   // new distances below this would break the bucket invariant
};

// Wave-staged appends to the light queue tiers 0-1 and the heavy wave-tier
// list: a wave buffers its entries in LDS and writes a run with one atomic
// when the buffer is full; at the end of a phase the block sums its waves'
// counts and makes one atomic per list. The counters then take one atomic per
// block and phase instead of one per wave and relaxation round.
constexpr u32 kSCap = 64;
constexpr int kSLists = 3;  // light tier 0, light tier 1, heavy wave tier
struct StageLds {
    u32 v[1024 / kWave][kSLists][kSCap];
    u32 cnt[1024 / kWave][kSLists];
};
struct Stage {
    StageLds *L;
    int w;            // wave of the block
    u32 n[kSLists];   // staged entries per list
    int buf, slot;    // light queue buffer and counter slot of the phase's pushes
    int hs;           // heavy list slot of a fused step
};

__device__ __forceinline__ u32 *stage_dst(const SsspArgs &a, const Stage &q, int l) {
    if (l == 2) return a.hq1;
    return q.buf ? a.fq[1][l] : a.fq[0][l];  // a select: a runtime index would copy the arguments to scratch
}
__device__ __forceinline__ u32 *stage_counter(const SsspArgs &a, const Stage &q, int l) {
    return l == 2 ? &a.ctl->hlen[q.hs].v : &a.ctl->flen[q.slot][l].v;
}
__device__ __forceinline__ void stage_copy(const volatile u32 *buf, u32 cnt, u32 *dst, u32 base, u64 cap) {
    for (u32 i = lane_id(); i < cnt; i += kWave) {
        const u64 o = (u64)base + i;
        if (o < cap) dst[o] = buf[i];
    }
}
template <int L>
__device__ __forceinline__ void stage_push(const SsspArgs &a, Stage &q, bool want, u32 val) {
    const u64 m = ballot(want);
    if (!m) return;
    const u32 k = (u32)__popcll(m);
    volatile u32 *buf = q.L->v[q.w][L];
    if (q.n[L] + k > kSCap) {
        u32 base = 0;
        if (lane_id() == 0) base = atomicAdd(stage_counter(a, q, L), q.n[L]);
        base = __shfl(base, 0);
        stage_copy(buf, q.n[L], stage_dst(a, q, L), base, a.n);
        q.n[L] = 0;
    }
    if (want) buf[q.n[L] + mask_rank(m)] = val;
    q.n[L] += k;
}
// End of a phase, called by every thread of the block.
__device__ __forceinline__ void stage_flush(const SsspArgs &a, Stage &q) {
    volatile u32 *cnt = &q.L->cnt[0][0];
    const int nw = (int)(blockDim.x / kWave);
    if (lane_id() == 0) {
#pragma unroll
        for (int l = 0; l < kSLists; ++l) cnt[q.w * kSLists + l] = q.n[l];
    }
    __syncthreads();
    if (threadIdx.x < (u32)kSLists) {
        const int l = (int)threadIdx.x;
        u32 tot = 0;
        for (int i = 0; i < nw; ++i) {
            const u32 c = cnt[i * kSLists + l];
            cnt[i * kSLists + l] = tot;
            tot += c;
        }
        u32 b = 0;
        if (tot) b = atomicAdd(stage_counter(a, q, l), tot);
        for (int i = 0; i < nw; ++i) cnt[i * kSLists + l] += b;
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < kSLists; ++l) {
        if (q.n[l]) stage_copy((volatile u32 *)q.L->v[q.w][l], q.n[l], stage_dst(a, q, l), cnt[q.w * kSLists + l], a.n);
        q.n[l] = 0;
    }
    __syncthreads();
}

// Wave-converged: queues v (lanes with `want`) into the light queue of `t`,
// once per queue (stamp tag), in the work tier of its light segment.
__device__ __forceinline__ void push_light(const SsspArgs &a, bool want, u32 v, const Target &t, Stage &sq) {
    const u64 wm = ballot(want);
    if (!wm) return;
    SsspCtl *ctl = a.ctl;
    // the flag only tells that the next bucket is not empty: set it once
    if (lane_id() == (u32)(__ffsll((unsigned long long)wm) - 1) && !ld_relaxed(&ctl->nnext[t.slot]))
        st_relaxed(&ctl->nnext[t.slot], 1u);
    // a vertex without light entries has nothing to do in light steps
    bool q = false;
    u32 lw = 0;
    if (want) {
        lw = a.split[v];
        if (lw) q = atomicExch(&a.stamp[v], t.tag) != t.tag;
    }
    if (!ballot(q)) return;
    stage_push<0>(a, sq, q && tier_of(lw) == 0, v);
    stage_push<1>(a, sq, q && tier_of(lw) == 1, v);
    hub_append(&ctl->fhub[t.slot].v, a.fq[t.buf][2], a.fq_off[t.buf], a.fq_len[t.buf], q && tier_of(lw) == 2, v, lw,
               a.n, &ctl->bar.error);
}

// kQ relaxations per lane, in phases (columns and weights, settled bits,
// probes, atomic mins, pushes) so that each wave keeps kQ independent chains
// of loads in flight; then the wave-converged pushes.
constexpr int kQ = 2;

__device__ __forceinline__ void visit_edges(const SsspArgs &a, const Target &t, const bool (&act)[kQ],
                                            const u32 (&u)[kQ], const f32 (&du)[kQ], const u64 (&k)[kQ],
                                            u64 &improved, Stage &sq) {
    u32 v[kQ];
    f32 cand[kQ];
    bool live[kQ];
    // Branch-free loads (an inactive lane reads entry 0, valid whenever a
    // segment is walked, or vertex 0, and its result is ignored), so that the
    // kQ chains are in flight together.
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        live[q] = act[q];
        const u64 kk = live[q] ? k[q] : 0;
        v[q] = a.col[kk];
        cand[q] = du[q] + a.w[kk];
    }
    // a settled vertex has d_v < floor <= cand: skip its label
    if (a.screen) {
#pragma unroll
        for (int q = 0; q < kQ; ++q) live[q] = live[q] && !((a.settled[v[q] >> 6] >> (v[q] & 63)) & 1);
    }
    // a plain read may be stale, i.e. larger: it only lets an atomic through
    u32 old[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) old[q] = a.dl[live[q] ? v[q] : 0u];
    bool push[kQ], first[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const u32 cb = __float_as_uint(cand[q]);
        bool ok = false;
        if (live[q] && old[q] > cb) {
            const u32 prev = atomicMin(&a.dl[v[q]], cb);
            ok = prev > cb;
            if (ok) {
                atomicMin((unsigned long long *)&a.lab[v[q]], ((unsigned long long)cb << 32) | u[q]);
                old[q] = prev;
            }
        }
        first[q] = ok && old[q] >= kInfBits;
        push[q] = ok && cand[q] < t.below;
        if (ok) {
            ++improved;
            if (cand[q] < t.floor) atomicOr(&a.ctl->bar.error, 16u);
        }
    }
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        if (first[q]) atomicOr((unsigned long long *)&a.reached[v[q] >> 6], 1ull << (v[q] & 63));
        push_light(a, push[q], v[q], t, sq);
    }
}

// Segment of u's row: the light prefix or the heavy suffix.
template <bool LIGHT>
__device__ __forceinline__ void segment(const SsspArgs &a, u32 u, u64 &b, u64 &e) {
    const RowRec r = load_rec(a.rows, u);
    b = LIGHT ? r.start : r.start + r.light;
    e = LIGHT ? r.start + r.light : r.start + r.len;
}

// Relaxes the concatenation of the wave's 64 segments (lane j: entries
// [b, b + len) of u's row, u at distance du), 64 entries per step, so short
// and uneven segments keep every lane busy.
__device__ __forceinline__ void walk_segments(const SsspArgs &a, const Target &t, u32 u, f32 du, u64 b, u32 len,
                                              u64 &improved, volatile int *slot, Stage &sq) {
    if (!ballot(len != 0)) return;
    const u32 incl = wave_inclusive_scan(len);
    const u32 off = incl - len, total = __shfl(incl, kWave - 1);
    const u32 lane = lane_id();
    for (u32 c = 0; c < total; c += kQ * kWave) {
        bool act[kQ];
        u32 uu[kQ];
        f32 dd[kQ];
        u64 kk[kQ];
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const u32 cq = c + (u32)q * kWave;
            act[q] = false;
            uu[q] = 0;
            dd[q] = 0.f;
            kk[q] = 0;
            if (cq < total) {
                bool inter;
                int last;
                int j = chunk_owner(slot, off, len, cq, inter, last);
                if (j < 0) j = 0;
                const u32 offj = __shfl(off, j);
                uu[q] = __shfl(u, j);
                dd[q] = __shfl(du, j);
                const u64 bj = shfl64(b, j);
                const u32 pos = cq + lane;
                act[q] = pos < total;
                kk[q] = bj + (pos - offj);
            }
        }
        visit_edges(a, t, act, uu, dd, kk, improved, sq);
    }
}

// Vertices list[0, count) of work tier 0 (64 vertices per wave, segments
// walked together) or 1 (a wave each).
template <bool LIGHT>
__device__ __forceinline__ void process_tier(const SsspArgs &a, const u32 *list, u32 count, int tier,
                                             const Target &t, u64 &improved, volatile int *slot, Stage &sq) {
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u32 lane = lane_id();
    const u64 wave_gid = gtid / kWave, nwaves = gthreads / kWave;
    if (tier == 0) {
        // 64 vertices per wave, or fewer when 64-vertex chunks would leave
        // waves idle (a short list then spreads over the whole grid)
        const u64 per = count >= nwaves * kWave ? (u64)kWave : max((u64)1, ((u64)count + nwaves - 1) / nwaves);
        for (u64 i0 = wave_gid * per; i0 < count; i0 += nwaves * per) {
            const u64 i = i0 + lane;
            u64 b = 0, e = 0;
            u32 u = 0;
            f32 du = 0.f;
            if (lane < per && i < count) {
                u = list[i];
                if (u < a.n) {
                    segment<LIGHT>(a, u, b, e);
                    du = __uint_as_float(a.dl[u]);
                } else {
                    atomicOr(&a.ctl->bar.error, 4u);
                }
            }
            walk_segments(a, t, u, du, b, (u32)(e - b), improved, slot, sq);
        }
    } else {
        for (u64 i = wave_gid; i < count; i += nwaves) {
            const u32 u = list[i];
            if (u >= a.n) {
                atomicOr(&a.ctl->bar.error, 4u);
                continue;
            }
            u64 b, e;
            segment<LIGHT>(a, u, b, e);
            const f32 du = __uint_as_float(a.dl[u]);
            for (u64 k0 = b; k0 < e; k0 += kQ * kWave) {
                bool act[kQ];
                u32 uu[kQ];
                f32 dd[kQ];
                u64 kk[kQ];
#pragma unroll
                for (int q = 0; q < kQ; ++q) {
                    kk[q] = k0 + (u64)q * kWave + lane;
                    act[q] = kk[q] < e;
                    uu[q] = u;
                    dd[q] = du;
                }
                visit_edges(a, t, act, uu, dd, kk, improved, sq);
            }
        }
    }
}

// Hub tier: the flattened work space [0, total) of hubs [0, count); each wave
// takes kHubChunk-edge chunks, so a long row is spread over the whole grid.
template <bool LIGHT>
__device__ __forceinline__ void process_hubs(const SsspArgs &a, const u32 *hv, const u64 *hoff, const u32 *hlen,
                                             u64 packed, const Target &t, u64 &improved, Stage &sq) {
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u32 lane = lane_id();
    const u64 wave_gid = gtid / kWave, nwaves = gthreads / kWave;
    const u64 count = hub_count(packed), total = hub_total(packed);
    u64 i = 0;
    for (u64 c = wave_gid; c * kHubChunk < total; c += nwaves) {
        const u64 p = c * kHubChunk;
        i = hub_locate(hoff, i, count, p);
        const u32 u = hv[i];
        if (u >= a.n) {
            atomicOr(&a.ctl->bar.error, 4u);
            continue;
        }
        const u64 local = p - hoff[i];
        const u32 len = hlen[i];
        u64 b, e;
        segment<LIGHT>(a, u, b, e);
        const f32 du = __uint_as_float(a.dl[u]);
        static_assert(kHubChunk % ((u64)kQ * kWave) == 0, "a hub chunk is whole rounds of visit_edges");
#pragma unroll
        for (u64 j = 0; j < kHubChunk; j += (u64)kQ * kWave) {
            bool act[kQ];
            u32 uu[kQ];
            f32 dd[kQ];
            u64 kk[kQ];
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const u64 off = local + j + (u64)q * kWave + lane;
                act[q] = off < len;
                kk[q] = b + off;
                uu[q] = u;
                dd[q] = du;
            }
            visit_edges(a, t, act, uu, dd, kk, improved, sq);
        }
    }
}

// Calls f(v, probe distance bits) for every reached vertex that is not
// settled, with the whole wave converged: each lane reads one 64-vertex word
// of reached & ~settled, a ballot picks the words with such vertices, and
// the wave walks them one word at a time (lane = vertex of the word; lanes of
// other vertices get +inf bits). Words in which everything is
// settled or unreached cost one load per 64 of them.
template <typename F>
__device__ __forceinline__ void scan_open(const SsspArgs &a, F &&f) {
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u32 lane = lane_id();
    const u64 wave_gid = gtid / kWave, nwaves = gthreads / kWave;
    const u64 nwords = (a.n + 63) / 64;
    // Words are dealt to waves round-robin (lane l of wave g reads word
    // base + l * nwaves + g): open vertices cluster in the degree order, and
    // runs of 64 consecutive words per wave would leave most waves idle while
    // a few walk dense runs.
    for (u64 base = 0; base < nwords; base += nwaves * kWave) {
        const u64 w = base + (u64)lane * nwaves + wave_gid;
        const u64 open_l = w < nwords ? (a.reached[w] & ~a.settled[w]) : 0ull;
        u64 todo = ballot(open_l != 0);
        // software pipelined: the next word's labels load while f runs
        auto word_of = [&](int j) { return base + (u64)j * nwaves + wave_gid; };
        auto take = [&](int &j, u32 &lbits) {
            j = todo ? __ffsll((unsigned long long)todo) - 1 : -1;
            todo &= todo - 1;
            const u64 open = shfl64(open_l, j < 0 ? 0 : j);
            const u64 v = word_of(j < 0 ? 0 : j) * 64 + lane;
            lbits = (j >= 0 && ((open >> lane) & 1)) ? a.dl[v] : kInfBits;
        };
        int j;
        u32 lbits;
        take(j, lbits);
        while (j >= 0) {
            int jn;
            u32 ln;
            take(jn, ln);
            f(word_of(j) * 64 + lane, lbits);
            j = jn;
            lbits = ln;
        }
    }
}

// Pull form of fused step x for the bucket [lo, hi): settles the bucket and
// lowers every other unsettled vertex to its best candidate through a heavy
// entry from the bucket; vertices ending in [hi, hi2) are queued for the next
// bucket. Two phases: a lane per vertex (rows of up to kPullLane heavy
// entries; longer rows go to the wave list), then a wave per listed vertex.
// Counters: settled vertices into nset[(x + 1) % 3], the wave list in
// hlen[(x + 1) % 3]. Ends after the second barrier.
__device__ __forceinline__ void pull_step(const SsspArgs &a, u32 x, f32 lo, f32 hi, f32 hi2, u64 &improved,
                                       Stage &sq) {
    SsspCtl *ctl = a.ctl;
    const u32 lane = lane_id();
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u64 wave_gid = gtid / kWave, nwaves = gthreads / kWave;
    const int hs = (int)((x + 1) % 3);
    const Target t{(int)((x + 1) & 1), hs, x + 2, hi2, hi};
    sq.buf = t.buf;
    sq.slot = t.slot;
    sq.hs = hs;
    const u32 lo_b = __float_as_uint(lo), hi_b = __float_as_uint(hi), hi2_b = __float_as_uint(hi2);
    u64 nset = 0;
    for (u64 v0 = wave_gid * kWave; v0 < a.nact; v0 += nwaves * kWave) {  // the wave owns the word of v0
        const u64 v = v0 + lane, word = v0 >> 6;
        const u64 sw = a.settled[word];
        const bool open = v < a.nact && !((sw >> lane) & 1);
        const u32 lb = open ? a.dl[v] : kInfBits;
        const bool mine = open && lb >= lo_b && lb < hi_b;
        const u64 sm = ballot(mine);
        if (sm && lane == 0) {
            atomicOr((unsigned long long *)&a.settled[word], (unsigned long long)sm);
            nset += (u64)__popcll(sm);
        }
        const bool act = open && !mine;
        u64 b = 0, e = 0;
        if (act) segment<false>(a, (u32)v, b, e);
        const bool longrow = act && e - b > kPullLane;
        stage_push<2>(a, sq, longrow, (u32)v);
        u32 best = lb, bu = 0;
        if (act && !longrow) {
            for (u64 k = b; k < e; ++k) {
                const f32 wk = a.w[k];
                if (__float_as_uint(__fadd_rn(lo, wk)) >= best) break;
                const u32 u = a.col[k];
                const u32 du = a.dl[u];
                if (du >= lo_b && du < hi_b) {
                    const u32 c = __float_as_uint(__fadd_rn(__uint_as_float(du), wk));
                    if (c < best) {
                        best = c;
                        bu = u;
                    }
                }
            }
        }
        const bool imp = best < lb;
        if (imp) {
            a.dl[v] = best;
            a.lab[v] = ((u64)best << 32) | bu;
            ++improved;
        }
        const u64 rm = ballot(imp && lb >= kInfBits);
        if (rm && lane == 0) atomicOr((unsigned long long *)&a.reached[word], (unsigned long long)rm);
        push_light(a, act && !longrow && best >= hi_b && best < hi2_b, (u32)v, t, sq);
    }
    stage_flush(a, sq);
    {  // the block sums its waves' settled counts and makes one atomic (every thread is here)
        __shared__ u64 bset;
        if (threadIdx.x == 0) bset = 0;
        __syncthreads();
        if (lane == 0 && nset) atomicAdd((unsigned long long *)&bset, (unsigned long long)nset);
        __syncthreads();
        if (threadIdx.x == 0 && bset) atomicAdd((unsigned long long *)&ctl->nset[hs], (unsigned long long)bset);
    }
    sync_traced(ctl, kPhaseFused, lo_b, x);
    // long rows: a wave per vertex, 64 entries per chunk, lightest chunk first
    const u32 nlong = min(ld_relaxed(&ctl->hlen[hs].v), (u32)a.n);
    for (u64 i = wave_gid; i < nlong; i += nwaves) {
        const u32 v = a.hq1[i];
        if (v >= a.n) {
            if (lane == 0) atomicOr(&ctl->bar.error, 4u);
            continue;
        }
        const u32 lb = a.dl[v];
        u64 b, e;
        segment<false>(a, v, b, e);
        u64 best = ((u64)lb << 32) | 0xFFFFFFFFull;  // (distance bits, parent)
        for (u64 k0 = b; k0 < e; k0 += kWave) {
            if (__float_as_uint(__fadd_rn(lo, a.w[k0])) >= (u32)(best >> 32)) break;
            const u64 k = k0 + lane;
            u64 c = ~0ull;
            if (k < e) {
                const f32 wk = a.w[k];
                const u32 u = a.col[k];
                const u32 du = a.dl[u];
                if (du >= lo_b && du < hi_b) c = ((u64)__float_as_uint(__fadd_rn(__uint_as_float(du), wk)) << 32) | u;
            }
            best = wave_min64(best < c ? best : c);
        }
        const u32 bb = (u32)(best >> 32);
        if (lane == 0 && bb < lb) {
            a.dl[v] = bb;
            a.lab[v] = best;
            ++improved;
            if (lb >= kInfBits) atomicOr((unsigned long long *)&a.reached[v >> 6], 1ull << (v & 63));
        }
        push_light(a, lane == 0 && bb >= hi_b && bb < hi2_b, v, t, sq);
    }
    stage_flush(a, sq);
    sync_traced(ctl, kPhaseHeavy, lo_b, x);
}

__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void sssp_kernel(SsspArgs a) {
    SsspCtl *ctl = a.ctl;
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u32 lane = lane_id();
    __shared__ int slots[1024 / kWave][kWave];
    volatile int *slot = slots[threadIdx.x / kWave];
    __shared__ StageLds slds;
    Stage sq;
    sq.L = &slds;
    sq.w = (int)(threadIdx.x / kWave);
#pragma unroll
    for (int l = 0; l < kSLists; ++l) sq.n[l] = 0;
    sq.buf = sq.slot = sq.hs = 0;
    auto stage_to = [&](const Target &t) {
        sq.buf = t.buf;
        sq.slot = t.slot;
    };
    u64 improved = 0;
    u64 bucket_items = 0;            // light-queue items of the current bucket (the same in every thread)
    u64 settled_total = 0;           // vertices settled so far (the same in every thread)
    u32 x = 0;                       // step
    f32 lo = 0.0f, hi = next_bound(0.0f, a.delta);

    const bool timer = gtid == 0;
    u64 t_prev = timer ? wall_clock64() : 0, t_start = t_prev, t_acc[3] = {0, 0, 0};
    auto tick = [&](int phase) {
        if (timer) {
            const u64 t = wall_clock64();
            t_acc[phase] += t - t_prev;
            if (x < kStepLog) st_relaxed(&ctl->step_ticks[x], t - t_prev);
            t_prev = t;
        }
    };
    auto log_step = [&](u32 kind, u32 n0, u32 n1, u64 hubp) {
        if (timer && x < kStepLog) {
            st_relaxed(&ctl->step_kind[x], kind);
            st_relaxed(&ctl->step_n0[x], n0);
            st_relaxed(&ctl->step_n1[x], n1);
            st_relaxed(&ctl->step_hub[x], hubp);
        }
    };
    auto clear_ahead = [&](u32 step) {
        if (gtid == 0) {
            const int s = (int)((step + 2) % 3);
            st_relaxed(&ctl->flen[s][0].v, 0u);
            st_relaxed(&ctl->flen[s][1].v, 0u);
            st_relaxed(&ctl->fhub[s].v, 0ull);
            st_relaxed(&ctl->hlen[s].v, 0u);
            st_relaxed(&ctl->hhub[s].v, 0ull);
            st_relaxed(&ctl->nmin[s], 0xFFFFFFFFu);
            st_relaxed(&ctl->nnext[s], 0u);
            st_relaxed(&ctl->nset[s], 0ull);
        }
    };
    // Light queue of step `step`: counts per tier and the packed hub counter.
    auto queue_size = [&](u32 step, u32 *cnt, u64 *hub) {
        u64 total = 0;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            u32 c = ld_relaxed(&ctl->flen[step % 3][t].v);
            if (c > a.n) {
                if (gtid == 0) atomicOr(&ctl->bar.error, 8u);
                c = 0;
            }
            cnt[t] = c;
            total += c;
        }
        *hub = ld_relaxed(&ctl->fhub[step % 3].v);
        return total + hub_count(*hub);
    };
    auto target = [&](u32 step, f32 below, f32 floor) {
        return Target{(int)((step + 1) & 1), (int)((step + 1) % 3), step + 2, below, floor};
    };

    if (gtid == 0) {
        const u32 r = trav_id(a.perm, a.root);
        st_relaxed(&a.lab[r], (u64)r);  // distance 0.0f, parent = root
        st_relaxed(&a.dl[r], 0u);
        atomicOr((unsigned long long *)&a.reached[r >> 6], 1ull << (r & 63));
        for (int i = 0; i < 3; ++i) st_relaxed(&ctl->nmin[i], 0xFFFFFFFFu);
    }
    sync_traced(ctl, kPhaseInit, 0, 0);
    stage_to(target(0, 0.f, 0.f));
    if (gtid < kWave)  // step 0 queues the root
        push_light(a, lane == 0, trav_id(a.perm, a.root), target(0, 0.f, 0.f), sq);
    stage_flush(a, sq);
    sync_traced(ctl, kPhaseInit, 1, 0);
    x = 1;

    for (;;) {  // buckets [lo, hi)
        for (;;) {  // light steps
            u32 cnt[2];
            u64 hub;
            const u64 total = queue_size(x, cnt, &hub);
            clear_ahead(x);
            if (total == 0 || ld_relaxed(&ctl->bar.error)) break;
            bucket_items += total;
            if (gtid == 0) atomicAdd((unsigned long long *)&ctl->light_iters, 1ull);
            log_step(0, cnt[0], cnt[1], hub);
            const Target t = target(x, hi, lo);
            stage_to(t);
#pragma unroll
            for (int tier = 0; tier < 2; ++tier)
                process_tier<true>(a, a.fq[x & 1][tier], cnt[tier], tier, t, improved, slot, sq);
            process_hubs<true>(a, a.fq[x & 1][2], a.fq_off[x & 1], a.fq_len[x & 1], hub, t, improved, sq);
            stage_flush(a, sq);
            sync_traced(ctl, kPhaseLight, __float_as_uint(lo), x);
            tick(0);
            ++x;
        }
        if (ld_relaxed(&ctl->bar.error)) break;
        // fused step x (its light queue was empty): heavy edges of [lo, hi) and
        // the next bucket [hi, hi2) queued
        if (gtid == 0) atomicAdd((unsigned long long *)&ctl->buckets, 1ull);
        log_step(1, 0, 0, 0);
        const f32 hi2 = next_bound(hi, a.delta);
        const u64 unsettled = a.nact > settled_total ? a.nact - settled_total : 0;
        if (a.pull_ratio > 0.f && (f32)bucket_items >= a.pull_ratio * (f32)unsettled) {
            pull_step(a, x, lo, hi, hi2, improved, sq);
            settled_total += ld_relaxed(&ctl->nset[(x + 1) % 3]);
            if (gtid == 0) atomicAdd((unsigned long long *)&ctl->pulls, 1ull);
            tick(1);
            ++x;
        } else {
            const Target t = target(x, hi2, hi);
            const int hs = (int)((x + 1) % 3);
            stage_to(t);
            sq.hs = hs;
            u64 nset_local = 0;
            scan_open(a, [&](u64 v, u32 lbits) {
                bool mine = false, next = false;
                u64 b = 0, e = 0;
                f32 dv = 0.f;
                if (lbits < kInfBits) {
                    const f32 d = __uint_as_float(lbits);
                    if (d >= lo && d < hi) {
                        mine = true;
                        dv = d;
                        segment<false>(a, (u32)v, b, e);
                    } else if (d >= hi && d < hi2) {
                        next = true;
                    }
                }
                if (!ballot(mine || next)) return;  // the word holds only later buckets
                push_light(a, next, (u32)v, t, sq);
                // the bucket's vertices are final now: the wave owns the word of v
                const u64 sm = ballot(mine);
                if (sm && lane == 0) {
                    a.settled[v >> 6] |= sm;
                    nset_local += (u64)__popcll(sm);
                }
                const u32 len = (u32)(e - b);
                const bool medium = mine && len > kSmall && len <= kHub;
                stage_push<2>(a, sq, medium, (u32)v);
                hub_append(&ctl->hhub[hs].v, a.hqh, a.hq_off, a.hq_len, mine && len > kHub, (u32)v, len, a.n,
                           &ctl->bar.error);
                walk_segments(a, t, (u32)v, dv, b, mine && len <= kSmall ? len : 0, improved, slot, sq);
            });
            stage_flush(a, sq);
            {  // the block sums its waves' settled counts and makes one atomic (every thread is here)
                __shared__ u64 bset;
                if (threadIdx.x == 0) bset = 0;
                __syncthreads();
                if (nset_local) atomicAdd((unsigned long long *)&bset, (unsigned long long)nset_local);
                __syncthreads();
                if (threadIdx.x == 0 && bset) atomicAdd((unsigned long long *)&ctl->nset[hs], (unsigned long long)bset);
            }
            sync_traced(ctl, kPhaseFused, __float_as_uint(lo), x);
            settled_total += ld_relaxed(&ctl->nset[hs]);
            const u32 nmed = ld_relaxed(&ctl->hlen[hs].v);
            const u64 hh = ld_relaxed(&ctl->hhub[hs].v);
            if (nmed || hub_count(hh)) {  // second phase of the same step: longer heavy rows
                process_tier<false>(a, a.hq1, min(nmed, (u32)a.n), 1, t, improved, slot, sq);
                process_hubs<false>(a, a.hqh, a.hq_off, a.hq_len, hh, t, improved, sq);
                stage_flush(a, sq);
                sync_traced(ctl, kPhaseHeavy, __float_as_uint(lo), x);
            }
            tick(1);
            if (gtid == 0) atomicAdd((unsigned long long *)&ctl->scans, 1ull);
            ++x;
        }
        {
            u32 cnt[2];
            u64 hub;
            // next bucket not empty: queued light work, or labels that entered it
            if (queue_size(x, cnt, &hub) || ld_relaxed(&ctl->nnext[x % 3])) {
                lo = hi;
                hi = hi2;
                bucket_items = 0;
                continue;
            }
        }
        if (ld_relaxed(&ctl->bar.error)) break;
        // min step: smallest distance >= hi among all labels (nothing changes now)
        log_step(2, 0, 0, 0);
        clear_ahead(x);
        {
            u32 bits = 0xFFFFFFFFu;
            scan_open(a, [&](u64 v, u32 lbits) {
                (void)v;
                if (lbits < kInfBits && __uint_as_float(lbits) >= hi) bits = min(bits, lbits);
            });
            bits = wave_min(bits);
            if (lane == 0 && bits != 0xFFFFFFFFu) atomicMin(&ctl->nmin[(x + 1) % 3], bits);
        }
        sync_traced(ctl, kPhaseMin, __float_as_uint(lo), x);
        ++x;
        const u32 nm = ld_relaxed(&ctl->nmin[x % 3]);
        if (gtid == 0) atomicAdd((unsigned long long *)&ctl->scans, 1ull);
        if (nm == 0xFFFFFFFFu) {
            tick(2);
            break;
        }
        const f32 dmin = __uint_as_float(nm);
        bucket_items = 0;
        lo = hi;
        hi = next_bound(lo, a.delta);
        while (dmin >= hi) {
            lo = hi;
            hi = next_bound(lo, a.delta);
        }
        // collect step: queue the bucket [lo, hi) that holds dmin
        log_step(3, 0, 0, 0);
        clear_ahead(x);
        {
            const Target t = target(x, hi, lo);
            stage_to(t);
            scan_open(a, [&](u64 v, u32 lbits) {
                bool want = false;
                if (lbits < kInfBits) {
                    const f32 d = __uint_as_float(lbits);
                    want = d >= lo && d < hi;
                }
                if (!ballot(want)) return;
                push_light(a, want, (u32)v, t, sq);
            });
        }
        stage_flush(a, sq);
        sync_traced(ctl, kPhaseCollect, __float_as_uint(lo), x);
        if (gtid == 0) atomicAdd((unsigned long long *)&ctl->scans, 1ull);
        tick(2);
        ++x;
    }
    improved = wave_sum(improved);
    if (lane == 0 && improved) atomicAdd((unsigned long long *)&ctl->improvements, (unsigned long long)improved);
    if (timer) {
        st_relaxed(&ctl->nsteps, x);
        for (int i = 0; i < 3; ++i) st_relaxed(&ctl->ticks[i], t_acc[i]);
        st_relaxed(&ctl->ticks[3], wall_clock64() - t_start);
    }
    // export in original ids (traversal.cu:333-340): unreached -> -1.0f / -1.
    // The probe is no longer read after the last barrier, so dist (its
    // storage) can take the output. Lanes walk consecutive original ids, so
    // the stores coalesce and the label reads are the scattered side.
    for (u64 o = gtid; o < a.n; o += gthreads) {
        const u64 l = a.lab[trav_id(a.perm, (u32)o)];
        const bool reached = (u32)(l >> 32) < kInfBits;
        if (reached && label_dist(l) >= a.dbound) atomicOr(&ctl->bar.error, 32u);
        a.dist[o] = reached ? label_dist(l) : -1.0f;
        a.pred[o] = reached ? (i64)orig_id(a.iperm, (u32)l) : -1;
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

__global__ __launch_bounds__(256) void sssp_rows_kernel(const u64 *__restrict__ rs, const u32 *__restrict__ split,
                                                        u64 n, RowRec *__restrict__ rows) {
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x)
        rows[v] = RowRec{rs[v], split[v], (u32)(rs[v + 1] - rs[v])};
}

__global__ __launch_bounds__(256) void sssp_clear_kernel(u64 *__restrict__ lab, u32 *__restrict__ dl,
                                                         u32 *__restrict__ stamp, u64 *__restrict__ settled,
                                                         u64 *__restrict__ reached, u64 n) {
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x) {
        lab[v] = kInfLabel;
        dl[v] = kInfBits;
        stamp[v] = 0;
        if ((v & 63) == 0) {
            settled[v >> 6] = 0;
            reached[v >> 6] = 0;
        }
    }
}

void compute_split(const u64 *rs, const f32 *w, u64 n, f32 delta, u32 *split, hipStream_t stream) {
    Context &c = ctx();
    sssp_split_kernel<<<grid_for(n, 256, (u64)c.num_cus * 32), 256, 0, stream ? stream : c.stream>>>(rs, w, n,
                                                                                                  delta, split);
    G500_HIP(hipGetLastError());
}

// Highest degree of the graph's rows packed with the smallest id reaching it:
// (degree << 32) | (~id), maximised.
__global__ __launch_bounds__(256) void hub_kernel(const u64 *__restrict__ rs, u64 n, u64 *__restrict__ best) {
    u64 m = 0;
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x)
        m = max(m, (MIN(rs[v + 1] - rs[v], (u64)0xFFFFFFFFu) << 32) | (0xFFFFFFFFu - (u32)v));
    for (int o = 32; o > 0; o >>= 1) {
        const u32 lo = __shfl_xor((u32)m, o), hi = __shfl_xor((u32)(m >> 32), o);
        m = max(m, ((u64)hi << 32) | lo);
    }
    if (lane_id() == 0 && m) atomicMax((unsigned long long *)best, (unsigned long long)m);
}

__global__ __launch_bounds__(256) void gather_dist_kernel(const f32 *__restrict__ dist, const u64 *__restrict__ ids,
                                                          u64 count, f32 *__restrict__ out) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (u64)gridDim.x * blockDim.x)
        out[i] = dist[ids[i]];
}

bool hub_distances(Worker_State *ws, const Oned_Graph *g, f32 *dist_dev) {
    Context &c = ctx();
    const u64 n = g->nlocalverts;
    const u64 mark = ws->gpu_arena.used;
    u64 *best = (u64 *)arena_alloc(ws, sizeof(u64), "hub search");
    if (failed()) return false;
    G500_HIP(hipMemsetAsync(best, 0, sizeof(u64), c.stream));
    hub_kernel<<<grid_for(n, 256, (u64)c.num_cus * 16), 256, 0, c.stream>>>(g->rowstarts, n, best);
    G500_HIP(hipGetLastError());
    u64 h = 0;
    g500_copy_to_host(&h, best, sizeof(h));
    const u64 hub = 0xFFFFFFFFu - (u32)h;
    Sssp_State s = oned_graph_sssp_create(ws, g);
    if (!failed() && hub < n) {
        s.delta = 0.05f;
        oned_graph_sssp_run(ws, g, &s, hub);
        G500_HIP(hipMemcpyAsync(dist_dev, s.dist, n * sizeof(f32), hipMemcpyDeviceToDevice, c.stream));
        sync();
    }
    ws->gpu_arena.used = mark;
    return !failed();
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
    s.heavy = (u32 *)arena_alloc(ws, 2 * n * sizeof(u32), "SSSP heavy lists");
    s.hub_offsets = (u64 *)arena_alloc(ws, 3 * n * sizeof(u64), "SSSP hub offsets");
    s.hub_lengths = (u32 *)arena_alloc(ws, 3 * n * sizeof(u32), "SSSP hub lengths");
    s.split = (u32 *)arena_alloc(ws, n * sizeof(u32), "SSSP light split");
    s.rows = arena_alloc(ws, n * sizeof(RowRec), "SSSP row records");
    s.stamp = (u32 *)arena_alloc(ws, n * sizeof(u32), "SSSP stamps");
    s.settled = (u64 *)arena_alloc(ws, (n + 63) / 64 * sizeof(u64), "SSSP settled bitmap");
    s.reached = (u64 *)arena_alloc(ws, (n + 63) / 64 * sizeof(u64), "SSSP reached bitmap");
    s.ctl = arena_alloc(ws, sizeof(SsspCtl), "SSSP control");
    if (failed()) return s;
    oned_graph_sssp_clear(ws, &s);
    return s;
}

void oned_graph_sssp_clear(Worker_State *ws, Sssp_State *s) {
    (void)ws;
    Context &c = ctx();
    const u64 n = s->max_nlocalverts;
    sssp_clear_kernel<<<grid_for(n, 256, (u64)c.num_cus * 32), 256, 0, c.stream>>>(s->dist_pred, (u32 *)s->dist,
                                                                                  s->stamp, s->settled, s->reached,
                                                                                  n);
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
    if (!(s->delta > 0.0f) || !(s->delta < 3.0e38f)) {
        fail("SSSP delta must be positive and finite (got %g)", (double)s->delta);
        return;
    }
    const u64 n = g->nlocalverts;
    const TraversalView tv = sssp_view(g);
    if (s->split_delta != s->delta || s->split_graph != (const void *)tv.rs) {  // light/heavy split of every row
        compute_split(tv.rs, tv.w, n, s->delta, s->split);
        sssp_rows_kernel<<<grid_for(n, 256, (u64)c.num_cus * 32), 256, 0, c.stream>>>(tv.rs, s->split, n,
                                                                                     (RowRec *)s->rows);
        G500_HIP(hipGetLastError());
        s->split_delta = s->delta;
        s->split_graph = (const void *)tv.rs;
    }
    SsspArgs a;
    a.n = n;
    a.rs = tv.rs;
    a.col = tv.col;
    a.w = tv.w;
    a.perm = tv.perm;
    a.iperm = tv.iperm;
    a.split = s->split;
    a.rows = (const RowRec *)s->rows;
    a.lab = s->dist_pred;
    a.dl = (u32 *)s->dist;
    for (int b = 0; b < 2; ++b) {
        for (int t = 0; t < 3; ++t) a.fq[b][t] = s->frontier + (u64)(b * 3 + t) * n;
        a.fq_off[b] = s->hub_offsets + (u64)b * n;
        a.fq_len[b] = s->hub_lengths + (u64)b * n;
    }
    a.hq1 = s->heavy;
    a.hqh = s->heavy + n;
    a.hq_off = s->hub_offsets + 2 * n;
    a.hq_len = s->hub_lengths + 2 * n;
    a.stamp = s->stamp;
    a.settled = s->settled;
    // with the degree-ordered copy, the isolated vertices are the last ids
    a.nact = (g->perm && g->notisolated) ? MIN(g->notisolated, n) : n;
    a.pull_ratio = (f32)c.sssp_pull;
    a.screen = (u32)ctx().sssp_screen;
    a.reached = s->reached;
    a.ctl = (SsspCtl *)s->ctl;
    a.pred = s->pred;
    a.dist = s->dist;
    a.root = root;
    a.delta = s->delta;
    a.dbound = g->rowstarts_s ? kPruneDistBound : __builtin_inff();
    const int block = ctx().sssp_block;
    int grid = cooperative_blocks((const void *)sssp_kernel, block, 0);
    if (failed()) return;
    void *args[] = {&a};
    G500_HIP(hipMemsetAsync(s->ctl, 0, sizeof(SsspCtl), c.stream));
    timer_start();
    launch_persistent((const void *)sssp_kernel, grid, block, args);
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
        fail("SSSP kernel error %u (1 barrier timeout, 2 hub list overflow, 4 bad queue entry, 8 queue overflow, "
             "16 bucket invariant, 32 distance at or above 2048 on the pruned SSSP copy: build with sssp_prune 0); %d blocks; barriers passed min %u (block %d phase %u tags %u/%u) max %u "
             "(block %d phase %u tags %u/%u)",
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

i32 g500_sssp_step_log(const Sssp_State *s, i32 *kind, u64 *items, u64 *hub_entries, f64 *ms, i32 max_steps) {
    static SsspCtl h;
    g500_copy_to_host(&h, s->ctl, sizeof(h));
    const f64 khz = ctx().wall_clock_khz;
    i32 k = 0;
    for (u32 x = 1; x < MIN(h.nsteps, kStepLog) && k < max_steps; ++x) {
        if (kind) kind[k] = (i32)h.step_kind[x];
        if (items) items[k] = (u64)h.step_n0[x] + h.step_n1[x] + (h.step_hub[x] >> kHubCountShift);
        if (hub_entries) hub_entries[k] = h.step_hub[x] & kHubTotalMask;
        if (ms) ms[k] = (f64)h.step_ticks[x] / khz;
        ++k;
    }
    return k;
}

void g500_sssp_phase_ms(const Sssp_State *s, f64 *ms) {
    static SsspCtl h;
    g500_copy_to_host(&h, s->ctl, sizeof(h));
    const f64 khz = ctx().wall_clock_khz;
    for (int i = 0; i < 4; ++i) ms[i] = (f64)h.ticks[i] / khz;
    ms[4] = (f64)h.scans;
}

i32 oned_graph_sssp_hub_distances(Worker_State *ws, const Oned_Graph *g, const u64 *roots, i32 nroots, f32 *out) {
    Context &c = ctx();
    if (nroots < 0 || (nroots && (!roots || !out))) {
        fail("oned_graph_sssp_hub_distances: bad root list (%d roots)", nroots);
        return -1;
    }
    for (i32 i = 0; i < nroots; ++i)
        if (roots[i] >= g->nglobalverts) {
            fail("SSSP root %llu is outside graph vertex range [0, %llu)", (unsigned long long)roots[i],
                 (unsigned long long)g->nglobalverts);
            return -1;
        }
    if (!nroots) return 0;
    const u64 n = g->nlocalverts;
    const u64 mark = ws->gpu_arena.used;
    f32 *dh = (f32 *)arena_alloc(ws, n * sizeof(f32), "hub distances");
    u64 *ids = (u64 *)arena_alloc(ws, (u64)nroots * sizeof(u64), "root ids");
    f32 *dev = (f32 *)arena_alloc(ws, (u64)nroots * sizeof(f32), "root hub distances");
    if (!failed() && hub_distances(ws, g, dh)) {
        g500_copy_to_device(ids, roots, (u64)nroots * sizeof(u64));
        gather_dist_kernel<<<grid_for((u64)nroots, 256, 64), 256, 0, c.stream>>>(dh, ids, (u64)nroots, dev);
        G500_HIP(hipGetLastError());
        g500_copy_to_host(out, dev, (u64)nroots * sizeof(f32));
    }
    ws->gpu_arena.used = mark;
    return failed() ? -1 : 0;
}

i32 oned_graph_sssp_root_order(Worker_State *ws, const Oned_Graph *g, const u64 *roots, i32 nroots, i32 *order) {
    if (nroots < 0 || (nroots && (!roots || !order))) {
        fail("oned_graph_sssp_root_order: bad root list (%d roots)", nroots);
        return -1;
    }
    std::vector<f32> d(nroots);
    if (oned_graph_sssp_hub_distances(ws, g, roots, nroots, d.data())) return -1;
    for (i32 i = 0; i < nroots; ++i) order[i] = i;
    std::stable_sort(order, order + nroots, [&](i32 x, i32 y) { return d[x] < d[y]; });
    return 0;
}

}  // extern "C"
