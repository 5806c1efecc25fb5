// Batched delta-stepping SSSP: up to 32 roots in one persistent cooperative
// kernel with one bucket sequence shared by all roots. No counterpart in the
// reference, whose runner solves one root at a time
// (examples/graph500_runner.cu:393-416); each root's distances equal the
// single-root kernel's, because every terminating label-correcting schedule
// reaches the same FP32 fixed point.
//
// Labels are vertex-major: lab[v * 32 + r] = (distance bits << 32) | parent of
// vertex v from root r, with a u32 probe dl[v * 32 + r] of the distance bits
// alone (the state's dist array while the kernel runs), so the 32 probes of a
// vertex fill one 128-byte line. A relaxation reads the probe, lowers it with
// a 32-bit atomic min and, only when that strictly lowered it, commits
// (distance, parent) to the label with a 64-bit atomic min; strict lowerings
// of a probe are by distinct values, so the label ends as (final distance, a
// parent that achieved it), as in the single-root kernel.
// Buckets are [lo, hi) with hi = fl(lo + delta) (traversal.cu:693-694), the
// same sequence for every root. Work items are (vertex, root mask) pairs: one
// pass over a row serves every root of the mask, and relaxing an edge for
// several roots touches one line of the target. Per bucket:
//   light steps: items (u, M) with M = pend[u], the roots whose label of u went
//     below hi in the previous step; their light edges (w < delta, a prefix of
//     the weight-sorted row) are relaxed for every root of M; a label lowered
//     below hi ORs its root into the target's pending mask for the next step,
//     and the first OR queues the target;
//   dense step: a lane per vertex with open work reads its probes; roots with
//     d in [lo, hi) become final and form the vertex's medium item, roots with
//     d in [hi, fl(hi + delta)) are queued for the next bucket; after a barrier
//     the medium items relax their medium edges (delta <= w < W_0, never into
//     [lo, hi), as for one root);
//   mid edges (W_0 <= w < W_1) and long edges (w >= W_1, the row's suffix)
//     are deferred, each class c with its own queue: the dense step ORs the
//     settled roots into the vertex's class-c mask and queues the vertex once.
//     A class-c relaxation from a root settled at d >= e_c (e_c: the floor of
//     the first bucket since the class's last flush) lands at
//     fl(d + w) >= fl(e_c + W_c), so a class is flushed (its edges relaxed for
//     every collected root) only before a bucket reaching past fl(e_c + W_c)
//     starts, and before a min step. A vertex settles a handful of roots per
//     window instead of one per bucket, and each flushed entry probes one line
//     for all of them;
//   when the next bucket is empty: a min step over all labels and a collect
//     step queueing the bucket that holds the minimum.
// Work is mapped by segment length: a lane per item (<= 32 entries), a wave
// per item (<= 1024) or 256-entry chunks over the grid. Every chunk of a hub
// item reads its mask; the last chunk to finish clears it.
// Counters rotate over three slots per step (written by the step's producer,
// read by its consumer, cleared two steps ahead), as in the single-root kernel.
#include <algorithm>
#include <cmath>

#include "sssp_common.hpp"

namespace g500 {

constexpr int kBatchRoots = kSsspBatchRoots;  // label stride and largest batch
constexpr u32 kBSmall = 32;
constexpr u32 kBHub = 1024;
constexpr u32 kBStepLog = 4096;
constexpr int kDefer = 3;  // deferred edge classes: 0 mid, 1 long, 2 far

// Error bits: 1 barrier timeout, 2 hub list overflow, 4 queue entry out of
// range, 8 queue length above n, 16 bucket invariant violated, 32 a distance
// at or above the pruned SSSP copy's bound (the copy may then be inexact).
struct SsspBatchCtl {
    GridBarrier bar;
    // Every counter that waves append to has a 256-byte line of its own, so
    // that appends to different lists do not queue behind each other.
    CounterLine<u32> flen[3][2];  // light queue lengths [slot][tier 0-1]
    CounterLine<u64> fhub[3];     // light queue hub counters [slot]
    CounterLine<u32> hlen[3][2];  // medium item list lengths [slot][tier 0-1]
    CounterLine<u64> hhub[3];     // medium hub counters [slot]
    CounterLine<u32> llen[kDefer][2][2];  // deferred queue lengths [class][set][tier 0-1], set = the class's flush parity
    CounterLine<u64> lhub[kDefer][2];     // deferred queue hub counters [class][set]
    alignas(256) u32 rmin[3][kBatchRoots];  // min step: smallest open distance bits per root, per slot
    CounterLine<u32> plen;        // pull flush: vertices handed from stage A to stage B
    CounterLine<u32> grab[2];     // light steps of long lists: next 64-item chunk [step parity]
    CounterLine<u32> mgrab;       // the same for medium items
    CounterLine<u32> fgrab[kDefer];  // and for each deferred class's flush
    alignas(256) u32 rw[kDefer][2];   // roots that some source settled in the class's window [class][set]
    u32 wcnt[2][kBatchRoots];     // far class: sources each root settled in the window [set][root]
    u64 pull_work[4];             // pull flushes, live (vertex, root) pairs, (vertex, entry) scan steps, stage-B vertices
    u32 nnext[3];                 // 1: a label entered [hi, hi + delta) (the bucket is not empty), per slot
    alignas(256) u64 buckets, light_iters, improvements, scans;
    u64 flushes;
    u64 ticks[6];                 // constant-clock ticks of block 0: light, dense scan, medium, flush,
                                  // min+collect, kernel
    u64 work[3 + kDefer][3];      // [light, medium][items, entries, (entry, root) pairs];
                                  // [dense scans][open words walked, vertices reading their line, settling];
                                  // [3 + deferred class][items, entries, pairs]
    // per-phase log (block 0): phase (0 light, 1 dense scan, 2 medium, 3 flush,
    // 4 min + collect), items processed, constant-clock ticks
    u32 nlog;
    u32 log_kind[kBStepLog], log_items[kBStepLog];
    u64 log_ticks[kBStepLog];
    alignas(256) BlockTrace trace[kMaxTraceBlocks];
};

// A row's segment bounds in one 32-byte record, so that an item's segment
// costs one memory transaction: row start, the light, mid, long and far
// splits (entries below delta, W_mid, W_long, W_far, from the start), row end.
struct RowInfo {
    u64 start;
    u32 light, cut[kDefer];
    u64 end;
};
static_assert(sizeof(RowInfo) == 32, "one 32-byte record per row");

__device__ __forceinline__ RowInfo load_row_info(const RowInfo *ri, u32 u) {
    const uint4 *p = reinterpret_cast<const uint4 *>(ri + u);
    const uint4 x = p[0], y = p[1];
    RowInfo r;
    r.start = ((u64)x.y << 32) | x.x;
    r.light = x.z;
    r.cut[0] = x.w;
    r.cut[1] = y.x;
    r.cut[2] = y.y;
    r.end = ((u64)y.w << 32) | y.z;
    return r;
}

struct SsspBatchArgs {
    u64 n;
    const u64 *rs;
    const u32 *col;
    const f32 *w;
    const u32 *split;
    const u32 *dsplit[kDefer];  // [n] entries with w < W_c: class c is [dsplit[c], dsplit[c+1]), the last to the row end
    const struct RowInfo *ri;   // [n] the row bounds and splits above in one 32-byte record
    const u32 *perm, *iperm;  // traversal <-> original ids (null: identity)
    u64 *lab;          // [n * kBatchRoots]
    u32 *dl;           // [n * kBatchRoots] probes (aliases dist)
    u32 *smask;        // [n] roots whose label of the vertex is final
    u32 *vmin;         // [n] smallest probe bits of the vertex's open roots
    u64 *openw;        // [ceil(n/64)] vertex has a reached root that is not final
    u32 *pend[2];      // [n] pending root masks, by step parity
    u32 *fq[2][3];     // light queues [buffer][tier]; tier 2 holds hub vertices
    u64 *fq_off[2];    // light hub tier: work offsets
    u32 *fq_len[2];    //   segment lengths
    u32 *fq_done[2];   //   finished chunks
    u64 *hq[2];        // heavy items (mask << 32 | v), tiers 0 and 1
    u32 *hqh;          // heavy hub tier: vertices,
    u64 *hq_off;       //   work offsets,
    u32 *hq_len;       //   segment lengths,
    u32 *hq_mask;      //   root masks
    u32 *dpend[kDefer];     // [n] deferred root masks per class
    u32 *dq[kDefer][3];     // deferred queue tiers per class; tier 2 holds hub vertices
    u64 *dq_off[kDefer];    //   hub work offsets,
    u32 *dq_len[kDefer];    //   segment lengths,
    u32 *dq_done[kDefer];   //   finished chunks
    SsspBatchCtl *ctl;
    f32 *dist;         // [nroots * n], root-major
    i64 *pred;         // [nroots * n]
    u64 roots[kBatchRoots];
    u32 lag[kBatchRoots];   // buckets root r lags the shared bucket counter (0 for at least one root)
    f32 uoff[kBatchRoots];  // fl(lag[r] * delta): the reduced-value offset of root r
    i32 nroots;
    f32 delta;
    f32 pull_ratio;    // the far class is pulled when its window items >= pull_ratio * nact (0: never)
    u32 pull_roots;    // roots a pull may serve: those with an offset (in the hub's component)
    u64 nact;          // vertices that can have entries: [0, nact) (the degree order puts isolated ones last)
    f32 dbound;        // distances at or above it set error 32 (+inf without the pruned copy)
    f32 dw[kDefer];    // class floor weights delta <= W_0 (mid) <= W_1 (long) <= W_2 (far)
};

__device__ __forceinline__ void bsync(SsspBatchCtl *ctl, u32 phase, u32 a, u32 b) {
    grid_sync(&ctl->bar, &ctl->bar.error);
    if (threadIdx.x == 0 && blockIdx.x < kMaxTraceBlocks) {
        BlockTrace *t = &ctl->trace[blockIdx.x];
        st_relaxed(&t->barriers, ld_relaxed(&t->barriers) + 1);
        st_relaxed(&t->phase, phase);
        st_relaxed(&t->a, a);
        st_relaxed(&t->b, b);
    }
}

// Per-root bucket state, the same in every block (each block keeps its copy
// in LDS and updates it identically). Root r runs the bucket sequence
// B(0) = 0, B(k + 1) = fl(B(k) + delta) (B(k) = 0 for k <= 0), lagging the
// shared bucket counter g by lag_r buckets: its bucket index is
// k_r = g - lag_r, and b[0..2][r] = B(k_r), B(k_r + 1), B(k_r + 2) are its
// lo, hi and next-bucket top. Each root's own schedule is plain
// delta-stepping, so every root ends at the same FP32 fixed point whatever its
// lag. A root whose distances run ahead of the others' by about lag_r * delta
// (the lags are set from the roots' distances to a hub) then settles each
// vertex in the same shared buckets as the other roots, so a row pass serves
// more roots at once.
// Every write to this state follows a block barrier that every read of the
// previous values precedes.
// Reduced values fl(d + u_r), with u_r = fl(lag_r * delta), put the roots'
// distances on the shared scale for the smallest-open-probe screen: rounding
// is monotone, so d < z_r implies fl(d + u_r) <= fl(z_r + u_r).
struct alignas(16) RootBuckets {
    f32 b[3][kBatchRoots];
    f32 u[kBatchRoots];
    f32 ep[3][kBatchRoots];  // deferred classes: lo_r of the first bucket since the class's last flush
    i32 k[kBatchRoots];
    u32 jr[kBatchRoots];     // min step: advances until the root's bucket holds its smallest open distance
    u32 smin[kBatchRoots];   // min step: this block's smallest open distance bits per root
};

// Where the pushes of a step go: pending-mask and queue buffer, counter slot,
// and the root bounds: a lowered label is queued below b[lvl + 1][r], and no
// candidate may fall below the floor b[lvl][r] (lvl 0: current bucket, 1: next).
struct BTarget {
    int buf, slot, lvl;
    const RootBuckets *rb;
};

__device__ __forceinline__ int btier(u32 len) { return len <= kBSmall ? 0 : (len <= kBHub ? 1 : 2); }

// Wave indices as wave-uniform (scalar) values: the compiler cannot tell on
// its own that threadIdx.x / 64 is the same for every lane of a wave.
__device__ __forceinline__ int wave_in_block() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave)); }
__device__ __forceinline__ u64 wave_index() {
    return (u64)blockIdx.x * (blockDim.x / kWave) + (u64)wave_in_block();
}

// Wave-buffered appends. A wave stages what it appends to the light queue
// tiers, the deferred queue tiers (u32 vertices) and the medium lists (u64
// items) in LDS and writes them out in runs: when its buffer for a list is
// full (one atomic for the run) and at the end of the phase, where the block
// sums its waves' counts and makes one atomic per list. Counters then see one
// atomic per block and phase instead of one per wave and relaxation round.
//
// The buffers are dynamic LDS sized by the block's wave count, so that blocks
// of fewer waves take proportionally less LDS and as many of them stay
// co-resident as the waves-per-CU limit allows.
constexpr u32 kQCap = 64;
enum : int { kQLight0 = 0, kQDef0 = 2, kQ32 = 2 + 2 * kDefer, kQMed0 = 0, kQ64 = 2 };  // u32: light 0-1, class c 2+2c..3+2c

struct QWave {  // one wave's staging buffers
    u32 v32[kQ32][kQCap];
    u64 v64[kQ64][kQCap];
};

// Dynamic LDS of a block of `block` threads: the waves' buffers, then the
// per-wave counts [waves][kQ32 + kQ64] (counts at the phase end, then the
// waves' offsets).
constexpr size_t q_lds_bytes(int block) {
    return (size_t)(block / kWave) * (sizeof(QWave) + (kQ32 + kQ64) * sizeof(u32));
}

struct WaveQ {
    QWave *W;              // this wave's buffers
    volatile u32 *cnt;     // the block's counts
    int w;                 // wave of the block
    u32 n[kQ32 + kQ64];    // staged entries per list (u32 lists first)
    int buf, slot;         // light queue buffer and counter slot of the phase's pushes
    int hs;                // medium list slot of a dense scan
    int ls[kDefer];        // deferred queue set per class of a dense scan
};

// Destination array and counter of list l (u32 lists 0..kQ32-1, u64 lists
// after). Selects over constant indexes, not runtime ones: a runtime index into
// the argument struct or a register array would move it to scratch.
__device__ __forceinline__ u32 *q_dst32(const SsspBatchArgs &a, const WaveQ &q, int l) {
    if (l < kQDef0) return q.buf ? (l ? a.fq[1][1] : a.fq[1][0]) : (l ? a.fq[0][1] : a.fq[0][0]);
    u32 *p = nullptr;
#pragma unroll
    for (int c = 0; c < kDefer; ++c) {
        if (l == kQDef0 + 2 * c) p = a.dq[c][0];
        if (l == kQDef0 + 2 * c + 1) p = a.dq[c][1];
    }
    return p;
}
__device__ __forceinline__ u32 *q_counter(const SsspBatchArgs &a, const WaveQ &q, int l) {
    SsspBatchCtl *ctl = a.ctl;
    if (l < kQDef0) return &ctl->flen[q.slot][l].v;
    if (l >= kQ32) return &ctl->hlen[q.hs][l - kQ32].v;
    u32 *p = nullptr;
#pragma unroll
    for (int c = 0; c < kDefer; ++c)
        if ((l - kQDef0) >> 1 == c) p = &ctl->llen[c][q.ls[c]][(l - kQDef0) & 1].v;
    return p;
}

// Writes the staged entries [0, cnt) of one list at dst[base...].
template <typename T>
__device__ __forceinline__ void q_copy(const volatile T *buf, u32 cnt, T *dst, u32 base, u64 cap) {
    for (u32 i = lane_id(); i < cnt; i += kWave) {
        const u64 o = (u64)base + i;
        if (o < cap) dst[o] = buf[i];
    }
}

template <int L>
__device__ __forceinline__ void q_push32(const SsspBatchArgs &a, WaveQ &q, bool want, u32 val) {
    const u64 m = ballot(want);
    if (!m) return;
    const u32 k = (u32)__popcll(m);
    volatile u32 *buf = q.W->v32[L];
    if (q.n[L] + k > kQCap) {  // spill the full buffer as one run
        u32 base = 0;
        if (lane_id() == 0) base = atomicAdd(q_counter(a, q, L), q.n[L]);
        base = __shfl(base, 0);
        q_copy(buf, q.n[L], q_dst32(a, q, L), base, a.n);
        q.n[L] = 0;
    }
    if (want) buf[q.n[L] + mask_rank(m)] = val;
    q.n[L] += k;
}

template <int L>
__device__ __forceinline__ void q_push64(const SsspBatchArgs &a, WaveQ &q, bool want, u64 val) {
    const u64 m = ballot(want);
    if (!m) return;
    const u32 k = (u32)__popcll(m);
    volatile u64 *buf = q.W->v64[L];
    if (q.n[kQ32 + L] + k > kQCap) {
        u32 base = 0;
        if (lane_id() == 0) base = atomicAdd(q_counter(a, q, kQ32 + L), q.n[kQ32 + L]);
        base = __shfl(base, 0);
        q_copy(buf, q.n[kQ32 + L], a.hq[L], base, a.n);
        q.n[kQ32 + L] = 0;
    }
    if (want) buf[q.n[kQ32 + L] + mask_rank(m)] = val;
    q.n[kQ32 + L] += k;
}

// End of a phase, called by every thread of the block: one atomic per list
// with staged entries, then each wave writes its runs.
__device__ __forceinline__ void q_flush(const SsspBatchArgs &a, WaveQ &q) {
    constexpr int kL = kQ32 + kQ64;
    volatile u32 *cnt = q.cnt;
    const int nw = (int)(blockDim.x / kWave);
    if (lane_id() == 0) {
#pragma unroll
        for (int l = 0; l < kL; ++l) cnt[q.w * kL + l] = q.n[l];
    }
    __syncthreads();
    if (threadIdx.x < (u32)kL) {
        const int l = (int)threadIdx.x;
        u32 tot = 0;
        for (int i = 0; i < nw; ++i) {
            const u32 c = cnt[i * kL + l];
            cnt[i * kL + l] = tot;
            tot += c;
        }
        u32 b = 0;
        if (tot) b = atomicAdd(q_counter(a, q, l), tot);
        for (int i = 0; i < nw; ++i) cnt[i * kL + l] += b;
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < kQ32; ++l) {
        if (q.n[l]) q_copy((volatile u32 *)q.W->v32[l], q.n[l], q_dst32(a, q, l), cnt[q.w * kL + l], a.n);
        q.n[l] = 0;
    }
#pragma unroll
    for (int l = 0; l < kQ64; ++l) {
        if (q.n[kQ32 + l])
            q_copy((volatile u64 *)q.W->v64[l], q.n[kQ32 + l], a.hq[l], cnt[q.w * kL + kQ32 + l], a.n);
        q.n[kQ32 + l] = 0;
    }
    __syncthreads();  // the LDS buffers and counts are reused by the next phase
}

// Queues v for the light step of target t with root mask `mask`; lw is v's
// light-entry count (split[v]), read by the caller together with its other
// loads of v (any value when mask is 0).
__device__ __forceinline__ void push_item_lw(const SsspBatchArgs &a, u32 v, u32 mask, u32 lw, const BTarget &t,
                                             WaveQ &wq) {
    const u64 wm = ballot(mask != 0);
    if (!wm) return;
    SsspBatchCtl *ctl = a.ctl;
    // the flag only tells that the next bucket is not empty: set it once
    if (lane_id() == (u32)(__ffsll((unsigned long long)wm) - 1) && !ld_relaxed(&ctl->nnext[t.slot]))
        st_relaxed(&ctl->nnext[t.slot], 1u);
    // a vertex without light entries has nothing to do in light steps
    bool q = false;
    if (!mask) lw = 0;
    if (lw) q = atomicOr(&(t.buf ? a.pend[1] : a.pend[0])[v], mask) == 0;
    if (!ballot(q)) return;
    q_push32<kQLight0>(a, wq, q && btier(lw) == 0, v);
    q_push32<kQLight0 + 1>(a, wq, q && btier(lw) == 1, v);
    hub_append(&ctl->fhub[t.slot].v, t.buf ? a.fq[1][2] : a.fq[0][2], t.buf ? a.fq_off[1] : a.fq_off[0],
               t.buf ? a.fq_len[1] : a.fq_len[0], q && btier(lw) == 2, v, lw,
               a.n, &ctl->bar.error);
}

__device__ __forceinline__ void push_item(const SsspBatchArgs &a, u32 v, u32 mask, const BTarget &t, WaveQ &wq) {
    push_item_lw(a, v, mask, mask ? a.split[v] : 0u, t, wq);
}

// A relaxation in two parts, so that a lane can have several relaxations in
// flight: relax_load reads the target and its probe for entry k (u at
// distance du, root r); relax_commit_push lowers the probes of a lane's
// rounds and queues the targets lowered below t.below.
struct Probe {
    u32 v, cb, cur;  // target, candidate distance bits, probe as read
};
// Relaxation rounds a lane has in flight: group tier, and the wave and hub
// tiers (one item's row, so a round costs fewer registers there).
constexpr int kRounds = 4;
constexpr int kWideRounds = 8;

// Branch-free, so that the loads of both relaxations of a lane are issued
// before the first one is waited for: an inactive lane reads entry 0 (valid
// whenever a segment is walked) and its result is ignored.
__device__ __forceinline__ void relax_load(const SsspBatchArgs &a, f32 du, int r, u64 k, bool act, Probe &p) {
    const u64 kk = act ? k : 0;
    p.v = a.col[kk];
    p.cb = __float_as_uint(du + a.w[kk]);
    p.cur = a.dl[(u64)p.v * kBatchRoots + r];  // a plain read may be stale, i.e. larger: it only lets an atomic through
}


// Commits R loaded relaxations of a lane and pushes the lowered targets, one
// phase at a time (probe atomics, label and mask updates, light-row counts,
// pending-mask ORs, queue appends), so that the R atomics of a phase are in
// flight together instead of one relaxation's chain after the other's.
template <int R>
__device__ __forceinline__ void relax_commit_push(const SsspBatchArgs &a, const u32 (&u)[R], const int (&r)[R],
                                                  const Probe (&pr)[R], const bool (&act)[R], const BTarget &t,
                                                  u64 &improved, WaveQ &wq) {
    bool tried[R];
    u32 prev[R], out[R], ow[R], lw[R];
    f32 below[R];
#pragma unroll
    for (int h = 0; h < R; ++h) tried[h] = act[h] && pr[h].cur > pr[h].cb;
#pragma unroll
    for (int h = 0; h < R; ++h) below[h] = t.rb->b[t.lvl + 1][r[h]];
#pragma unroll
    for (int h = 0; h < R; ++h)
        prev[h] = tried[h] ? atomicMin(&a.dl[(u64)pr[h].v * kBatchRoots + r[h]], pr[h].cb) : 0u;
    // Reads that depend on the target alone go out with the probe atomics, not
    // after them, so that the chain of dependent accesses is shorter. The
    // open-bit word (its 32-bit half holding v) is read when the plain probe
    // read showed v unreached from r: probes only fall, so only then can the
    // atomic report it unreached. The light split is read when the candidate
    // lands below t.below, the only case in which v may be queued.
#pragma unroll
    for (int h = 0; h < R; ++h)
        ow[h] = (tried[h] && pr[h].cur >= kInfBits) ? ((const u32 *)a.openw)[pr[h].v >> 5] : ~0u;
#pragma unroll
    for (int h = 0; h < R; ++h)
        lw[h] = (tried[h] && __uint_as_float(pr[h].cb) < below[h]) ? a.split[pr[h].v] : 0u;
    u32 nimp = 0;  // lowered probes of this lane; summed per wave below (a wave-uniform counter)
#pragma unroll
    for (int h = 0; h < R; ++h) {
        out[h] = 0;
        const u32 v = pr[h].v, cb = pr[h].cb;
        if (tried[h] && prev[h] > cb) {
            const f32 cand = __uint_as_float(cb);
            atomicMin((unsigned long long *)&a.lab[(u64)v * kBatchRoots + r[h]], ((unsigned long long)cb << 32) | u[h]);
            atomicMin(&a.vmin[v], __float_as_uint(cand + t.rb->u[r[h]]));
            if (prev[h] >= kInfBits) {  // v reached from root r: it has open work (again)
                // bits are cleared only by dense scans, never during a relaxation phase, so a
                // set bit read in this phase is current; a clear one may race a concurrent OR,
                // which is harmless
                const u32 bit = 1u << (v & 31);
                if (!(ow[h] & bit)) atomicOr(&((u32 *)a.openw)[v >> 5], bit);
            }
            ++nimp;
            if (cand < t.rb->b[t.lvl][r[h]]) atomicOr(&a.ctl->bar.error, 16u);
            out[h] = cand < below[h] ? 1u << r[h] : 0u;
        }
    }
    if (ballot(nimp != 0)) improved += wave_sum((u64)nimp);
    u64 any = 0;
#pragma unroll
    for (int h = 0; h < R; ++h) any |= ballot(out[h] != 0);
    if (!any) return;
    SsspBatchCtl *ctl = a.ctl;
    // the flag only tells that the next bucket is not empty: set it once
    if (lane_id() == (u32)(__ffsll((unsigned long long)any) - 1) && !ld_relaxed(&ctl->nnext[t.slot]))
        st_relaxed(&ctl->nnext[t.slot], 1u);
    // a vertex without light entries has nothing to do in light steps
    u32 *pend = t.buf ? a.pend[1] : a.pend[0];
    u32 po[R];
#pragma unroll
    for (int h = 0; h < R; ++h) lw[h] = out[h] ? lw[h] : 0u;
#pragma unroll
    for (int h = 0; h < R; ++h) po[h] = (out[h] && lw[h]) ? atomicOr(&pend[pr[h].v], out[h]) : 1u;
#pragma unroll
    for (int h = 0; h < R; ++h) {
        const bool q = out[h] && lw[h] && po[h] == 0;
        if (!ballot(q)) continue;
        const u32 v = pr[h].v;
        q_push32<kQLight0>(a, wq, q && btier(lw[h]) == 0, v);
        q_push32<kQLight0 + 1>(a, wq, q && btier(lw[h]) == 1, v);
        hub_append(&ctl->fhub[t.slot].v, t.buf ? a.fq[1][2] : a.fq[0][2], t.buf ? a.fq_off[1] : a.fq_off[0],
                   t.buf ? a.fq_len[1] : a.fq_len[0], q && btier(lw[h]) == 2, v, lw[h], a.n, &ctl->bar.error);
    }
}

// Work counts (items, entries, pairs) of one segment kind, kept per block in
// LDS and summed into the control block at the end.
__device__ __forceinline__ void work_add(u64 *wk, u64 items, u64 entries, u64 pairs) {
    atomicAdd((unsigned long long *)&wk[0], (unsigned long long)items);
    atomicAdd((unsigned long long *)&wk[1], (unsigned long long)entries);
    atomicAdd((unsigned long long *)&wk[2], (unsigned long long)pairs);
}

// Index of the i-th set bit of M (M has more than i bits).
__device__ __forceinline__ int nth_root(u32 M, u32 i) {
    for (; i; --i) M &= M - 1;
    return __ffs(M) - 1;
}

// Row segments (rows are sorted by weight): light w < delta, medium
// delta <= w < W_0, and deferred class c W_c <= w < W_{c+1} (the last class
// to the row end): segment kSegDef + c. Light and deferred items take their
// root mask from a pending array (cleared when read); medium items carry it.
enum : int { kSegLight = 0, kSegMedium = 1, kSegDef = 2 };

template <int K>
__device__ __forceinline__ void bsegment(const SsspBatchArgs &a, u32 u, u64 &b, u64 &e) {
    const RowInfo r = load_row_info(a.ri, u);
    if (K == kSegLight) {
        b = r.start;
        e = r.start + r.light;
    } else if (K == kSegMedium) {
        b = r.start + r.light;
        e = r.start + r.cut[0];
    } else {
        constexpr int c = K - kSegDef;
        b = r.start + r.cut[c];
        e = c + 1 < kDefer ? r.start + r.cut[c + 1 < kDefer ? c + 1 : c] : r.end;
    }
}

// The work of an item (u, M) over a segment is its (entry, root) pairs,
// entry-major: pair q = entry * k + i, i-th root of M, k = |M|. The k lanes of
// one entry read the same column and weight and relax k labels in the same
// 128-byte line of the target.

// Pull form of the far class's flush, for windows whose items are a large
// share of the graph. Every vertex v with a root r that is not final at v,
// that some source of the window settled (bit r of rw) and whose probe is
// above fl(e_r + W_c) (the smallest candidate a class entry from the window
// can give; e_r is the class epoch of root r) scans the class segment of its
// own row, lightest first. The SSSP rows are symmetric, so entry (v, u, w)
// stands for the pushed relaxation of (u, v, w): when u's window mask
// dpend[u] holds r, the candidate is fl(d_r(u) + w) with d_r(u) final. Root r
// stops at the first entry with fl(e_r + w) >= its best candidate so far:
// every later candidate comes from a source settled at >= e_r through an
// entry at least as heavy, and rounding is monotone. A half wave takes one
// vertex, one lane per root, so an entry's source line is read once for all
// 32 roots. A vertex whose roots still scan after kPullLane entries goes to a
// list that whole waves finish (two entries per step).
constexpr u32 kPullLane = 32;

// Lane r of a half wave holding v lowers root r's probe of v to `best`
// (parent `par`) and queues v for the light steps when it lands below
// b[lvl + 1][r]; `mine`: this lane commits.
__device__ __forceinline__ void pull_commit(const SsspBatchArgs &a, const RootBuckets &rb, u32 v, bool mine, int r,
                                            u32 p, u32 best, u32 par, const BTarget &t, u64 &improved, WaveQ &wq) {
    bool lowered = false;
    if (mine && best < p) {
        const u32 prev = atomicMin(&a.dl[(u64)v * kBatchRoots + r], best);
        if (prev > best) {
            lowered = true;
            const f32 d = __uint_as_float(best);
            atomicMin((unsigned long long *)&a.lab[(u64)v * kBatchRoots + r], ((unsigned long long)best << 32) | par);
            atomicMin(&a.vmin[v], __float_as_uint(d + rb.u[r]));
            if (prev >= kInfBits) atomicOr(&((u32 *)a.openw)[v >> 5], 1u << (v & 31));
            if (d < rb.b[t.lvl][r]) atomicOr(&a.ctl->bar.error, 16u);
        }
    }
    const u64 lm = ballot(lowered);
    if (!lm) return;
    improved += (u64)__popcll(lm);
    const u64 om = ballot(lowered && __uint_as_float(best) < rb.b[t.lvl + 1][r]);
    const u32 half = (u32)(om >> (lane_id() & 32u));
    push_item(a, v, (lane_id() & 31u) == 0 ? half : 0u, t, wq);
}

// Stage A: a half wave per vertex of [0, nact), at most kPullLane entries.
template <int C>
__device__ __forceinline__ void pull_stage_a(const SsspBatchArgs &a, const RootBuckets &rb, u32 rw, const BTarget &t,
                                             u64 &improved, WaveQ &wq, u64 *pw) {
    constexpr int K = kSegDef + C;
    const u32 lane = lane_id();
    const int r = (int)(lane & 31u), hl = (int)(lane >> 5);
    const u64 wave_gid = wave_index(), nwaves = (u64)gridDim.x * blockDim.x / kWave;
    const f32 er = rb.ep[C][r];
    const u32 lowb = __float_as_uint(er + a.dw[C]);
    const bool inw = (rw >> r) & 1u;
    SsspBatchCtl *ctl = a.ctl;
    for (u64 v0 = wave_gid * 2; v0 < a.nact; v0 += nwaves * 2) {
        const u64 v = v0 + (u64)hl;
        const bool valid = v < a.nact;
        const u32 fin = valid ? a.smask[v] : ~0u;
        const u32 p = valid ? a.dl[v * kBatchRoots + r] : 0u;
        bool act = valid && inw && !((fin >> r) & 1u) && p > lowb;
        const u64 lm = ballot(act);
        if (!lm) continue;
        u64 b = 0, e = 0;
        if ((u32)(lm >> (lane & 32u))) bsegment<K>(a, (u32)v, b, e);
        const u64 cap = min(e, b + kPullLane);
        u32 best = p, par = 0;
        bool cont = false;  // still scanning at the cap, with entries left
        u64 steps = 0;
        for (u64 k = b;; ++k) {
            if (act && k >= cap) {
                cont = k < e;
                act = false;
            }
            const f32 wk = act ? a.w[k] : 0.0f;
            act = act && __float_as_uint(er + wk) < best;
            const u64 am = ballot(act);
            if (!am) break;
            steps += (u64)((u32)am != 0) + (u64)((u32)(am >> 32) != 0);
            if (act) {
                const u32 u = a.col[k];
                if ((a.dpend[C][u] >> r) & 1u) {
                    const u32 cb = __float_as_uint(__uint_as_float(a.dl[(u64)u * kBatchRoots + r]) + wk);
                    if (cb < best) {
                        best = cb;
                        par = u;
                    }
                }
            }
        }
        const u64 cm = ballot(cont);
        if ((lane & 31u) == 0 && (u32)(cm >> (lane & 32u))) {
            const u32 at = atomicAdd(&ctl->plen.v, 1u);
            if (at < a.n) a.hqh[at] = (u32)v;
            else atomicOr(&ctl->bar.error, 2u);
        }
        pull_commit(a, rb, (u32)v, valid, r, p, best, par, t, improved, wq);
        if (lane == 0) {
            atomicAdd((unsigned long long *)&pw[1], (unsigned long long)__popcll(lm));
            atomicAdd((unsigned long long *)&pw[2], (unsigned long long)steps);
            atomicAdd((unsigned long long *)&pw[3], (unsigned long long)(((u32)cm != 0) + ((u32)(cm >> 32) != 0)));
        }
    }
}

// Stage B: a whole wave per listed vertex, from entry kPullLane on, the two
// halves on alternate entries; the halves' best candidates are merged at the end.
template <int C>
__device__ __forceinline__ void pull_stage_b(const SsspBatchArgs &a, const RootBuckets &rb, u32 rw, u32 count,
                                             const BTarget &t, u64 &improved, WaveQ &wq, u64 *pw) {
    constexpr int K = kSegDef + C;
    const u32 lane = lane_id();
    const int r = (int)(lane & 31u), hl = (int)(lane >> 5);
    const u64 wave_gid = wave_index(), nwaves = (u64)gridDim.x * blockDim.x / kWave;
    const f32 er = rb.ep[C][r];
    const u32 lowb = __float_as_uint(er + a.dw[C]);
    const bool inw = (rw >> r) & 1u;
    for (u64 i = wave_gid; i < count; i += nwaves) {
        const u32 v = a.hqh[i];
        if (v >= a.nact) {
            if (lane == 0) atomicOr(&a.ctl->bar.error, 4u);
            continue;
        }
        const u32 fin = a.smask[v];
        const u32 p = a.dl[(u64)v * kBatchRoots + r];
        bool act = inw && !((fin >> r) & 1u) && p > lowb;
        if (!ballot(act)) continue;
        u64 b, e;
        bsegment<K>(a, v, b, e);
        u32 best = p, par = 0;
        u64 steps = 0;
        for (u64 k0 = b + kPullLane;; k0 += 2) {
            const u64 k = k0 + (u64)hl;
            if (act && k >= e) act = false;
            const f32 wk = act ? a.w[k] : 0.0f;
            act = act && __float_as_uint(er + wk) < best;
            const u64 am = ballot(act);
            if (!am) break;
            steps += (u64)((u32)am != 0) + (u64)((u32)(am >> 32) != 0);
            if (act) {
                const u32 u = a.col[k];
                if ((a.dpend[C][u] >> r) & 1u) {
                    const u32 cb = __float_as_uint(__uint_as_float(a.dl[(u64)u * kBatchRoots + r]) + wk);
                    if (cb < best) {
                        best = cb;
                        par = u;
                    }
                }
            }
        }
        const u32 ob = __shfl_xor(best, 32), op = __shfl_xor(par, 32);
        if (ob < best) {
            best = ob;
            par = op;
        }
        pull_commit(a, rb, v, hl == 0, r, p, best, par, t, improved, wq);
        if (lane == 0) atomicAdd((unsigned long long *)&pw[2], (unsigned long long)steps);
    }
}

// After a pull, clears the window masks of the class's queued sources.
template <int C>
__device__ __forceinline__ void pull_clear(const SsspBatchArgs &a, u32 n0, u32 n1, u64 nh) {
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    for (u64 i = gtid; i < n0 + n1 + nh; i += gthreads) {
        const u32 u = i < n0 ? a.dq[C][0][i] : (i < (u64)n0 + n1 ? a.dq[C][1][i - n0] : a.dq[C][2][i - n0 - n1]);
        if (u < a.n) a.dpend[C][u] = 0;
    }
}

// Tier 0 (segments <= 32 entries): a wave takes 64 items and walks their
// concatenated pair ranges 64 pairs at a time. Light and long items read their
// mask from pend and clear it; medium items carry it in the list entry.
template <int K>
__device__ __forceinline__ void group_tier(const SsspBatchArgs &a, u32 *pend, const u32 *lq, const u64 *hq, u32 count,
                           const BTarget &t, u64 &improved, volatile int *slot, u64 *wk, WaveQ &wq,
                           u32 keep = ~0u, u32 *grab = nullptr) {
    const u64 gthreads = (u64)gridDim.x * blockDim.x;
    const u32 lane = lane_id();
    const u64 wave_gid = wave_index(), nwaves = gthreads / kWave;
    // Items per wave: 64, or fewer when 64-item chunks would leave waves idle
    // (a short list then spreads over the whole grid). A list of several
    // chunks per wave is dealt dynamically when the caller gives a counter
    // (zero at the start, grab): the cost of a chunk varies with its rows, and
    // a fixed stride leaves the waves with the heaviest chunks running alone.
    const u64 per = count >= nwaves * kWave ? (u64)kWave : max((u64)1, ((u64)count + nwaves - 1) / nwaves);
    const bool dyn = grab != nullptr && (u64)count >= 4 * nwaves * kWave;
    auto next_chunk = [&]() -> u64 {
        u32 c = 0;
        if (lane == 0) c = atomicAdd(grab, 1u);
        return (u64)__builtin_amdgcn_readfirstlane(c) * kWave;
    };
    for (u64 i0 = dyn ? next_chunk() : wave_gid * per; i0 < count; i0 = dyn ? next_chunk() : i0 + nwaves * per) {
        const u64 i = i0 + lane;
        u32 u = 0, M = 0;
        u64 b = 0, e = 0;
        if (lane < per && i < count) {
            if (K != kSegMedium) {
                u = lq[i];
                if (u < a.n) {
                    M = pend[u] & keep;
                    if (keep == ~0u) pend[u] = 0;
                    // the row record does not depend on the mask: its load goes out with the mask's
                    bsegment<K>(a, u, b, e);
                    if (!M) b = e = 0;
                } else {
                    atomicOr(&a.ctl->bar.error, 4u);
                }
            } else {
                const u64 it = hq[i];
                u = (u32)it;
                M = (u32)(it >> 32);
                if (u >= a.n) {
                    atomicOr(&a.ctl->bar.error, 4u);
                    M = 0;
                }
                if (M) bsegment<K>(a, u, b, e);
            }
        }
        const u32 pairs = (u32)(e - b) * (u32)__popc(M);
        const u32 incl = wave_inclusive_scan(pairs);
        const u32 off = incl - pairs, total = __shfl(incl, kWave - 1);
        {
            const u64 nitems = (u64)__popcll(ballot(M != 0)), nent = wave_sum((u64)(e - b));
            if (lane == 0) work_add(wk, nitems, nent, total);
        }
        for (u32 c = 0; c < total; c += kRounds * kWave) {
            u32 uu[kRounds];
            int rr[kRounds];
            bool act[kRounds];
            Probe pr[kRounds];
#pragma unroll
            for (int h = 0; h < kRounds; ++h) {
                const u32 ch = c + (u32)h * kWave;
                uu[h] = 0;
                rr[h] = 0;
                act[h] = false;
                u64 k = 0;
                if (ch < total) {
                    bool inter;
                    int last;
                    int j = chunk_owner(slot, off, pairs, ch, inter, last);
                    if (j < 0) j = 0;
                    const u32 uj = __shfl(u, j), Mj = __shfl(M, j), offj = __shfl(off, j);
                    const u64 bj = shfl64(b, j);
                    const u32 pos = ch + lane;
                    if (pos < total) {
                        const u32 kk = (u32)__popc(Mj), q = pos - offj;
                        const u32 entry = q / kk;
                        uu[h] = uj;
                        rr[h] = nth_root(Mj, q - entry * kk);
                        k = bj + entry;
                        act[h] = true;
                    }
                }
                const f32 du = __uint_as_float(a.dl[(u64)uu[h] * kBatchRoots + rr[h]]);
                relax_load(a, du, rr[h], k, act[h], pr[h]);
            }
            relax_commit_push<kRounds>(a, uu, rr, pr, act, t, improved, wq);
        }
    }
}

// Lane roles for a wave-uniform mask M: `per` entries per step, lane -> entry
// offset and root.
struct PairLanes {
    u32 per, eoff;
    int r;
    bool active;
};
__device__ __forceinline__ PairLanes pair_lanes(u32 M) {
    const u32 lane = lane_id(), k = (u32)__popc(M);
    PairLanes p;
    p.per = kWave / k;
    p.eoff = lane / k;
    p.active = lane < p.per * k;
    p.r = p.active ? nth_root(M, lane - p.eoff * k) : 0;
    return p;
}

// Tier 1 (segments <= 1024 entries): a wave per item.
template <int K>
__device__ __forceinline__ void wave_tier(const SsspBatchArgs &a, u32 *pend, const u32 *lq, const u64 *hq, u32 count,
                          const BTarget &t, u64 &improved, u64 *wk, WaveQ &wq, u32 keep = ~0u) {
    const u64 gthreads = (u64)gridDim.x * blockDim.x;
    const u32 lane = lane_id();
    const u64 wave_gid = wave_index(), nwaves = gthreads / kWave;
    for (u64 i = wave_gid; i < count; i += nwaves) {
        u32 u, M;
        if (K != kSegMedium) {
            u = lq[i];
            if (u >= a.n) {
                if (lane == 0) atomicOr(&a.ctl->bar.error, 4u);
                continue;
            }
            M = __builtin_amdgcn_readfirstlane(pend[u]) & keep;
            if (lane == 0 && keep == ~0u) pend[u] = 0;
        } else {
            const u64 it = hq[i];
            u = (u32)it;
            M = (u32)(it >> 32);
            if (u >= a.n) {
                if (lane == 0) atomicOr(&a.ctl->bar.error, 4u);
                continue;
            }
        }
        u64 b, e;
        bsegment<K>(a, u, b, e);  // before the mask test, so that its load goes out with the mask's
        if (!M) continue;
        if (lane == 0) work_add(wk, 1, e - b, (e - b) * (u64)__popc(M));
        const PairLanes p = pair_lanes(M);
        const f32 du = p.active ? __uint_as_float(a.dl[(u64)u * kBatchRoots + p.r]) : 0.f;
        u32 uk[kWideRounds];
        int rk[kWideRounds];
#pragma unroll
        for (int h = 0; h < kWideRounds; ++h) {
            uk[h] = u;
            rk[h] = p.r;
        }
        for (u64 e0 = b; e0 < e; e0 += kWideRounds * p.per) {
            bool act[kWideRounds];
            Probe pr[kWideRounds];
#pragma unroll
            for (int h = 0; h < kWideRounds; ++h) {
                const u64 k = e0 + (u64)h * p.per + p.eoff;
                act[h] = p.active && k < e;
                relax_load(a, du, p.r, k, act[h], pr[h]);
            }
            relax_commit_push<kWideRounds>(a, uk, rk, pr, act, t, improved, wq);
        }
    }
}

// Hub tier: chunks of the flattened work space. Light and long hubs read their
// mask from pend (cleared by the item's last chunk; `aux` counts finished
// chunks); medium hubs from `aux`.
template <int K>
__device__ __forceinline__ void hub_tier(const SsspBatchArgs &a, const u32 *hv, const u64 *hoff, const u32 *hlen, u32 *aux,
                         u32 *pend, u64 packed, const BTarget &t, u64 &improved, u64 *wk, WaveQ &wq,
                         u32 keep = ~0u) {
    const u64 gthreads = (u64)gridDim.x * blockDim.x;
    const u32 lane = lane_id();
    const u64 wave_gid = wave_index(), nwaves = gthreads / kWave;
    const u64 count = hub_count(packed), total = hub_total(packed);
    u64 i = 0;
    for (u64 c = wave_gid; c * kHubChunk < total; c += nwaves) {
        const u64 p0 = c * kHubChunk;
        i = hub_locate(hoff, i, count, p0);
        const u32 u = hv[i];
        if (u >= a.n) {
            if (lane == 0) atomicOr(&a.ctl->bar.error, 4u);
            continue;
        }
        const u32 M = __builtin_amdgcn_readfirstlane(K != kSegMedium ? pend[u] & keep : aux[i]);
        const u64 local = p0 - hoff[i];
        const u32 len = hlen[i];
        u64 b, e;
        bsegment<K>(a, u, b, e);
        if (M) {
            const u64 inchunk = min((u64)kHubChunk, (u64)len - min((u64)len, local));
            if (lane == 0) work_add(wk, local == 0 ? 1 : 0, inchunk, inchunk * (u64)__popc(M));
            const PairLanes p = pair_lanes(M);
            const f32 du = p.active ? __uint_as_float(a.dl[(u64)u * kBatchRoots + p.r]) : 0.f;
            u32 uk[kWideRounds];
            int rk[kWideRounds];
#pragma unroll
            for (int h = 0; h < kWideRounds; ++h) {
                uk[h] = u;
                rk[h] = p.r;
            }
            for (u32 j = 0; j < (u32)kHubChunk; j += kWideRounds * p.per) {
                bool act[kWideRounds];
                Probe pr[kWideRounds];
#pragma unroll
                for (int h = 0; h < kWideRounds; ++h) {
                    const u32 jj = j + (u32)h * p.per + p.eoff;
                    const u64 off = local + jj;
                    act[h] = p.active && jj < (u32)kHubChunk && off < len;
                    relax_load(a, du, p.r, b + off, act[h], pr[h]);
                }
                relax_commit_push<kWideRounds>(a, uk, rk, pr, act, t, improved, wq);
            }
        }
        if (K != kSegMedium && lane == 0) {
            const u32 chunks = (len + (u32)kHubChunk - 1) / (u32)kHubChunk;
            if (atomicAdd(&aux[i], 1u) + 1 == chunks) {  // every chunk has read the mask
                if (keep == ~0u) pend[u] = 0;
                aux[i] = 0;
            }
        }
    }
}

__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void sssp_batch_kernel(SsspBatchArgs a) {
    SsspBatchCtl *ctl = a.ctl;
    const u64 gtid = (u64)blockIdx.x * blockDim.x + threadIdx.x, gthreads = (u64)gridDim.x * blockDim.x;
    const u32 lane = lane_id();
    __shared__ int slots[512 / kWave][kWave];
    volatile int *slot = slots[wave_in_block()];
    const u64 nwords = (a.n + 63) / 64;
    const u64 wave_gid = wave_index(), nwaves = gthreads / kWave;
    u64 improved = 0;
    // work counts of this block: [light, medium, dense scans, deferred class c at 3 + c][3]
    __shared__ u64 wks[3 + kDefer][3];
    if (threadIdx.x < 3 * (3 + kDefer)) wks[threadIdx.x / 3][threadIdx.x % 3] = 0;
    extern __shared__ u64 q_lds[];
    WaveQ wq;
    wq.w = wave_in_block();
    wq.W = (QWave *)q_lds + wq.w;
    wq.cnt = (volatile u32 *)((QWave *)q_lds + blockDim.x / kWave);
#pragma unroll
    for (int l = 0; l < kQ32 + kQ64; ++l) wq.n[l] = 0;
    wq.buf = wq.slot = wq.hs = 0;
#pragma unroll
    for (int c = 0; c < kDefer; ++c) wq.ls[c] = 0;
    // the light queue tiers that a phase's pushes go to
    auto light_dest = [&](const BTarget &t) {
        wq.buf = t.buf;
        wq.slot = t.slot;
    };
    // per-root bucket state: thread r < 32 of every block keeps root r's entries
    __shared__ RootBuckets rbk;
    __shared__ u32 rws[kDefer];       // dense step: roots this block deferred into each class
    __shared__ u32 wcs[kBatchRoots];  // dense step: far-class sources this block deferred, per root
    __shared__ u64 pws[4];            // pull flush counts of this block (see g500_sssp_batch_pull_work)
    if (threadIdx.x < 4) pws[threadIdx.x] = 0;
    const int rt = (int)threadIdx.x;  // the root this thread keeps (when < kBatchRoots)
    if (rt < kBatchRoots) {
        // selects, not a runtime index into the argument struct (which would copy it to scratch)
        u32 lag = 0;
        f32 uo = 0.0f;
#pragma unroll
        for (int i = 0; i < kBatchRoots; ++i)
            if (rt == i) {
                lag = a.lag[i];
                uo = a.uoff[i];
            }
        const f32 b1 = next_bound(0.0f, a.delta), b2 = next_bound(b1, a.delta);
        const i32 k = -(i32)lag;  // <= 0
        rbk.k[rt] = k;
        rbk.b[0][rt] = 0.0f;             // B(k), k <= 0
        rbk.b[1][rt] = k + 1 >= 1 ? b1 : 0.0f;
        rbk.b[2][rt] = k + 2 >= 2 ? b2 : (k + 2 == 1 ? b1 : 0.0f);
        rbk.u[rt] = uo;
#pragma unroll
        for (int c = 0; c < kDefer; ++c) rbk.ep[c][rt] = 0.0f;
    }
    // one bucket further for root rt: B(k + 3) = fl(B(k + 2) + delta) once k + 2 >= 0
    auto advance_root = [&]() {
        const i32 k = rbk.k[rt];
        const f32 b2 = rbk.b[2][rt];
        rbk.b[0][rt] = rbk.b[1][rt];
        rbk.b[1][rt] = b2;
        rbk.b[2][rt] = k + 2 >= 0 ? next_bound(b2, a.delta) : 0.0f;
        rbk.k[rt] = k + 1;
    };
    // the screen threshold of a scan: max_r fl(b[L][r] + u_r) over the roots
    // with b[L][r] > 0 (block-uniform); a root with b[L][r] = 0 has not started,
    // and its ranges [lo_r, hi_r) and [hi_r, b[L][r]) of the scan are empty
    auto screen_top = [&](int L) {
        f32 z = 0.0f;
        for (int r = 0; r < kBatchRoots; ++r)
            if (rbk.b[L][r] > 0.0f) z = fmaxf(z, rbk.b[L][r] + rbk.u[r]);
        return z;
    };
    __syncthreads();
    u32 x = 0;
    u32 gb = 0;  // shared bucket counter
    // per deferred class: flushes so far (selects the counter set) and whether
    // the queue holds work (indexed with constants only, so they stay in registers)
    u32 nflush[kDefer];
    bool lpending[kDefer];
#pragma unroll
    for (int c = 0; c < kDefer; ++c) {
        nflush[c] = 0;
        lpending[c] = false;
    }

    const bool timer = gtid == 0;
    // timer state of thread 0 lives in LDS: [0] previous tick, [1] start,
    // [2..6] per-phase sums, [7] log entries (registers held by every lane
    // for one thread's bookkeeping would cost the relaxation loops)
    __shared__ u64 tst[8];
    if (timer) {
        tst[0] = tst[1] = wall_clock64();
        for (int i = 2; i < 8; ++i) tst[i] = 0;
    }
    auto tick = [&](int phase, u64 items = 0) {
        if (timer) {
            const u64 t = wall_clock64(), d = t - tst[0];
            const u32 nlog = (u32)tst[7];
            tst[2 + phase] += d;
            if (nlog < kBStepLog) {
                st_relaxed(&ctl->log_kind[nlog], (u32)phase);
                st_relaxed(&ctl->log_items[nlog], (u32)min(items, (u64)0xFFFFFFFFu));
                st_relaxed(&ctl->log_ticks[nlog], d);
            }
            tst[7] = nlog + 1;
            tst[0] = t;
        }
    };
    auto clear_ahead = [&](u32 step) {
        if (gtid == 0) {
            const int s = (int)((step + 2) % 3);
            st_relaxed(&ctl->flen[s][0].v, 0u);
            st_relaxed(&ctl->flen[s][1].v, 0u);
            st_relaxed(&ctl->fhub[s].v, 0ull);
            st_relaxed(&ctl->hlen[s][0].v, 0u);
            st_relaxed(&ctl->hlen[s][1].v, 0u);
            st_relaxed(&ctl->hhub[s].v, 0ull);
            for (int r = 0; r < kBatchRoots; ++r) st_relaxed(&ctl->rmin[s][r], 0xFFFFFFFFu);
            st_relaxed(&ctl->nnext[s], 0u);
        }
    };
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
    // pushes of step `step`; lvl 0: queued below hi_r (floor lo_r), 1: below hi2_r (floor hi_r)
    auto target = [&](u32 step, int lvl) {
        return BTarget{(int)((step + 1) & 1), (int)((step + 1) % 3), lvl, &rbk};
    };
    // Word-parallel pass over the vertices with open work: each lane reads one
    // word of openw, the wave walks the words that have open vertices, one
    // lane per vertex. vmin[v], the smallest reduced value fl(d + u_r) of v's
    // open roots (exact between phases), screens the vertex: only one with
    // vmin <= max_r fl(z_r + u_r) can hold a root with d < z_r, and only it
    // reads its probe line, in the 16-byte pieces that hold a root not yet
    // final. Per root r, with x_r = lo_r, y_r = hi_r and z_r = b[zl][r] (zl 1:
    // hi_r, 2: hi2_r), f(word, v, A, B, dmin, act, R) gets the roots whose open
    // probe is in [x_r, y_r) (A) and in [y_r, z_r) (B), the smallest reduced
    // value of the open probes >= y_r (vmin itself for a screened-out open
    // vertex, all ones for a closed one), whether the line was read, and the
    // reached roots that are not final (R, read lines only).
    auto scan_open = [&](int zl, auto &&f) {
        const u32 zb = __float_as_uint(screen_top(zl));
        // Words are dealt to waves round-robin (lane l of wave g reads word
        // base + l * nwaves + g): open vertices cluster in the degree order, and
        // runs of 64 consecutive words per wave would leave most waves idle
        // while a few walk dense runs.
        for (u64 base = 0; base < nwords; base += nwaves * kWave) {
            const u64 wl = base + (u64)lane * nwaves + wave_gid;
            const u64 open_l = wl < nwords ? a.openw[wl] : 0ull;
            u64 todo = ballot(open_l != 0);
            // The walk is software-pipelined: the screen loads of the next open
            // word (vmin and the settled mask, both read for every open vertex)
            // go out before the current word's line is read and f runs, so
            // that their round trip overlaps the current word's. f touches
            // only the current word's vertices, whose loads are its own.
            int jn = -1;
            u32 vmn = 0xFFFFFFFFu, smn = 0xFFFFFFFFu;
            bool openn = false;
            auto prefetch = [&]() {
                jn = todo ? __ffsll((unsigned long long)todo) - 1 : -1;
                vmn = smn = 0xFFFFFFFFu;
                openn = false;
                if (jn >= 0) {
                    todo &= todo - 1;
                    const u64 vn = (base + (u64)jn * nwaves + wave_gid) * 64 + lane;
                    openn = (shfl64(open_l, jn) >> lane) & 1;
                    if (openn) {
                        vmn = a.vmin[vn];
                        smn = a.smask[vn];
                    }
                }
            };
            prefetch();
            while (jn >= 0) {
                const int j = jn;
                const u32 vm = vmn, sm = smn;
                const bool isopen = openn;
                prefetch();
                const u64 word = base + (u64)j * nwaves + wave_gid;
                const u64 v = word * 64 + lane;
                const bool act = isopen && vm <= zb;
                const u32 open = act ? ~sm : 0u;
                u32 A = 0, B = 0, R = 0, dmin = act ? 0xFFFFFFFFu : vm;
                if (open) {
                    const uint4 *line = reinterpret_cast<const uint4 *>(a.dl + v * kBatchRoots);
#pragma unroll 2
                    for (int q = 0; q < kBatchRoots / 4; ++q) {
                        const u32 lq = (open >> (4 * q)) & 0xFu;
                        if (!lq) continue;
                        const uint4 c = line[q];
                        const u32 bits[4] = {c.x, c.y, c.z, c.w};
                        const float4 x4 = *reinterpret_cast<const float4 *>(&rbk.b[0][4 * q]);
                        const float4 y4 = *reinterpret_cast<const float4 *>(&rbk.b[1][4 * q]);
                        const float4 z4 = *reinterpret_cast<const float4 *>(&rbk.b[zl][4 * q]);
                        const float4 u4 = *reinterpret_cast<const float4 *>(&rbk.u[4 * q]);
                        const f32 xs[4] = {x4.x, x4.y, x4.z, x4.w}, ys[4] = {y4.x, y4.y, y4.z, y4.w};
                        const f32 zs[4] = {z4.x, z4.y, z4.z, z4.w}, us[4] = {u4.x, u4.y, u4.z, u4.w};
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            if (!((lq >> i) & 1) || bits[i] >= kInfBits) continue;  // final or unreached
                            const f32 d = __uint_as_float(bits[i]);
                            const u32 bit = 1u << (4 * q + i);
                            R |= bit;
                            if (d >= xs[i] && d < ys[i]) A |= bit;
                            if (d >= ys[i] && d < zs[i]) B |= bit;
                            if (d >= ys[i]) dmin = min(dmin, __float_as_uint(d + us[i]));
                        }
                    }
                }
                f(word, v, A, B, dmin, act, R);
            }
        }
    };

    // this thread's root (a select chain: a runtime index into the argument
    // struct would copy it to scratch)
    u64 my_root = 0;
#pragma unroll
    for (int i = 0; i < kBatchRoots; ++i)
        if (gtid == (u64)i) my_root = a.roots[i];
    if (gtid < (u64)a.nroots) {  // block 0: thread gtid keeps root gtid's bucket state
        const u64 root = trav_id(a.perm, my_root);
        st_relaxed(&a.lab[root * kBatchRoots + gtid], root);  // distance 0.0f, parent = root
        st_relaxed(&a.dl[root * kBatchRoots + gtid], 0u);
        atomicMin(&a.vmin[root], __float_as_uint(rbk.u[rt]));  // reduced value of distance 0
        atomicOr((unsigned long long *)&a.openw[root >> 6], 1ull << (root & 63));
    }
    if (gtid == 0)
        for (int i = 0; i < 3; ++i)
            for (int r = 0; r < kBatchRoots; ++r) st_relaxed(&ctl->rmin[i][r], 0xFFFFFFFFu);
    bsync(ctl, 1, 0, 0);
    light_dest(target(0, 0));
    if (gtid < kWave) {  // step 0 queues the roots without lag (repeated roots share one item)
        // a lagging root is queued by the dense step before its bucket 0 (d = 0 in [hi_r, hi2_r) = [0, delta))
        const bool r = (i32)gtid < a.nroots && rbk.k[rt] == 0;
        push_item(a, r ? trav_id(a.perm, my_root) : 0u, r ? 1u << gtid : 0u, target(0, 0), wq);
    }
    q_flush(a, wq);
    bsync(ctl, 1, 1, 0);
    x = 1;

    for (;;) {  // shared buckets: root r's [lo_r, hi_r)
        for (;;) {  // light steps
            u32 cnt[2];
            u64 hub;
            const u64 total = queue_size(x, cnt, &hub);
            clear_ahead(x);
            if (total == 0 || ld_relaxed(&ctl->bar.error)) break;
            if (gtid == 0) atomicAdd((unsigned long long *)&ctl->light_iters, 1ull);
            const BTarget t = target(x, 0);
            const int buf = (int)(x & 1);
            light_dest(t);
            // selects, not runtime indexes into the argument struct (which would copy it to scratch)
            u32 *pb = buf ? a.pend[1] : a.pend[0];
            group_tier<kSegLight>(a, pb, buf ? a.fq[1][0] : a.fq[0][0], nullptr, cnt[0], t, improved, slot, wks[0], wq,
                                  ~0u, &ctl->grab[x & 1].v);
            wave_tier<kSegLight>(a, pb, buf ? a.fq[1][1] : a.fq[0][1], nullptr, cnt[1], t, improved, wks[0], wq);
            hub_tier<kSegLight>(a, buf ? a.fq[1][2] : a.fq[0][2], buf ? a.fq_off[1] : a.fq_off[0],
                                buf ? a.fq_len[1] : a.fq_len[0], buf ? a.fq_done[1] : a.fq_done[0], pb, hub, t,
                                improved, wks[0], wq);
            q_flush(a, wq);
            bsync(ctl, 2, gb, x);
            // every wave has finished this step's chunks; the counter's next user is step x + 2
            if (gtid == 0) st_relaxed(&ctl->grab[x & 1].v, 0u);
            tick(0, total);
            ++x;
        }
        if (ld_relaxed(&ctl->bar.error)) break;
        // dense step x: medium items of [lo_r, hi_r), mid, long and far edges
        // deferred, next bucket [hi_r, hi2_r) queued
        if (gtid == 0) atomicAdd((unsigned long long *)&ctl->buckets, 1ull);
        if (rt < kBatchRoots) {
#pragma unroll
            for (int c = 0; c < kDefer; ++c)
                if (!lpending[c]) rbk.ep[c][rt] = rbk.b[0][rt];
        }
        if (threadIdx.x < (u32)kDefer) rws[threadIdx.x] = 0u;
        if (rt < kBatchRoots) wcs[rt] = 0u;
        if (gtid == 0) st_relaxed(&ctl->plen.v, 0u);
        __syncthreads();
        {
            const BTarget t = target(x, 1);
            const int hs = (int)((x + 1) % 3);
            light_dest(t);
            wq.hs = hs;
#pragma unroll
            for (int c = 0; c < kDefer; ++c) wq.ls[c] = (int)(nflush[c] & 1);
            // defers class c's segment of v (the lane owns v in this pass: no
            // atomic needed): ORs H into its pending mask, queues v once
            // (`old`: the class's pending mask of v, read with v's other loads)
            auto defer = [&](auto cls, u32 v, u32 H, u32 len, u32 old) {
                constexpr int c = decltype(cls)::value;
                bool q = false;
                if (H && len) {
                    a.dpend[c][v] = old | H;
                    q = old == 0;
                    if (c == kDefer - 1) atomicOr(&rws[c], H);  // the window's roots (pull flushes read them)
                }
                if (c == kDefer - 1) {  // far-class sources per root of the window (pull flushes read them)
                    const u32 hf = (H && len) ? H : 0u;
                    u32 any = hf;
#pragma unroll
                    for (int o = 32; o >= 1; o >>= 1) any |= __shfl_xor(any, o);
                    while (any) {
                        const int r = __ffs(any) - 1;
                        any &= any - 1;
                        const u64 bm = ballot((hf >> r) & 1u);
                        if (lane == 0) atomicAdd(&wcs[r], (u32)__popcll(bm));
                    }
                }
                q_push32<kQDef0 + 2 * c>(a, wq, q && btier(len) == 0, v);
                q_push32<kQDef0 + 2 * c + 1>(a, wq, q && btier(len) == 1, v);
                hub_append(&ctl->lhub[c][wq.ls[c]].v, a.dq[c][2], a.dq_off[c], a.dq_len[c], q && btier(len) == 2,
                           v, len, a.n, &ctl->bar.error);
            };
            scan_open(2, [&](u64 word, u64 v, u32 H, u32 N, u32 dmin, bool act, u32 R) {
                if (act) a.vmin[v] = dmin;  // the open roots left are those >= hi_r
                {
                    const u64 na = __popcll(ballot(act)), nh = __popcll(ballot(H != 0));
                    if (lane == 0) work_add(wks[2], 1, na, nh);
                }
                if (!ballot(H | N)) return;
                // Every load of v goes out here, before the first store: the
                // compiler cannot move a load above a store to another array
                // (they may alias), so loads interleaved with the updates
                // below would each cost a round trip of their own.
                RowInfo r = {};
                u32 dp[kDefer] = {}, sm = 0, lw = 0;
                if (H) {
                    r = load_row_info(a.ri, (u32)v);
#pragma unroll
                    for (int c = 0; c < kDefer; ++c) dp[c] = a.dpend[c][v];
                    sm = a.smask[v];
                }
                if (N) lw = a.split[v];
                // the bucket's roots of v are final now; v closes when no reached root is open
                bool closed = false;
                if (H) {
                    a.smask[v] = sm | H;
                    closed = (R & ~H) == 0;
                }
                const u64 cm = ballot(closed);
                if (cm && lane == 0) atomicAnd((unsigned long long *)&a.openw[word], ~(unsigned long long)cm);
                push_item_lw(a, (u32)v, N, lw, t, wq);
                u64 b = 0, e = 0;
                u64 cut[kDefer + 1];  // class c is [cut[c], cut[c + 1])
                if (H) {
                    b = r.start + r.light;
                    e = r.start + r.cut[0];
#pragma unroll
                    for (int c = 0; c < kDefer; ++c) cut[c] = r.start + r.cut[c];
                    cut[kDefer] = r.end;
                } else {
#pragma unroll
                    for (int c = 0; c <= kDefer; ++c) cut[c] = 0;
                }
                defer(std::integral_constant<int, 0>(), (u32)v, H, (u32)(cut[1] - cut[0]), dp[0]);
                defer(std::integral_constant<int, 1>(), (u32)v, H, (u32)(cut[2] - cut[1]), dp[1]);
                defer(std::integral_constant<int, 2>(), (u32)v, H, (u32)(cut[3] - cut[2]), dp[2]);
                static_assert(kDefer == 3, "one defer call per class");
                const u32 len = (u32)(e - b);
                const bool item = H && len > 0;
                q_push64<kQMed0>(a, wq, item && btier(len) == 0, ((u64)H << 32) | (u32)v);
                q_push64<kQMed0 + 1>(a, wq, item && btier(len) == 1, ((u64)H << 32) | (u32)v);
                const u64 hslot = hub_append(&ctl->hhub[hs].v, a.hqh, a.hq_off, a.hq_len, item && btier(len) == 2,
                                             (u32)v, len, a.n, &ctl->bar.error);
                if (hslot != ~0ull) a.hq_mask[hslot] = H;
            });
            q_flush(a, wq);  // ends with a block barrier: rws is complete
            if (threadIdx.x == 0 && rws[kDefer - 1])
                atomicOr(&ctl->rw[kDefer - 1][wq.ls[kDefer - 1]], rws[kDefer - 1]);
            if (rt < kBatchRoots && wcs[rt]) atomicAdd(&ctl->wcnt[wq.ls[kDefer - 1]][rt], wcs[rt]);
            bsync(ctl, 3, gb, x);
            // The deferred queues are read between the two barriers of this
            // step: the next appends to these sets come after the second one.
            u32 lc[kDefer][2];
            u64 lh[kDefer];
#pragma unroll
            for (int c = 0; c < kDefer; ++c) {
                for (int tier = 0; tier < 2; ++tier) lc[c][tier] = min(ld_relaxed(&ctl->llen[c][wq.ls[c]][tier].v), (u32)a.n);
                lh[c] = ld_relaxed(&ctl->lhub[c][wq.ls[c]].v);
            }
            tick(1);
            group_tier<kSegMedium>(a, nullptr, nullptr, a.hq[0], min(ld_relaxed(&ctl->hlen[hs][0].v), (u32)a.n), t,
                                   improved, slot, wks[1], wq, ~0u, &ctl->mgrab.v);
            wave_tier<kSegMedium>(a, nullptr, nullptr, a.hq[1], min(ld_relaxed(&ctl->hlen[hs][1].v), (u32)a.n), t,
                                  improved, wks[1], wq);
            hub_tier<kSegMedium>(a, a.hqh, a.hq_off, a.hq_len, a.hq_mask, nullptr, ld_relaxed(&ctl->hhub[hs].v), t,
                                 improved, wks[1], wq);
            q_flush(a, wq);
            bsync(ctl, 4, gb, x);
            if (gtid == 0) st_relaxed(&ctl->mgrab.v, 0u);  // its next user is the next dense step
            tick(2, (u64)ld_relaxed(&ctl->hlen[hs][0].v) + ld_relaxed(&ctl->hlen[hs][1].v) +
                        hub_count(ld_relaxed(&ctl->hhub[hs].v)));
            // Flush a deferred class c when the next bucket reaches past
            // fl(epoch_c + W_c), where its relaxations may land, or when the
            // next bucket looks empty (a min step must see every label).
            bool anyp = false;
#pragma unroll
            for (int c = 0; c < kDefer; ++c) {
                lpending[c] = lc[c][0] || lc[c][1] || hub_count(lh[c]);
                anyp = anyp || lpending[c];
            }
            if (anyp) {
                u32 qc[2];
                u64 qh;
                const bool next_empty =
                    queue_size(x + 1, qc, &qh) == 0 && ld_relaxed(&ctl->nnext[(x + 1) % 3]) == 0;
                bool fl[kDefer], anyf = false;
                u64 items = 0;
#pragma unroll
                for (int c = 0; c < kDefer; ++c) {
                    // due when some root's next bucket reaches past fl(epoch_r + W_c)
                    bool due = next_empty;
                    for (int r = 0; r < a.nroots && !due; ++r) due = rbk.b[2][r] > rbk.ep[c][r] + a.dw[c];
                    fl[c] = lpending[c] && due;
                    anyf = anyf || fl[c];
                    if (fl[c]) items += (u64)lc[c][0] + lc[c][1] + hub_count(lh[c]);
                }
                if (anyf) {
                    // the far class is pulled when its window's items are a large share of the graph
                    constexpr int cf = kDefer - 1;
                    // A pull serves the roots at their peak: those with an offset (in the hub's
                    // component) that settled at least pull_ratio * nact sources in the window. A
                    // root that has reached few vertices (just started, or in another component)
                    // would keep every other vertex scanning its whole segment; it is pushed.
                    const u32 rwf = ld_relaxed(&ctl->rw[cf][wq.ls[cf]]);
                    u32 elig = 0;  // roots that settled at least pull_ratio * nact sources in the window
                    if (fl[cf] && a.pull_ratio > 0.0f) {
                        const f32 need = a.pull_ratio * (f32)a.nact;
                        for (int r = 0; r < a.nroots; ++r)
                            if (((a.pull_roots >> r) & 1u) && (f32)ld_relaxed(&ctl->wcnt[wq.ls[cf]][r]) >= need)
                                elig |= 1u << r;
                    }
                    const bool pull = elig != 0u;
                    // the window's other roots are pushed, their masks kept for the pull (cleared after it)
                    const u32 keep = pull ? ~elig : ~0u;
                    if (gtid == 0) {  // a class's other set was last read by its previous flush
#pragma unroll
                        for (int c = 0; c < kDefer; ++c)
                            if (fl[c]) {
                                const int o = wq.ls[c] ^ 1;
                                st_relaxed(&ctl->llen[c][o][0].v, 0u);
                                st_relaxed(&ctl->llen[c][o][1].v, 0u);
                                st_relaxed(&ctl->lhub[c][o].v, 0ull);
                                st_relaxed(&ctl->rw[c][o], 0u);
                                if (c == kDefer - 1)
                                    for (int r = 0; r < kBatchRoots; ++r) st_relaxed(&ctl->wcnt[o][r], 0u);
                            }
                        atomicAdd((unsigned long long *)&ctl->flushes, 1ull);
                        if (pull) pws[0] += 1;
                    }
                    auto flush_class = [&](auto cls) {
                        constexpr int c = decltype(cls)::value;
                        if (!fl[c] || (c == cf && pull && (rwf & keep) == 0u)) return;
                        const u32 kc = c == cf ? keep : ~0u;
                        group_tier<kSegDef + c>(a, a.dpend[c], a.dq[c][0], nullptr, lc[c][0], t, improved, slot,
                                                wks[3 + c], wq, kc, &ctl->fgrab[c].v);
                        wave_tier<kSegDef + c>(a, a.dpend[c], a.dq[c][1], nullptr, lc[c][1], t, improved, wks[3 + c],
                                               wq, kc);
                        hub_tier<kSegDef + c>(a, a.dq[c][2], a.dq_off[c], a.dq_len[c], a.dq_done[c], a.dpend[c], lh[c],
                                              t, improved, wks[3 + c], wq, kc);
                    };
                    flush_class(std::integral_constant<int, 0>());
                    flush_class(std::integral_constant<int, 1>());
                    flush_class(std::integral_constant<int, 2>());
                    static_assert(kDefer == 3, "one flush_class call per class");
                    if (pull) pull_stage_a<cf>(a, rbk, elig, t, improved, wq, pws);
                    q_flush(a, wq);
                    bsync(ctl, 7, gb, x);
                    if (gtid == 0)  // their next users are the next flushes
                        for (int c = 0; c < kDefer; ++c) st_relaxed(&ctl->fgrab[c].v, 0u);
                    if (pull) {
                        pull_stage_b<cf>(a, rbk, elig, min(ld_relaxed(&ctl->plen.v), (u32)a.n), t, improved, wq,
                                         pws);
                        q_flush(a, wq);
                        bsync(ctl, 8, gb, x);
                        pull_clear<cf>(a, lc[cf][0], lc[cf][1], hub_count(lh[cf]));
                        bsync(ctl, 9, gb, x);
                    }
                    tick(3, items);
#pragma unroll
                    for (int c = 0; c < kDefer; ++c)
                        if (fl[c]) {
                            ++nflush[c];
                            lpending[c] = false;
                        }
                }
            }
            if (gtid == 0) atomicAdd((unsigned long long *)&ctl->scans, 1ull);
            ++x;
        }
        {
            u32 cnt[2];
            u64 hub;
            // next bucket not empty: queued light work, or labels that entered it
            if (queue_size(x, cnt, &hub) || ld_relaxed(&ctl->nnext[x % 3])) {
                // every thread of the block has read this bucket's bounds (the
                // flush decision reads them after the last grid barrier) before
                // they change
                __syncthreads();
                if (rt < kBatchRoots) advance_root();
                __syncthreads();
                ++gb;
                continue;
            }
        }
        if (ld_relaxed(&ctl->bar.error)) break;
        // min step: each root's smallest open distance (every open label is >= hi2_r now)
        clear_ahead(x);
        {
            if (rt < kBatchRoots) rbk.smin[rt] = 0xFFFFFFFFu;
            __syncthreads();
            for (u64 w0 = wave_gid * kWave; w0 < nwords; w0 += nwaves * kWave) {
                const u64 wl = w0 + lane;
                const u64 open_l = wl < nwords ? a.openw[wl] : 0ull;
                u64 todo = ballot(open_l != 0);
                while (todo) {
                    const int j = __ffsll((unsigned long long)todo) - 1;
                    todo &= todo - 1;
                    const u64 v = (w0 + (u64)j) * 64 + lane;
                    const bool isopen = (shfl64(open_l, j) >> lane) & 1;
                    const u32 open = isopen ? ~a.smask[v] : 0u;
                    if (!open) continue;
                    const uint4 *line = reinterpret_cast<const uint4 *>(a.dl + v * kBatchRoots);
                    for (int q = 0; q < kBatchRoots / 4; ++q) {
                        const u32 lq = (open >> (4 * q)) & 0xFu;
                        if (!lq) continue;
                        const uint4 c = line[q];
                        const u32 bits[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            if (((lq >> i) & 1) && bits[i] < kInfBits) atomicMin(&rbk.smin[4 * q + i], bits[i]);
                    }
                }
            }
            __syncthreads();
            if (rt < kBatchRoots && rbk.smin[rt] != 0xFFFFFFFFu) atomicMin(&ctl->rmin[(x + 1) % 3][rt], rbk.smin[rt]);
        }
        bsync(ctl, 5, gb, x);
        ++x;
        if (gtid == 0) atomicAdd((unsigned long long *)&ctl->scans, 1ull);
        // jump: the fewest bucket advances J >= 1 after which some root's
        // bucket [lo_r, hi_r) holds that root's smallest open distance; no
        // root skips a bucket with open labels (each needs >= J advances)
        if (rt < kBatchRoots) {
            const u32 m = ld_relaxed(&ctl->rmin[x % 3][rt]);
            u32 j = 0xFFFFFFFFu;
            if (m < kInfBits) {
                i32 k = rbk.k[rt];
                f32 b1 = rbk.b[1][rt], b2 = rbk.b[2][rt];
                j = 0;
                do {
                    const f32 nb = k + 2 >= 0 ? next_bound(b2, a.delta) : 0.0f;
                    b1 = b2;
                    b2 = nb;
                    ++k;
                    ++j;
                } while (!(__uint_as_float(m) < b1));
            }
            rbk.jr[rt] = j;
        }
        __syncthreads();
        u32 jump = 0xFFFFFFFFu;
        for (int r = 0; r < kBatchRoots; ++r) jump = min(jump, rbk.jr[r]);
        if (jump == 0xFFFFFFFFu) {  // no open label left
            tick(4);
            break;
        }
        if (rt < kBatchRoots)
            for (u32 i = 0; i < jump; ++i) advance_root();
        __syncthreads();
        gb += jump;
        // collect step: queue every root's bucket [lo_r, hi_r)
        clear_ahead(x);
        {
            const BTarget t = target(x, 0);
            light_dest(t);
            scan_open(1, [&](u64 word, u64 v, u32 M, u32 B, u32 dmin, bool act, u32 R) {
                (void)word;
                (void)B;
                (void)dmin;
                (void)act;
                (void)R;
                push_item(a, (u32)v, M, t, wq);
            });
            q_flush(a, wq);
        }
        bsync(ctl, 6, gb, x);
        if (gtid == 0) atomicAdd((unsigned long long *)&ctl->scans, 1ull);
        tick(4);
        ++x;
    }
    if (lane == 0 && improved) atomicAdd((unsigned long long *)&ctl->improvements, (unsigned long long)improved);
    __syncthreads();
    if (threadIdx.x < 4 && pws[threadIdx.x])
        atomicAdd((unsigned long long *)&ctl->pull_work[threadIdx.x], (unsigned long long)pws[threadIdx.x]);
    if (threadIdx.x < 3 * (3 + kDefer) && wks[threadIdx.x / 3][threadIdx.x % 3])
        atomicAdd((unsigned long long *)&ctl->work[threadIdx.x / 3][threadIdx.x % 3],
                  (unsigned long long)wks[threadIdx.x / 3][threadIdx.x % 3]);
    if (timer) {
        st_relaxed(&ctl->nlog, (u32)tst[7]);
        for (int i = 0; i < 5; ++i) st_relaxed(&ctl->ticks[i], tst[2 + i]);
        st_relaxed(&ctl->ticks[5], wall_clock64() - tst[1]);
    }
    // the labels are exported by sssp_batch_export_kernel, launched next on the same stream
}

// Export in original ids, root-major (traversal.cu:333-340): unreached ->
// -1.0f / -1. A block takes 64 consecutive original ids, reads their label
// lines (the 32 roots' labels of a vertex are 256 contiguous bytes), turns
// them root-major in LDS and writes each root's 64 distances and parents as
// contiguous runs. Writing straight from the vertex-major labels would scatter
// single 4- and 8-byte stores over 32 output arrays. Runs after the traversal
// kernel, whose probe array is the dist storage.
constexpr int kExportIds = 64;
__global__ __launch_bounds__(256) void sssp_batch_export_kernel(const u64 *__restrict__ lab,
                                                                const u32 *__restrict__ perm,
                                                                const u32 *__restrict__ iperm, u64 n, int nroots,
                                                                f32 dbound, f32 *__restrict__ dist,
                                                                i64 *__restrict__ pred, SsspBatchCtl *ctl) {
    __shared__ f32 sd[kBatchRoots][kExportIds + 1];
    __shared__ i64 sp[kBatchRoots][kExportIds + 1];
    bool over = false;
    // load mapping: thread t takes id t / 4 and roots 8 (t % 4) .. 8 (t % 4) + 7
    const int j = (int)(threadIdx.x >> 2), q = (int)(threadIdx.x & 3);
    for (u64 o0 = (u64)blockIdx.x * kExportIds; o0 < n; o0 += (u64)gridDim.x * kExportIds) {
        const u64 o = o0 + (u64)j;
        if (o < n && 8 * q < nroots) {
            const u64 v = perm ? perm[o] : o;
            const uint4 *p = reinterpret_cast<const uint4 *>(lab + v * kBatchRoots + 8 * q);
            u64 l8[8];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const uint4 x = p[h];
                l8[2 * h] = ((u64)x.y << 32) | x.x;
                l8[2 * h + 1] = ((u64)x.w << 32) | x.z;
            }
            // A vertex's parent is often the same for many roots: a parent equal
            // to the previous root's reuses its original id, and only the others
            // are looked up (all at once, not one after the other).
            u32 oid[8];
            bool fresh[8];
#pragma unroll
            for (int i = 0; i < 8; ++i)
                fresh[i] = i == 0 || (u32)l8[i] != (u32)l8[i - 1] || (u32)(l8[i - 1] >> 32) >= kInfBits;
#pragma unroll
            for (int i = 0; i < 8; ++i)
                oid[i] = fresh[i] && (u32)(l8[i] >> 32) < kInfBits ? orig_id(iperm, (u32)l8[i]) : 0u;
#pragma unroll
            for (int i = 1; i < 8; ++i)
                if (!fresh[i]) oid[i] = oid[i - 1];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int r = 8 * q + i;
                const u64 l = l8[i];
                const bool reached = (u32)(l >> 32) < kInfBits;
                over = over || (reached && label_dist(l) >= dbound);
                sd[r][j] = reached ? label_dist(l) : -1.0f;
                sp[r][j] = reached ? (i64)oid[i] : -1;
            }
        }
        __syncthreads();
        const u64 m = min((u64)kExportIds, n - o0);
        for (int k = (int)threadIdx.x; k < nroots * kExportIds; k += (int)blockDim.x) {
            const int r = k / kExportIds, jj = k % kExportIds;
            if ((u64)jj < m) {
                dist[(u64)r * n + o0 + (u64)jj] = sd[r][jj];
                pred[(u64)r * n + o0 + (u64)jj] = sp[r][jj];
            }
        }
        __syncthreads();
    }
    if (over) atomicOr(&ctl->bar.error, 32u);
}

// Packs each row's bounds and splits into its 32-byte record.
__global__ __launch_bounds__(256) void sssp_batch_rowinfo_kernel(const u64 *__restrict__ rs,
                                                                 const u32 *__restrict__ split,
                                                                 const u32 *__restrict__ d0,
                                                                 const u32 *__restrict__ d1,
                                                                 const u32 *__restrict__ d2, u64 n,
                                                                 RowInfo *__restrict__ ri) {
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x) {
        RowInfo r;
        r.start = rs[v];
        r.light = split[v];
        r.cut[0] = d0[v];
        r.cut[1] = d1[v];
        r.cut[2] = d2[v];
        r.end = rs[v + 1];
        ri[v] = r;
    }
}

__global__ __launch_bounds__(256) void sssp_batch_clear_kernel(u64 *__restrict__ lab, u32 *__restrict__ dl,
                                                               u32 *__restrict__ pend, u32 *__restrict__ done,
                                                               u32 *__restrict__ masks, u32 *__restrict__ mwork,
                                                               u32 *__restrict__ lwork, u32 *__restrict__ fwork,
                                                               u64 *__restrict__ openw, u64 n) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x; t < n * kBatchRoots; t += stride) {
        lab[t] = kInfLabel;
        dl[t] = kInfBits;
        if (t < 2 * n) {
            pend[t] = 0;
            done[t] = 0;
            masks[t] = t < n ? 0u : 0xFFFFFFFFu;  // final-root masks, then vmin
        }
        if (t < n) {  // deferred pending masks and hub chunk counts of every class
            mwork[t] = 0;
            mwork[5 * n + t] = 0;
            lwork[t] = 0;
            lwork[5 * n + t] = 0;
            fwork[t] = 0;
            fwork[5 * n + t] = 0;
        }
        if (t < (n + 63) / 64) openw[t] = 0;
    }
}

// Per-root summary of a finished batch: reached vertices and the distance sum
// in 2^-24 fixed point (an integer sum, so it does not depend on the order).
__global__ __launch_bounds__(256) void sssp_batch_summary_kernel(const u64 *__restrict__ lab, u64 n, i32 nroots,
                                                                 u64 *__restrict__ out) {
    __shared__ unsigned long long cnt[kBatchRoots], sum[kBatchRoots];
    if (threadIdx.x < kBatchRoots) {
        cnt[threadIdx.x] = 0;
        sum[threadIdx.x] = 0;
    }
    __syncthreads();
    for (u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x; t < n * kBatchRoots; t += (u64)gridDim.x * blockDim.x) {
        const int r = (int)(t & (kBatchRoots - 1));
        if (r >= nroots) continue;
        const u32 bits = (u32)(lab[t] >> 32);
        if (bits >= kInfBits) continue;
        atomicAdd(&cnt[r], 1ull);
        atomicAdd(&sum[r], (unsigned long long)(__uint_as_float(bits) * 16777216.0f));
    }
    __syncthreads();
    if (threadIdx.x < kBatchRoots && cnt[threadIdx.x]) {
        atomicAdd((unsigned long long *)&out[threadIdx.x], cnt[threadIdx.x]);
        atomicAdd((unsigned long long *)&out[kBatchRoots + threadIdx.x], sum[threadIdx.x]);
    }
}

void sssp_batch_summary_enqueue(const Oned_Graph *g, const Sssp_Batch_State *s, u64 *out, hipStream_t stream) {
    Context &c = ctx();
    const hipStream_t st = stream ? stream : c.stream;
    G500_HIP(hipMemsetAsync(out, 0, 2 * kBatchRoots * sizeof(u64), st));
    sssp_batch_summary_kernel<<<(unsigned)c.num_cus * 8, 256, 0, st>>>(s->labels, g->nlocalverts, s->nroots, out);
    G500_HIP(hipGetLastError());
}

void sssp_batch_prepare(const Oned_Graph *g, Sssp_Batch_State *s, hipStream_t st, f32 *wmid_out, f32 *wlong_out,
                        f32 *wfar_out) {
    const Context &c = ctx();
    // class floors: delta <= W_0 (mid) <= W_1 (long) <= W_2 (far); a class whose
    // floor is clamped onto the next one is empty
    auto clampw = [](f32 w, f32 lo, f32 hi) { return w < lo ? lo : (w > hi ? hi : w); };
    const f32 wfar = s->far_weight > s->delta ? s->far_weight : s->delta;
    const f32 wlong = clampw(s->long_weight, s->delta, wfar);
    const f32 wmid = clampw(s->mid_weight, s->delta, wlong);
    if (wmid_out) *wmid_out = wmid;
    if (wlong_out) *wlong_out = wlong;
    if (wfar_out) *wfar_out = wfar;
    const u64 n = g->nlocalverts;
    // the split depends on the graph too: a state reused across graphs recomputes it
    const TraversalView tv = sssp_view(g);
    if (s->split_delta != s->delta || s->long_split_weight != wlong || s->mid_split_weight != wmid ||
        s->far_split_weight != wfar || s->split_graph != (const void *)tv.rs) {
        compute_split(tv.rs, tv.w, n, s->delta, s->split, st);
        compute_split(tv.rs, tv.w, n, wmid, s->mid_split, st);
        compute_split(tv.rs, tv.w, n, wlong, s->long_split, st);
        compute_split(tv.rs, tv.w, n, wfar, s->far_split, st);
        sssp_batch_rowinfo_kernel<<<grid_for(n, 256, (u64)c.num_cus * 32), 256, 0, st>>>(
            tv.rs, s->split, s->mid_split, s->long_split, s->far_split, n, (RowInfo *)s->row_info);
        G500_HIP(hipGetLastError());
        s->split_delta = s->delta;
        s->mid_split_weight = wmid;
        s->long_split_weight = wlong;
        s->far_split_weight = wfar;
        s->split_graph = (const void *)tv.rs;
    }
}

// Bucket lags from distance offsets: the root with the largest offset runs
// without lag, root r lags round((max - offset_r) / delta) buckets (at most
// kMaxLag); roots without an offset (null, negative or not finite) lag none.
// Lags only order the work: every lag gives the same distances. with_offset:
// the roots that have an offset (the ones pulled flushes may serve).
constexpr u32 kMaxLag = 1u << 16;
void sssp_batch_lags(const f32 *offsets, i32 nroots, f32 delta, u32 *lag, f32 *uoff, u32 *with_offset) {
    *with_offset = 0;
    f64 omax = -1.0;
    for (i32 i = 0; offsets && i < nroots; ++i)
        if (std::isfinite(offsets[i]) && offsets[i] >= 0.0f) omax = std::max(omax, (f64)offsets[i]);
    for (int i = 0; i < kBatchRoots; ++i) {
        u32 l = 0;
        if (omax >= 0.0 && i < nroots && std::isfinite(offsets[i]) && offsets[i] >= 0.0f) {
            const f64 q = std::floor((omax - (f64)offsets[i]) / (f64)delta + 0.5);
            l = q >= (f64)kMaxLag ? kMaxLag : (u32)q;
            *with_offset |= 1u << i;
        }
        lag[i] = l;
        uoff[i] = (f32)((f64)l * (f64)delta);
    }
}

bool sssp_batch_enqueue(const Oned_Graph *g, Sssp_Batch_State *s, const u64 *roots, const f32 *offsets, i32 nroots,
                        hipStream_t stream, int grid) {
    Context &c = ctx();
    if (nroots < 1 || nroots > kBatchRoots) {
        fail("batched SSSP takes 1..%d roots (got %d)", kBatchRoots, nroots);
        return false;
    }
    for (i32 i = 0; i < nroots; ++i)
        if (roots[i] >= g->nglobalverts) {
            fail("SSSP root %llu is outside graph vertex range [0, %llu)", (unsigned long long)roots[i],
                 (unsigned long long)g->nglobalverts);
            return false;
        }
    if (!(s->delta > 0.0f) || !(s->delta < 3.0e38f)) {
        fail("SSSP delta must be positive and finite (got %g)", (double)s->delta);
        return false;
    }
    if (!(s->long_weight >= 0.0f) || !(s->long_weight < 3.0e38f)) {
        fail("batched SSSP long_weight must be non-negative and finite (got %g)", (double)s->long_weight);
        return false;
    }
    if (!(s->mid_weight >= 0.0f) || !(s->mid_weight < 3.0e38f)) {
        fail("batched SSSP mid_weight must be non-negative and finite (got %g)", (double)s->mid_weight);
        return false;
    }
    if (!(s->far_weight >= 0.0f) || !(s->far_weight < 3.0e38f)) {
        fail("batched SSSP far_weight must be non-negative and finite (got %g)", (double)s->far_weight);
        return false;
    }
    const u64 n = g->nlocalverts;
    if (n > s->max_nlocalverts) {
        fail("batched SSSP state holds %llu vertices, graph has %llu", (unsigned long long)s->max_nlocalverts,
             (unsigned long long)n);
        return false;
    }
    const bool library = stream == nullptr;
    const hipStream_t st = library ? c.stream : stream;
    f32 wmid, wlong, wfar;
    sssp_batch_prepare(g, s, st, &wmid, &wlong, &wfar);
    const TraversalView tv = sssp_view(g);
    sssp_batch_clear_kernel<<<grid_for(n * kBatchRoots, 256, (u64)c.num_cus * 64), 256, 0, st>>>(
        s->labels, (u32 *)s->dist, s->pending, s->hub_aux, s->root_masks, s->mid_work, s->long_work, s->far_work,
        s->open_words, n);
    G500_HIP(hipGetLastError());
    G500_HIP(hipMemsetAsync(s->ctl, 0, sizeof(SsspBatchCtl), st));
    SsspBatchArgs a;
    a.n = n;
    a.rs = tv.rs;
    a.col = tv.col;
    a.w = tv.w;
    a.perm = tv.perm;
    a.iperm = tv.iperm;
    a.split = s->split;
    // deferred classes: 0 mid, 1 long, 2 far; work arrays [6n]: pending masks,
    // queue tiers 0-2, hub lengths, hub chunk counts
    a.dsplit[0] = s->mid_split;
    a.dsplit[1] = s->long_split;
    a.dsplit[2] = s->far_split;
    a.ri = (const RowInfo *)s->row_info;
    u32 *dwork[kDefer] = {s->mid_work, s->long_work, s->far_work};
    u64 *doff[kDefer] = {s->mid_offsets, s->long_offsets, s->far_offsets};
    for (int d = 0; d < kDefer; ++d) {
        a.dpend[d] = dwork[d];
        for (int t = 0; t < 3; ++t) a.dq[d][t] = dwork[d] + (u64)(1 + t) * n;
        a.dq_len[d] = dwork[d] + 4 * n;
        a.dq_done[d] = dwork[d] + 5 * n;
        a.dq_off[d] = doff[d];
    }
    a.lab = s->labels;
    a.dl = (u32 *)s->dist;
    a.smask = s->root_masks;
    a.vmin = s->root_masks + n;
    a.openw = s->open_words;
    for (int b = 0; b < 2; ++b) {
        a.pend[b] = s->pending + (u64)b * n;
        for (int t = 0; t < 3; ++t) a.fq[b][t] = s->frontier + (u64)(b * 3 + t) * n;
        a.fq_off[b] = s->hub_offsets + (u64)b * n;
        a.fq_len[b] = s->hub_lengths + (u64)b * n;
        a.fq_done[b] = s->hub_aux + (u64)b * n;
        a.hq[b] = s->heavy + (u64)b * n;
    }
    a.hqh = s->hub_vertices;
    a.hq_off = s->hub_offsets + 2 * n;
    a.hq_len = s->hub_lengths + 2 * n;
    a.hq_mask = s->hub_aux + 2 * n;
    a.ctl = (SsspBatchCtl *)s->ctl;
    a.dist = s->dist;
    a.pred = s->pred;
    memset(a.roots, 0, sizeof(a.roots));
    for (i32 i = 0; i < nroots; ++i) a.roots[i] = roots[i];
    a.nroots = nroots;
    a.delta = s->delta;
    a.pull_ratio = (f32)c.batch_pull;
    a.nact = (g->perm && g->notisolated) ? std::min<u64>(g->notisolated, n) : n;
    sssp_batch_lags(offsets, nroots, s->delta, a.lag, a.uoff, &a.pull_roots);
    a.dbound = g->rowstarts_s ? kPruneDistBound : __builtin_inff();
    a.dw[0] = wmid;
    a.dw[1] = wlong;
    a.dw[2] = wfar;
    const int block = ctx().batch_sssp_block;
    const size_t dyn = q_lds_bytes(block);
    if (grid <= 0) grid = cooperative_blocks((const void *)sssp_batch_kernel, block, dyn);
    if (failed()) return false;
    void *args[] = {&a};
    s->nroots = nroots;
    // the export's grid: as many blocks as the persistent grid holds waves / 4, at most one per 64 ids
    const unsigned xgrid = (unsigned)std::max<u64>(1, std::min<u64>((n + kExportIds - 1) / kExportIds,
                                                                     (u64)grid * (u64)block / 256));
    auto export_labels = [&]() {
        sssp_batch_export_kernel<<<xgrid, 256, 0, st>>>(a.lab, a.perm, a.iperm, n, nroots, a.dbound, a.dist, a.pred,
                                                        a.ctl);
        G500_HIP(hipGetLastError());
    };
    if (library) {
        timer_start();
        launch_persistent((const void *)sssp_batch_kernel, grid, block, args, nullptr, true, dyn);
        export_labels();
        timer_stop();
    } else {
        launch_persistent((const void *)sssp_batch_kernel, grid, block, args, st, false, dyn);
        export_labels();
    }
    return !failed();
}

u32 sssp_batch_error(const Sssp_Batch_State *s, hipStream_t stream) {
    u32 e = 0;
    const hipStream_t st = stream ? stream : ctx().stream;
    G500_HIP(hipMemcpyAsync(&e, &((const SsspBatchCtl *)s->ctl)->bar.error, sizeof(u32), hipMemcpyDeviceToHost, st));
    G500_HIP(hipStreamSynchronize(st));
    return e;
}

int sssp_batch_grid() {
    const int block = ctx().batch_sssp_block;
    return cooperative_blocks((const void *)sssp_batch_kernel, block, q_lds_bytes(block));
}

}  // namespace g500

using namespace g500;

extern "C" {

Sssp_Batch_State oned_graph_sssp_batch_create(Worker_State *ws, const Oned_Graph *g) {
    Sssp_Batch_State s;
    memset(&s, 0, sizeof(s));
    const u64 n = g->nlocalverts;
    s.max_nlocalverts = n;
    s.delta = 0.1f;
    s.mid_weight = 0.075f;
    s.long_weight = 0.25f;
    s.far_weight = 0.5f;
    s.dist = (f32 *)arena_alloc(ws, kBatchRoots * n * sizeof(f32), "batched SSSP dist");
    s.pred = (i64 *)arena_alloc(ws, kBatchRoots * n * sizeof(i64), "batched SSSP pred");
    s.labels = (u64 *)arena_alloc(ws, kBatchRoots * n * sizeof(u64), "batched SSSP labels");
    s.pending = (u32 *)arena_alloc(ws, 2 * n * sizeof(u32), "batched SSSP pending masks");
    s.frontier = (u32 *)arena_alloc(ws, 6 * n * sizeof(u32), "batched SSSP queues");
    s.heavy = (u64 *)arena_alloc(ws, 2 * n * sizeof(u64), "batched SSSP heavy items");
    s.hub_vertices = (u32 *)arena_alloc(ws, n * sizeof(u32), "batched SSSP heavy hub vertices");
    s.hub_offsets = (u64 *)arena_alloc(ws, 3 * n * sizeof(u64), "batched SSSP hub offsets");
    s.hub_lengths = (u32 *)arena_alloc(ws, 3 * n * sizeof(u32), "batched SSSP hub lengths");
    s.hub_aux = (u32 *)arena_alloc(ws, 3 * n * sizeof(u32), "batched SSSP hub chunk counts / masks");
    s.split = (u32 *)arena_alloc(ws, n * sizeof(u32), "batched SSSP light split");
    s.root_masks = (u32 *)arena_alloc(ws, 2 * n * sizeof(u32), "batched SSSP final-root masks and smallest open probes");
    s.open_words = (u64 *)arena_alloc(ws, ((n + 63) / 64) * sizeof(u64), "batched SSSP open-vertex bitmap");
    s.mid_split = (u32 *)arena_alloc(ws, n * sizeof(u32), "batched SSSP mid split");
    s.mid_work = (u32 *)arena_alloc(ws, 6 * n * sizeof(u32), "batched SSSP deferred mid-edge queue");
    s.mid_offsets = (u64 *)arena_alloc(ws, n * sizeof(u64), "batched SSSP deferred mid hub offsets");
    s.far_split = (u32 *)arena_alloc(ws, n * sizeof(u32), "batched SSSP far split");
    s.far_work = (u32 *)arena_alloc(ws, 6 * n * sizeof(u32), "batched SSSP deferred far-edge queue");
    s.far_offsets = (u64 *)arena_alloc(ws, n * sizeof(u64), "batched SSSP deferred far hub offsets");
    s.long_split = (u32 *)arena_alloc(ws, n * sizeof(u32), "batched SSSP long split");
    s.long_work = (u32 *)arena_alloc(ws, 6 * n * sizeof(u32), "batched SSSP deferred long-edge queue");
    s.long_offsets = (u64 *)arena_alloc(ws, n * sizeof(u64), "batched SSSP deferred hub offsets");
    s.ctl = arena_alloc(ws, sizeof(SsspBatchCtl), "batched SSSP control");
    s.row_info = arena_alloc(ws, n * sizeof(RowInfo), "batched SSSP row records");
    return s;
}

void oned_graph_sssp_batch_run(Worker_State *ws, const Oned_Graph *g, Sssp_Batch_State *s, const u64 *roots,
                               i32 nroots) {
    oned_graph_sssp_batch_run_offsets(ws, g, s, roots, nullptr, nroots);
}

void oned_graph_sssp_batch_run_offsets(Worker_State *ws, const Oned_Graph *g, Sssp_Batch_State *s, const u64 *roots,
                                       const f32 *offsets, i32 nroots) {
    (void)ws;
    if (!sssp_batch_enqueue(g, s, roots, offsets, nroots, nullptr, 0)) return;
    const u32 e = sssp_batch_error(s, nullptr);
    if (e) {
        // where the blocks stood: the fewest and most grid barriers passed, with the phase and tags of each
        static SsspBatchCtl h;
        g500_copy_to_host(&h, s->ctl, sizeof(h));
        const int nb = std::min(sssp_batch_grid(), kMaxTraceBlocks);
        u32 lo = 0xFFFFFFFFu, hi = 0;
        int at_lo = 0, at_hi = 0;
        for (int b = 0; b < nb; ++b) {
            if (h.trace[b].barriers < lo) { lo = h.trace[b].barriers; at_lo = b; }
            if (h.trace[b].barriers > hi) { hi = h.trace[b].barriers; at_hi = b; }
        }
        const BlockTrace &tl = h.trace[at_lo], &th = h.trace[at_hi];
        fail("batched SSSP kernel error %u (1 barrier timeout, 2 hub list overflow, 4 bad queue entry, 8 queue "
             "overflow, 16 bucket invariant, 32 distance at or above 2048 on the pruned SSSP copy: build with "
             "sssp_prune 0); barriers passed min %u (block %d phase %u bucket %u step %u) max %u (block %d phase %u "
             "bucket %u step %u)",
             e, lo, at_lo, tl.phase, tl.a, tl.b, hi, at_hi, th.phase, th.a, th.b);
    }
}

void oned_graph_sssp_batch_destroy(Sssp_Batch_State *s) {
    if (s) memset(s, 0, sizeof(*s));
}

void g500_sssp_batch_phase_ms(const Sssp_Batch_State *s, f64 *ms) {
    static SsspBatchCtl h;
    g500_copy_to_host(&h, s->ctl, sizeof(h));
    const f64 khz = ctx().wall_clock_khz;
    ms[0] = (f64)h.ticks[0] / khz;
    ms[1] = (f64)(h.ticks[1] + h.ticks[2] + h.ticks[3]) / khz;
    ms[2] = (f64)h.ticks[4] / khz;
    ms[3] = (f64)h.ticks[5] / khz;
    ms[4] = (f64)h.scans;
    ms[5] = (f64)h.flushes;
    ms[6] = (f64)h.ticks[1] / khz;
    ms[7] = (f64)h.ticks[2] / khz;
    ms[8] = (f64)h.ticks[3] / khz;
}

i32 g500_sssp_batch_step_log(const Sssp_Batch_State *s, i32 *kind, u64 *items, f64 *ms, i32 max_steps) {
    static SsspBatchCtl h;
    g500_copy_to_host(&h, s->ctl, sizeof(h));
    const f64 khz = ctx().wall_clock_khz;
    i32 k = 0;
    for (u32 i = 0; i < MIN(h.nlog, kBStepLog) && k < max_steps; ++i, ++k) {
        if (kind) kind[k] = (i32)h.log_kind[i];
        if (items) items[k] = h.log_items[i];
        if (ms) ms[k] = (f64)h.log_ticks[i] / khz;
    }
    return k;
}

void g500_sssp_batch_work(const Sssp_Batch_State *s, u64 *out) {
    static SsspBatchCtl h;
    g500_copy_to_host(&h, s->ctl, sizeof(h));
    for (int k = 0; k < 3 + kDefer; ++k)
        for (int j = 0; j < 3; ++j) out[k * 3 + j] = h.work[k][j];
}

void g500_sssp_batch_pull_work(const Sssp_Batch_State *s, u64 *out) {
    static SsspBatchCtl h;
    g500_copy_to_host(&h, s->ctl, sizeof(h));
    for (int i = 0; i < 4; ++i) out[i] = h.pull_work[i];
}

void g500_sssp_batch_run_stats(const Sssp_Batch_State *s, u64 *buckets, u64 *light_iterations, u64 *relaxations) {
    static SsspBatchCtl h;
    g500_copy_to_host(&h, s->ctl, sizeof(h));
    if (buckets) *buckets = h.buckets;
    if (light_iterations) *light_iterations = h.light_iters;
    if (relaxations) *relaxations = h.improvements;
}

}  // extern "C"
