// CSR construction from the tuple list and the isolated-vertex query.
// Reference: src/traversal.cu:15-73 (degree / rowstart / fill kernels),
// :342-410 (oned_graph_from_tuple_graph, oned_graph_is_vertex_isolated).
//
// Steps: (1) degree histogram with u32 device atomics, self-loops skipped;
// (2) three-phase exclusive scan into u64 row offsets; (3) fill of both
// directions of every non-loop tuple edge through per-row cursors; (4) each
// row sorted ascending by the key (f32 weight bits << 32 | column), which
// makes the CSR deterministic and puts light edges first in every row. Rows
// are sorted by one of four tiers: a thread (<= 16 entries), a wave
// (<= 64), a block in LDS (<= 4096), or a block-wide LSD radix sort through a
// scratch buffer (longer rows).
#include <algorithm>
#include <cstdio>
#include <vector>

#include "common.hpp"

namespace g500 {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ u64 make_key(u32 col, f32 w) { return ((u64)__float_as_uint(w) << 32) | col; }

// Tuples with an endpoint >= n are skipped and flagged (the entry point then
// fails) instead of indexing out of bounds.
__global__ __launch_bounds__(256) void degree_kernel(const Packed_Edge *__restrict__ edges, u64 ne, u64 n,
                                                     u32 *__restrict__ deg, u32 *__restrict__ bad) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += (u64)gridDim.x * blockDim.x) {
        Packed_Edge e = edges[i];
        if (e.v0 >= n || e.v1 >= n) {
            atomicOr(bad, 1u);
            continue;
        }
        if (e.v0 == e.v1) continue;
        atomicAdd(&deg[e.v0], 1u);
        atomicAdd(&deg[e.v1], 1u);
    }
}

__device__ __forceinline__ u64 shfl_up64(u64 v, int d) {
    u32 lo = __shfl_up((u32)v, d), hi = __shfl_up((u32)(v >> 32), d);
    return ((u64)hi << 32) | lo;
}

// Exclusive scan over the block; *total receives the block sum.
template <int BLOCK>
__device__ __forceinline__ u64 block_exclusive_scan(u64 v, u64 *total) {
    __shared__ u64 wtot[BLOCK / kWave];
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    u64 x = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        u64 y = shfl_up64(x, o);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) wtot[wave] = x;
    __syncthreads();
    if (wave == 0) {
        u64 t = lane < BLOCK / kWave ? wtot[lane] : 0;
#pragma unroll
        for (int o = 1; o < BLOCK / kWave; o <<= 1) {
            u64 y = shfl_up64(t, o);
            if (lane >= o) t += y;
        }
        if (lane < BLOCK / kWave) wtot[lane] = t;
    }
    __syncthreads();
    u64 before = wave ? wtot[wave - 1] : 0;
    *total = wtot[BLOCK / kWave - 1];
    __syncthreads();
    return before + x - v;
}

__global__ __launch_bounds__(kScanBlock) void scan_tile_sums(const u32 *__restrict__ deg, u64 n,
                                                             u64 *__restrict__ tile_sums) {
    const u64 base = (u64)blockIdx.x * kScanTile;
    u64 s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        u64 i = base + threadIdx.x + (u64)j * kScanBlock;
        if (i < n) s += deg[i];
    }
    u64 tot;
    block_exclusive_scan<kScanBlock>(s, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// Single block: exclusive scan of the tile sums in place; rowstarts[n] = total.
__global__ __launch_bounds__(1024) void scan_tile_offsets(u64 *__restrict__ tile_sums, u64 ntiles,
                                                          u64 *__restrict__ rowstarts, u64 n) {
    u64 carry = 0;
    for (u64 b = 0; b < ntiles; b += 1024) {
        u64 i = b + threadIdx.x;
        u64 v = i < ntiles ? tile_sums[i] : 0;
        u64 tot;
        u64 ex = block_exclusive_scan<1024>(v, &tot);
        if (i < ntiles) tile_sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) rowstarts[n] = carry;
}

__global__ __launch_bounds__(kScanBlock) void scan_write(const u32 *__restrict__ deg, u64 n,
                                                         const u64 *__restrict__ tile_off,
                                                         u64 *__restrict__ rowstarts) {
    const u64 base = (u64)blockIdx.x * kScanTile + (u64)threadIdx.x * kScanItems;
    u32 d[kScanItems];
    u64 s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        d[j] = base + j < n ? deg[base + j] : 0;
        s += d[j];
    }
    u64 tot;
    u64 run = tile_off[blockIdx.x] + block_exclusive_scan<kScanBlock>(s, &tot);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < n) rowstarts[base + j] = run;
        run += d[j];
    }
}

__global__ __launch_bounds__(256) void fill_kernel(const Packed_Edge *__restrict__ edges,
                                                   const f32 *__restrict__ ew, u64 ne, u64 n,
                                                   const u64 *__restrict__ rowstarts,
                                                   u32 *__restrict__ cursor, u32 *__restrict__ col,
                                                   f32 *__restrict__ cw) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += (u64)gridDim.x * blockDim.x) {
        Packed_Edge e = edges[i];
        if (e.v0 == e.v1 || e.v0 >= n || e.v1 >= n) continue;
        f32 w = ew[i];
        u64 p0 = rowstarts[e.v0] + atomicAdd(&cursor[e.v0], 1u);
        u64 p1 = rowstarts[e.v1] + atomicAdd(&cursor[e.v1], 1u);
        col[p0] = (u32)e.v1;
        cw[p0] = w;
        col[p1] = (u32)e.v0;
        cw[p1] = w;
    }
}

// ------------------------------------------------------------------ row sort
struct SortLists {
    u32 n1, n2, n3;       // rows in tiers 1..3
    u32 pad;
    u64 scratch_total;    // entries of tier-3 rows
    u64 notisolated;
};

__global__ __launch_bounds__(256) void classify_rows(const u64 *__restrict__ rs, u64 n, SortLists *L,
                                                     u32 *__restrict__ t1, u32 *__restrict__ t2,
                                                     u32 *__restrict__ t3, u64 *__restrict__ t3_off) {
    u32 local_nonzero = 0;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    const u64 nround = (n + stride - 1) / stride * stride;  // every lane runs the same trip count
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < nround; v += stride) {
        u64 d = v < n ? rs[v + 1] - rs[v] : 0;
        local_nonzero += d > 0;
        bool a = d > 16 && d <= 64, b = d > 64 && d <= 4096, c = d > 4096;
        u32 s1 = wave_append(&L->n1, a);
        u32 s2 = wave_append(&L->n2, b);
        u32 s3 = wave_append(&L->n3, c);
        if (a) t1[s1] = (u32)v;
        if (b) t2[s2] = (u32)v;
        if (c) {
            t3[s3] = (u32)v;
            t3_off[s3] = atomicAdd((unsigned long long *)&L->scratch_total, (unsigned long long)d);
        }
    }
    u32 tot = wave_sum(local_nonzero);
    if (lane_id() == 0 && tot) atomicAdd((unsigned long long *)&L->notisolated, (unsigned long long)tot);
}

// Tier 0: rows of 2..16 entries, one thread each, bitonic network in registers.
__global__ __launch_bounds__(256) void sort_rows_thread(const u64 *__restrict__ rs, u64 n,
                                                        u32 *__restrict__ col, f32 *__restrict__ cw) {
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (u64)gridDim.x * blockDim.x) {
        const u64 b = rs[v];
        const u32 d = (u32)(rs[v + 1] - b);
        if (d < 2 || d > 16) continue;
        u64 k[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) k[i] = (u32)i < d ? make_key(col[b + i], cw[b + i]) : ~0ull;
#pragma unroll
        for (int size = 2; size <= 16; size <<= 1)
#pragma unroll
            for (int stride = size / 2; stride > 0; stride >>= 1)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int j = i ^ stride;
                    if (j > i) {
                        const bool up = (i & size) == 0;
                        if ((k[i] > k[j]) == up) {
                            u64 t = k[i];
                            k[i] = k[j];
                            k[j] = t;
                        }
                    }
                }
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if ((u32)i < d) {
                col[b + i] = (u32)k[i];
                cw[b + i] = __uint_as_float((u32)(k[i] >> 32));
            }
    }
}

// Tier 1: rows of 17..64 entries, one wave each, bitonic across lanes.
__global__ __launch_bounds__(256) void sort_rows_wave(const u64 *__restrict__ rs, const u32 *__restrict__ rows,
                                                      const SortLists *L, u32 *__restrict__ col,
                                                      f32 *__restrict__ cw) {
    const u32 count = L->n1;
    const int lane = lane_id();
    const u32 waves = gridDim.x * (blockDim.x / kWave);
    for (u32 r = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < count; r += waves) {
        const u32 v = rows[r];
        const u64 b = rs[v];
        const u32 d = (u32)(rs[v + 1] - b);
        u64 x = (u32)lane < d ? make_key(col[b + lane], cw[b + lane]) : ~0ull;
#pragma unroll
        for (int size = 2; size <= kWave; size <<= 1)
#pragma unroll
            for (int stride = size / 2; stride > 0; stride >>= 1) {
                u64 y = shfl_xor64(x, stride);
                const bool up = (lane & size) == 0, lower = (lane & stride) == 0;
                x = (lower == up) ? (x < y ? x : y) : (x < y ? y : x);
            }
        if ((u32)lane < d) {
            col[b + lane] = (u32)x;
            cw[b + lane] = __uint_as_float((u32)(x >> 32));
        }
    }
}

// Tier 2: rows of 65..4096 entries, one 256-thread block each, bitonic in LDS.
__global__ __launch_bounds__(256) void sort_rows_block(const u64 *__restrict__ rs, const u32 *__restrict__ rows,
                                                       const SortLists *L, u32 *__restrict__ col,
                                                       f32 *__restrict__ cw) {
    __shared__ u64 keys[4096];
    const u32 count = L->n2;
    for (u32 r = blockIdx.x; r < count; r += gridDim.x) {
        const u32 v = rows[r];
        const u64 b = rs[v];
        const u32 d = (u32)(rs[v + 1] - b);
        u32 P = 128;
        while (P < d) P <<= 1;
        for (u32 i = threadIdx.x; i < P; i += 256) keys[i] = i < d ? make_key(col[b + i], cw[b + i]) : ~0ull;
        __syncthreads();
        for (u32 size = 2; size <= P; size <<= 1)
            for (u32 stride = size >> 1; stride > 0; stride >>= 1) {
                for (u32 t = threadIdx.x; t < P / 2; t += 256) {
                    const u32 i = 2 * stride * (t / stride) + (t % stride), j = i + stride;
                    const bool up = (i & size) == 0;
                    const u64 a = keys[i], c = keys[j];
                    if ((a > c) == up) {
                        keys[i] = c;
                        keys[j] = a;
                    }
                }
                __syncthreads();
            }
        for (u32 i = threadIdx.x; i < d; i += 256) {
            col[b + i] = (u32)keys[i];
            cw[b + i] = __uint_as_float((u32)(keys[i] >> 32));
        }
        __syncthreads();
    }
}

// Tier 3: rows longer than 4096 entries, one 1024-thread block each, stable
// LSD radix sort on 8-bit digits ping-ponging between the row (column /
// weight-bit arrays) and a scratch pair. Digits cover the column bits, then
// the 30 significant weight bits.
constexpr int kRadixBlock = 1024;
constexpr int kRadixWaves = kRadixBlock / kWave;

__device__ __forceinline__ u32 digit_of(u32 lo, u32 hi, int shift) {
    return shift < 32 ? (lo >> shift) & 0xFF : (hi >> (shift - 32)) & 0xFF;
}

__global__ __launch_bounds__(kRadixBlock) void sort_rows_radix(
    const u64 *__restrict__ rs, const u32 *__restrict__ rows, const u64 *__restrict__ row_off,
    const SortLists *L, u32 *__restrict__ col, u32 *__restrict__ wbits, u32 *__restrict__ s_lo,
    u32 *__restrict__ s_hi, int col_passes) {
    __shared__ u32 base[256];
    __shared__ u32 wcnt[kRadixWaves][256];
    const u32 count = L->n3;
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    const int npass = col_passes + 4;
    for (u32 r = blockIdx.x; r < count; r += gridDim.x) {
        const u32 v = rows[r];
        const u64 b = rs[v];
        const u32 d = (u32)(rs[v + 1] - b);
        u32 *a_lo = col + b, *a_hi = wbits + b;
        u32 *b_lo = s_lo + row_off[r], *b_hi = s_hi + row_off[r];
        for (int pass = 0; pass < npass; ++pass) {
            const int shift = pass < col_passes ? 8 * pass : 32 + 8 * (pass - col_passes);
            const u32 *src_lo = (pass & 1) ? b_lo : a_lo, *src_hi = (pass & 1) ? b_hi : a_hi;
            u32 *dst_lo = (pass & 1) ? a_lo : b_lo, *dst_hi = (pass & 1) ? a_hi : b_hi;
            if (threadIdx.x < 256) base[threadIdx.x] = 0;
            __syncthreads();
            for (u32 i = threadIdx.x; i < d; i += kRadixBlock)
                atomicAdd(&base[digit_of(src_lo[i], src_hi[i], shift)], 1u);
            __syncthreads();
            if (threadIdx.x < 256) {  // exclusive scan of the 256 bins, 4 waves
                u32 x = base[threadIdx.x];
                u32 incl = x;
                for (int o = 1; o < kWave; o <<= 1) {
                    u32 y = __shfl_up(incl, o);
                    if (lane >= o) incl += y;
                }
                wcnt[0][threadIdx.x] = incl;  // staging of wave-inclusive sums
            }
            __syncthreads();
            if (threadIdx.x < 256) {
                u32 add = 0;
                for (int w = 0; w < (int)(threadIdx.x / kWave); ++w) add += wcnt[0][w * kWave + kWave - 1];
                base[threadIdx.x] = add + wcnt[0][threadIdx.x] - base[threadIdx.x];
            }
            __syncthreads();
            for (u32 t0 = 0; t0 < d; t0 += kRadixBlock) {
                const u32 i = t0 + threadIdx.x;
                const bool valid = i < d;
                u32 lo = 0, hi = 0, dg = 0;
                if (valid) {
                    lo = src_lo[i];
                    hi = src_hi[i];
                    dg = digit_of(lo, hi, shift);
                }
                for (int k = threadIdx.x; k < kRadixWaves * 256; k += kRadixBlock) (&wcnt[0][0])[k] = 0;
                u64 peers = ballot(valid);
#pragma unroll
                for (int bit = 0; bit < 8; ++bit) {
                    const u64 m = ballot((dg >> bit) & 1);
                    peers &= ((dg >> bit) & 1) ? m : ~m;
                }
                const u32 rank = mask_rank(peers);
                __syncthreads();
                if (valid && rank == 0) wcnt[wave][dg] = (u32)__popcll(peers);
                __syncthreads();
                if (threadIdx.x < 256) {
                    u32 run = base[threadIdx.x];
                    for (int w = 0; w < kRadixWaves; ++w) {
                        const u32 c = wcnt[w][threadIdx.x];
                        wcnt[w][threadIdx.x] = run;
                        run += c;
                    }
                    base[threadIdx.x] = run;
                }
                __syncthreads();
                if (valid) {
                    const u32 pos = wcnt[wave][dg] + rank;
                    dst_lo[pos] = lo;
                    dst_hi[pos] = hi;
                }
                __syncthreads();
            }
        }
        if (npass & 1) {  // result sits in the scratch pair: copy back
            for (u32 i = threadIdx.x; i < d; i += kRadixBlock) {
                a_lo[i] = b_lo[i];
                a_hi[i] = b_hi[i];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- relabeling
// Traversal copy of the CSR with vertices renumbered by decreasing degree
// (exact up to kDegKeys - 1, above that in one class). In Kronecker graphs a
// few percent of the vertices hold most edge endpoints; numbering them first
// packs the labels, bitmap words and masks that traversals probe most into a
// small, cache-resident range. Entry order inside each row is unchanged, so
// rows stay sorted by weight.
constexpr u32 kDegKeys = 4096;
constexpr int kRelabelBlock = 1024;
constexpr int kRelabelItems = 8;

__device__ __forceinline__ u32 degree_key(const u64 *rs, u64 v) {
    const u64 d = rs[v + 1] - rs[v];
    return kDegKeys - 1 - (u32)(d < kDegKeys - 1 ? d : kDegKeys - 1);
}

__global__ __launch_bounds__(kRelabelBlock) void relabel_count(const u64 *__restrict__ rs, u64 n,
                                                               u32 *__restrict__ counts) {
    __shared__ u32 h[kDegKeys];
    for (u32 k = threadIdx.x; k < kDegKeys; k += blockDim.x) h[k] = 0;
    __syncthreads();
    const u64 base = (u64)blockIdx.x * kRelabelBlock * kRelabelItems;
    for (int j = 0; j < kRelabelItems; ++j) {
        const u64 v = base + (u64)j * kRelabelBlock + threadIdx.x;
        if (v < n) atomicAdd(&h[degree_key(rs, v)], 1u);
    }
    __syncthreads();
    for (u32 k = threadIdx.x; k < kDegKeys; k += blockDim.x)
        if (h[k]) atomicAdd(&counts[k], h[k]);
}

// Single block: exclusive scan of the key counts into cursors.
__global__ __launch_bounds__(1024) void relabel_offsets(const u32 *__restrict__ counts, u32 *__restrict__ cursor) {
    u64 carry = 0;
    for (u32 b = 0; b < kDegKeys; b += 1024) {
        const u32 k = b + threadIdx.x;
        const u64 v = counts[k];
        u64 tot;
        const u64 ex = block_exclusive_scan<1024>(v, &tot);
        cursor[k] = (u32)(carry + ex);
        carry += tot;
    }
}

// Each block reserves one range per key for its tile, then ranks its
// vertices inside the ranges (order inside a key class is not specified).
__global__ __launch_bounds__(kRelabelBlock) void relabel_scatter(const u64 *__restrict__ rs, u64 n,
                                                                 u32 *__restrict__ cursor, u32 *__restrict__ perm,
                                                                 u32 *__restrict__ iperm, u32 *__restrict__ deg_t) {
    __shared__ u32 h[kDegKeys], base[kDegKeys];
    for (u32 k = threadIdx.x; k < kDegKeys; k += blockDim.x) h[k] = 0;
    __syncthreads();
    const u64 tile = (u64)blockIdx.x * kRelabelBlock * kRelabelItems;
    u32 key[kRelabelItems];
    for (int j = 0; j < kRelabelItems; ++j) {
        const u64 v = tile + (u64)j * kRelabelBlock + threadIdx.x;
        key[j] = v < n ? degree_key(rs, v) : 0;
        if (v < n) atomicAdd(&h[key[j]], 1u);
    }
    __syncthreads();
    for (u32 k = threadIdx.x; k < kDegKeys; k += blockDim.x) {
        base[k] = h[k] ? atomicAdd(&cursor[k], h[k]) : 0;
        h[k] = 0;
    }
    __syncthreads();
    for (int j = 0; j < kRelabelItems; ++j) {
        const u64 v = tile + (u64)j * kRelabelBlock + threadIdx.x;
        if (v >= n) continue;
        const u32 pos = base[key[j]] + atomicAdd(&h[key[j]], 1u);
        perm[v] = pos;
        iperm[pos] = (u32)v;
        deg_t[pos] = (u32)(rs[v + 1] - rs[v]);
    }
}

// Rows of the traversal copy: a lane per row up to kWave entries, a block per
// longer row (they come first: rows are in decreasing degree order).
__global__ __launch_bounds__(256) void relabel_rows_lane(const u64 *__restrict__ rs, const u32 *__restrict__ col,
                                                         const f32 *__restrict__ w, const u32 *__restrict__ perm,
                                                         const u32 *__restrict__ iperm, const u64 *__restrict__ rs_t,
                                                         u64 first, u64 n, u32 *__restrict__ col_t,
                                                         f32 *__restrict__ w_t) {
    for (u64 p = first + (u64)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (u64)gridDim.x * blockDim.x) {
        const u32 v = iperm[p];
        const u64 b = rs[v], e = rs[v + 1], o = rs_t[p];
        for (u64 k = b; k < e; ++k) {
            col_t[o + (k - b)] = perm[col[k]];
            w_t[o + (k - b)] = w[k];
        }
    }
}

__global__ __launch_bounds__(256) void relabel_rows_block(const u64 *__restrict__ rs, const u32 *__restrict__ col,
                                                          const f32 *__restrict__ w, const u32 *__restrict__ perm,
                                                          const u32 *__restrict__ iperm, const u64 *__restrict__ rs_t,
                                                          u64 count, u32 *__restrict__ col_t, f32 *__restrict__ w_t) {
    for (u64 p = blockIdx.x; p < count; p += gridDim.x) {
        const u32 v = iperm[p];
        const u64 b = rs[v], e = rs[v + 1], o = rs_t[p];
        for (u64 k = b + threadIdx.x; k < e; k += blockDim.x) {
            col_t[o + (k - b)] = perm[col[k]];
            w_t[o + (k - b)] = w[k];
        }
    }
}

// ---------------------------------------------------------------- SSSP copy
// An entry (u, v, w) of the traversal copy is left out of the SSSP copy when
// a two-hop path u-x-v through one of the kPruneHubs highest-degree vertices
// (ties to the smaller traversal id) is lighter by more than kPruneMargin:
// hw[u][x] + hw[v][x] + kPruneMargin < w, with hw[y][x] the lightest weight of
// an entry (y, x) (+inf without one), summed in f64 (exact for f32 weights
// below kPruneMaxWeight that are multiples of 2^-31, as generated weights are;
// otherwise off by far less than the margin).
//
// Why distances do not change: both hops of a dropped entry are lighter than
// it, so by induction on weight every dropped entry is beaten by a path of
// kept entries that is lighter by more than (hops - 1) margins. Along that
// path f32 rounding costs at most half an ulp per addition, and comparing
// with fl(d + w) costs another half: with every value below 2048 + 2, an ulp
// is at most 2^-12, and (h - 1) * 2^-10 > (h + 1) * 2^-13 for every h >= 2.
// So fl(d(u) + w) is never below what the kept path gives v, the entry never
// decides a distance, and the f32 fixpoint (the distances) is the same with
// or without it. Parents stay valid for the full graph's certificate.
constexpr int kPruneHubs = 32;
constexpr double kPruneMargin = 1.0 / 1024.0;
constexpr f32 kPruneMaxWeight = 2.0f;

// Hub slot of traversal id x (ids below nslot are looked up; kPruneHubs: none).
__device__ __forceinline__ u32 hub_slot_of(const u8 *slot, u32 nslot, u32 x) {
    return x < nslot ? (u32)slot[x] : (u32)kPruneHubs;
}

// hw[u][slot of x] (f32 bits; non-negative floats order as their bits): a
// lane per row from row `first` on; a block per row for the rows before it.
__global__ __launch_bounds__(256) void prune_hubs_lane(const u64 *__restrict__ rs, const u32 *__restrict__ col,
                                                       const f32 *__restrict__ w, const u8 *__restrict__ slot,
                                                       u32 nslot, u64 first, u64 n, u32 *__restrict__ hw) {
    for (u64 u = first + (u64)blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64)gridDim.x * blockDim.x) {
        const u64 b = rs[u], e = rs[u + 1];
        for (u64 k = b; k < e; ++k) {
            const u32 x = hub_slot_of(slot, nslot, col[k]);
            if (x < (u32)kPruneHubs) {
                u32 *p = &hw[u * kPruneHubs + x];
                const u32 wb = __float_as_uint(w[k]);
                if (wb < *p) *p = wb;  // only this lane writes row u
            }
        }
    }
}

__global__ __launch_bounds__(256) void prune_hubs_block(const u64 *__restrict__ rs, const u32 *__restrict__ col,
                                                        const f32 *__restrict__ w, const u8 *__restrict__ slot,
                                                        u32 nslot, u64 count, u32 *__restrict__ hw) {
    for (u64 u = blockIdx.x; u < count; u += gridDim.x) {
        const u64 b = rs[u], e = rs[u + 1];
        for (u64 k = b + threadIdx.x; k < e; k += blockDim.x) {
            const u32 x = hub_slot_of(slot, nslot, col[k]);
            if (x < (u32)kPruneHubs) atomicMin(&hw[u * kPruneHubs + x], __float_as_uint(w[k]));
        }
    }
}

// Row u's hub weights, and their minimum (an entry lighter than it plus the
// margin cannot be beaten).
struct HubRow {
    f32 h[kPruneHubs];
    f32 hmin;
};
__device__ __forceinline__ void load_hub_row(const u32 *hw, u64 u, HubRow &r) {
    const uint4 *p = reinterpret_cast<const uint4 *>(hw + u * kPruneHubs);
    r.hmin = __uint_as_float(0x7F800000u);
#pragma unroll
    for (int q = 0; q < kPruneHubs / 4; ++q) {
        const uint4 c = p[q];
        r.h[4 * q] = __uint_as_float(c.x);
        r.h[4 * q + 1] = __uint_as_float(c.y);
        r.h[4 * q + 2] = __uint_as_float(c.z);
        r.h[4 * q + 3] = __uint_as_float(c.w);
#pragma unroll
        for (int i = 0; i < 4; ++i) r.hmin = fminf(r.hmin, r.h[4 * q + i]);
    }
}

// Whether entry (u, v, w) of row u is beaten by a two-hop path through a hub.
__device__ __forceinline__ bool prune_entry(const HubRow &r, const u32 *hw, u32 v, f32 w) {
    if (!(w < kPruneMaxWeight) || !((double)r.hmin + kPruneMargin < (double)w)) return false;
    const uint4 *p = reinterpret_cast<const uint4 *>(hw + (u64)v * kPruneHubs);
    double best = __builtin_inf();
#pragma unroll
    for (int q = 0; q < kPruneHubs / 4; ++q) {
        const uint4 c = p[q];
        best = fmin(best, (double)r.h[4 * q] + (double)__uint_as_float(c.x));
        best = fmin(best, (double)r.h[4 * q + 1] + (double)__uint_as_float(c.y));
        best = fmin(best, (double)r.h[4 * q + 2] + (double)__uint_as_float(c.z));
        best = fmin(best, (double)r.h[4 * q + 3] + (double)__uint_as_float(c.w));
    }
    return best + kPruneMargin < (double)w;
}

// Kept entries per row (deg_s), or, with rs_s set, the kept entries written
// in row order at rs_s[u].
__global__ __launch_bounds__(256) void prune_rows_lane(const u64 *__restrict__ rs, const u32 *__restrict__ col,
                                                       const f32 *__restrict__ w, const u32 *__restrict__ hw,
                                                       u64 first, u64 n, u32 *__restrict__ deg_s,
                                                       const u64 *__restrict__ rs_s, u32 *__restrict__ col_s,
                                                       f32 *__restrict__ w_s) {
    for (u64 u = first + (u64)blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64)gridDim.x * blockDim.x) {
        const u64 b = rs[u], e = rs[u + 1];
        HubRow r;
        load_hub_row(hw, u, r);
        u64 o = rs_s ? rs_s[u] : 0;
        u32 kept = 0;
        for (u64 k = b; k < e; ++k) {
            const u32 v = col[k];
            const f32 wk = w[k];
            if (prune_entry(r, hw, v, wk)) continue;
            if (rs_s) {
                col_s[o] = v;
                w_s[o] = wk;
                ++o;
            }
            ++kept;
        }
        if (!rs_s) deg_s[u] = kept;
    }
}

__global__ __launch_bounds__(256) void prune_rows_block(const u64 *__restrict__ rs, const u32 *__restrict__ col,
                                                        const f32 *__restrict__ w, const u32 *__restrict__ hw,
                                                        u64 count, u32 *__restrict__ deg_s,
                                                        const u64 *__restrict__ rs_s, u32 *__restrict__ col_s,
                                                        f32 *__restrict__ w_s) {
    __shared__ u32 wave_kept[256 / kWave];
    const u32 wv = threadIdx.x / kWave, nw = blockDim.x / kWave;
    for (u64 u = blockIdx.x; u < count; u += gridDim.x) {
        const u64 b = rs[u], e = rs[u + 1];
        HubRow r;
        load_hub_row(hw, u, r);
        u64 o = rs_s ? rs_s[u] : 0;
        for (u64 base = b; base < e; base += blockDim.x) {
            const u64 k = base + threadIdx.x;
            bool keep = false;
            u32 v = 0;
            f32 wk = 0.0f;
            if (k < e) {
                v = col[k];
                wk = w[k];
                keep = !prune_entry(r, hw, v, wk);
            }
            const u64 m = ballot(keep);
            if (lane_id() == 0) wave_kept[wv] = (u32)__popcll(m);
            __syncthreads();
            u32 before = 0, total = 0;
            for (u32 i = 0; i < nw; ++i) {
                before += i < wv ? wave_kept[i] : 0u;
                total += wave_kept[i];
            }
            if (keep && rs_s) {
                const u64 at = o + before + mask_rank(m);
                col_s[at] = v;
                w_s[at] = wk;
            }
            o += total;
            __syncthreads();  // wave_kept is rewritten by the next chunk
        }
        if (!rs_s && threadIdx.x == 0) deg_s[u] = (u32)(o - (rs_s ? rs_s[u] : 0));
    }
}

// Builds g's SSSP copy from its traversal copy (rows longer than kWave
// entries, the first nlong, go to a block each; the first ntop rows are the
// degree class above the exact range, in no particular order). Two passes,
// each with its own hub table: kept counts and row offsets, then, with the
// exact sizes allocated, the kept entries. Scratch is released after each pass.
void build_sssp_copy(Worker_State *ws, Oned_Graph *g, u64 nlong, u64 ntop) {
    Context &c = ctx();
    const hipStream_t s = c.stream;
    const u64 n = g->nlocalverts;
    // The copy is an optimization: without arena room for its largest moment
    // (row offsets, entries at full size, hub table, degrees, alignment slack)
    // the graph keeps none, and SSSP relaxes the traversal copy.
    const u64 need = (n + 1) * 8 + g->nlocaledges * 8 + n * kPruneHubs * 4 + n * 8 + (8u << 20);
    if (ws->gpu_arena.reserved - ws->gpu_arena.used < need) {
        fprintf(stderr, "libg500: no arena room for the SSSP copy (%llu bytes); SSSP relaxes the traversal copy\n",
                (unsigned long long)need);
        return;
    }
    // the hubs: the kPruneHubs largest degrees among the first max(ntop, kPruneHubs)
    // rows (every later row is no longer), ties to the smaller id
    const u64 nhead = MIN(n, MAX(ntop, (u64)kPruneHubs));
    std::vector<u64> head(nhead + 1);
    G500_HIP(hipMemcpyAsync(head.data(), g->rowstarts_t, (nhead + 1) * sizeof(u64), hipMemcpyDeviceToHost, s));
    sync();
    if (failed()) return;
    std::vector<u32> ids(nhead);
    for (u64 i = 0; i < nhead; ++i) ids[i] = (u32)i;
    const u64 nh = MIN(nhead, (u64)kPruneHubs);
    std::partial_sort(ids.begin(), ids.begin() + nh, ids.end(), [&](u32 x, u32 y) {
        const u64 dx = head[x + 1] - head[x], dy = head[y + 1] - head[y];
        return dx != dy ? dx > dy : x < y;
    });
    u32 nslot = 0;
    for (u64 i = 0; i < nh; ++i) nslot = MAX(nslot, ids[i] + 1);
    std::vector<u8> slot_h(MAX(nslot, 1u), (u8)kPruneHubs);
    for (u64 i = 0; i < nh; ++i) slot_h[ids[i]] = (u8)i;
    u64 *rs_s = (u64 *)arena_alloc(ws, (n + 1) * sizeof(u64), "SSSP row offsets");
    if (failed()) return;
    const unsigned lane_grid = grid_for(n - nlong, 256, (u64)c.num_cus * 32);
    const unsigned block_grid = (unsigned)MIN(MAX(nlong, (u64)1), (u64)c.num_cus * 8);
    // hub table of the pass (slot map and weights, in the pass's scratch)
    auto hub_table = [&](u32 *hw, u8 *slot) {
        G500_HIP(hipMemcpyAsync(slot, slot_h.data(), slot_h.size(), hipMemcpyHostToDevice, s));
        G500_HIP(hipMemsetD32Async(hw, 0x7F800000u, n * kPruneHubs, s));
        if (nlong)
            prune_hubs_block<<<block_grid, 256, 0, s>>>(g->rowstarts_t, g->column_t, g->weights_t, slot, nslot, nlong,
                                                        hw);
        prune_hubs_lane<<<lane_grid, 256, 0, s>>>(g->rowstarts_t, g->column_t, g->weights_t, slot, nslot, nlong, n,
                                                  hw);
        G500_HIP(hipGetLastError());
        sync();  // slot_h is read by the copy
    };
    u64 mark = ws->gpu_arena.used;
    u32 *hw = (u32 *)arena_alloc(ws, n * kPruneHubs * sizeof(u32), "SSSP prune hub table");
    u8 *slot = (u8 *)arena_alloc(ws, slot_h.size(), "SSSP hub slots");
    u32 *deg_s = (u32 *)arena_alloc(ws, n * sizeof(u32), "SSSP degrees");
    const u64 ntiles = (n + kScanTile - 1) / kScanTile;
    u64 *tiles = (u64 *)arena_alloc(ws, (ntiles + 1) * sizeof(u64), "SSSP scan tiles");
    if (failed()) return;
    hub_table(hw, slot);
    if (nlong)
        prune_rows_block<<<block_grid, 256, 0, s>>>(g->rowstarts_t, g->column_t, g->weights_t, hw, nlong, deg_s,
                                                    nullptr, nullptr, nullptr);
    prune_rows_lane<<<lane_grid, 256, 0, s>>>(g->rowstarts_t, g->column_t, g->weights_t, hw, nlong, n, deg_s,
                                              nullptr, nullptr, nullptr);
    scan_tile_sums<<<(unsigned)ntiles, kScanBlock, 0, s>>>(deg_s, n, tiles);
    scan_tile_offsets<<<1, 1024, 0, s>>>(tiles, ntiles, rs_s, n);
    scan_write<<<(unsigned)ntiles, kScanBlock, 0, s>>>(deg_s, n, tiles, rs_s);
    G500_HIP(hipGetLastError());
    u64 m_s = 0;
    G500_HIP(hipMemcpyAsync(&m_s, rs_s + n, sizeof(u64), hipMemcpyDeviceToHost, s));
    sync();
    ws->gpu_arena.used = mark;
    if (failed()) return;
    u32 *col_s = (u32 *)arena_alloc(ws, MAX(m_s, (u64)1) * sizeof(u32), "SSSP columns");
    f32 *w_s = (f32 *)arena_alloc(ws, MAX(m_s, (u64)1) * sizeof(f32), "SSSP weights");
    mark = ws->gpu_arena.used;
    hw = (u32 *)arena_alloc(ws, n * kPruneHubs * sizeof(u32), "SSSP prune hub table");
    slot = (u8 *)arena_alloc(ws, slot_h.size(), "SSSP hub slots");
    if (failed()) return;
    hub_table(hw, slot);
    if (nlong)
        prune_rows_block<<<block_grid, 256, 0, s>>>(g->rowstarts_t, g->column_t, g->weights_t, hw, nlong, nullptr,
                                                    rs_s, col_s, w_s);
    prune_rows_lane<<<lane_grid, 256, 0, s>>>(g->rowstarts_t, g->column_t, g->weights_t, hw, nlong, n, nullptr,
                                              rs_s, col_s, w_s);
    G500_HIP(hipGetLastError());
    sync();
    ws->gpu_arena.used = mark;
    if (failed()) return;
    g->rowstarts_s = rs_s;
    g->column_s = col_s;
    g->weights_s = w_s;
}

// Builds g's traversal copy (perm, iperm, rowstarts_t, column_t, weights_t)
// on the library stream; scratch is taken above the arena mark and released.
void build_traversal_copy(Worker_State *ws, Oned_Graph *g) {
    Context &c = ctx();
    const hipStream_t s = c.stream;
    const u64 n = g->nlocalverts, m = g->nlocaledges;
    g->perm = (u32 *)arena_alloc(ws, n * sizeof(u32), "traversal ids");
    g->iperm = (u32 *)arena_alloc(ws, n * sizeof(u32), "original ids");
    g->rowstarts_t = (u64 *)arena_alloc(ws, (n + 1) * sizeof(u64), "traversal row offsets");
    g->column_t = (u32 *)arena_alloc(ws, m * sizeof(u32), "traversal columns");
    g->weights_t = (f32 *)arena_alloc(ws, m * sizeof(f32), "traversal weights");
    const u64 mark = ws->gpu_arena.used;
    u32 *counts = (u32 *)arena_alloc(ws, 2 * kDegKeys * sizeof(u32), "relabel counts");
    u32 *deg_t = (u32 *)arena_alloc(ws, n * sizeof(u32), "relabel degrees");
    const u64 ntiles = (n + kScanTile - 1) / kScanTile;
    u64 *tiles = (u64 *)arena_alloc(ws, (ntiles + 1) * sizeof(u64), "relabel scan tiles");
    if (failed()) return;
    u32 *cursor = counts + kDegKeys;
    const unsigned nblk = (unsigned)((n + (u64)kRelabelBlock * kRelabelItems - 1) / ((u64)kRelabelBlock * kRelabelItems));
    G500_HIP(hipMemsetAsync(counts, 0, kDegKeys * sizeof(u32), s));
    relabel_count<<<nblk, kRelabelBlock, 0, s>>>(g->rowstarts, n, counts);
    relabel_offsets<<<1, 1024, 0, s>>>(counts, cursor);
    relabel_scatter<<<nblk, kRelabelBlock, 0, s>>>(g->rowstarts, n, cursor, g->perm, g->iperm, deg_t);
    scan_tile_sums<<<(unsigned)ntiles, kScanBlock, 0, s>>>(deg_t, n, tiles);
    scan_tile_offsets<<<1, 1024, 0, s>>>(tiles, ntiles, g->rowstarts_t, n);
    scan_write<<<(unsigned)ntiles, kScanBlock, 0, s>>>(deg_t, n, tiles, g->rowstarts_t);
    G500_HIP(hipGetLastError());
    // rows longer than kWave entries: keys below kDegKeys - 1 - kWave
    u32 hc[kDegKeys];
    G500_HIP(hipMemcpyAsync(hc, counts, sizeof(hc), hipMemcpyDeviceToHost, s));
    sync();
    u64 nlong = 0;
    for (u32 k = 0; k < kDegKeys - 1 - (u32)kWave; ++k) nlong += hc[k];
    if (nlong)
        relabel_rows_block<<<(unsigned)MIN(nlong, (u64)c.num_cus * 8), 256, 0, s>>>(
            g->rowstarts, g->column, g->weights, g->perm, g->iperm, g->rowstarts_t, nlong, g->column_t,
            g->weights_t);
    relabel_rows_lane<<<grid_for(n - nlong, 256, (u64)c.num_cus * 32), 256, 0, s>>>(
        g->rowstarts, g->column, g->weights, g->perm, g->iperm, g->rowstarts_t, nlong, n, g->column_t,
        g->weights_t);
    G500_HIP(hipGetLastError());
    sync();
    ws->gpu_arena.used = mark;
    if (c.sssp_prune && !failed()) build_sssp_copy(ws, g, nlong, hc[0]);
}

}  // namespace g500

using namespace g500;

extern "C" {

Oned_Graph oned_graph_from_tuple_graph(Worker_State *ws, Tuple_Graph *tg, u64 nglobalverts) {
    Context &c = ctx();
    Oned_Graph g;
    memset(&g, 0, sizeof(g));
    const u64 n = nglobalverts;
    g.nglobalverts = n;
    g.nlocalverts = n;
    if (n == 0 || n > (1ull << 32)) {
        fail("oned_graph_from_tuple_graph: %llu vertices outside [1, 2^32]", (unsigned long long)n);
        return g;
    }
    const u64 ne = tg->edges_size;
    const u64 mark = ws->gpu_arena.used;  // scratch above this mark is released at the end
    g.rowstarts = (u64 *)arena_alloc(ws, (n + 1) * sizeof(u64), "CSR row offsets");
    if (failed()) return g;
    // Scratch: degrees/cursors, scan tile sums, sort lists.
    const u64 mark_scratch = ws->gpu_arena.used;
    u32 *deg = (u32 *)arena_alloc(ws, n * sizeof(u32), "CSR degrees");
    const u64 ntiles = (n + kScanTile - 1) / kScanTile;
    u64 *tiles = (u64 *)arena_alloc(ws, (ntiles + 1) * sizeof(u64), "CSR scan tiles");
    if (failed()) return g;
    u32 *bad = (u32 *)(tiles + ntiles);
    timer_start();
    const hipStream_t s = c.stream;
    G500_HIP(hipMemsetAsync(deg, 0, n * sizeof(u32), s));
    G500_HIP(hipMemsetAsync(bad, 0, sizeof(u32), s));
    if (ne) degree_kernel<<<grid_for(ne, 256, (u64)c.num_cus * 64), 256, 0, s>>>(tg->edges, ne, n, deg, bad);
    scan_tile_sums<<<(unsigned)ntiles, kScanBlock, 0, s>>>(deg, n, tiles);
    scan_tile_offsets<<<1, 1024, 0, s>>>(tiles, ntiles, g.rowstarts, n);
    scan_write<<<(unsigned)ntiles, kScanBlock, 0, s>>>(deg, n, tiles, g.rowstarts);
    G500_HIP(hipGetLastError());
    u64 m = 0;
    u32 hbad = 0;
    G500_HIP(hipMemcpyAsync(&m, g.rowstarts + n, sizeof(u64), hipMemcpyDeviceToHost, s));
    G500_HIP(hipMemcpyAsync(&hbad, bad, sizeof(u32), hipMemcpyDeviceToHost, s));
    sync();
    if (hbad) fail("oned_graph_from_tuple_graph: tuple endpoint outside [0, %llu)", (unsigned long long)n);
    if (failed()) return g;
    // Column / weight arrays are allocated below the scratch so that the
    // scratch can be released afterwards: move the scratch region up.
    ws->gpu_arena.used = mark_scratch;
    g.nlocaledges = m;
    g.column = (u32 *)arena_alloc(ws, m * sizeof(u32), "CSR columns");
    g.weights = (f32 *)arena_alloc(ws, m * sizeof(f32), "CSR weights");
    const u64 mark_sort = ws->gpu_arena.used;
    u32 *cursor = (u32 *)arena_alloc(ws, n * sizeof(u32), "CSR cursors");
    SortLists *L = (SortLists *)arena_alloc(ws, sizeof(SortLists), "sort lists");
    u32 *t1 = (u32 *)arena_alloc(ws, n * sizeof(u32), "sort tier 1");
    u32 *t2 = (u32 *)arena_alloc(ws, n * sizeof(u32), "sort tier 2");
    u32 *t3 = (u32 *)arena_alloc(ws, n * sizeof(u32), "sort tier 3");
    u64 *t3_off = (u64 *)arena_alloc(ws, n * sizeof(u64), "sort tier 3 offsets");
    if (failed()) return g;
    G500_HIP(hipMemsetAsync(cursor, 0, n * sizeof(u32), s));
    G500_HIP(hipMemsetAsync(L, 0, sizeof(SortLists), s));
    if (ne)
        fill_kernel<<<grid_for(ne, 256, (u64)c.num_cus * 64), 256, 0, s>>>(tg->edges, tg->weights, ne, n,
                                                                          g.rowstarts, cursor, g.column,
                                                                          g.weights);
    classify_rows<<<grid_for(n, 256, (u64)c.num_cus * 16), 256, 0, s>>>(g.rowstarts, n, L, t1, t2, t3, t3_off);
    sort_rows_thread<<<grid_for(n, 256, (u64)c.num_cus * 32), 256, 0, s>>>(g.rowstarts, n, g.column, g.weights);
    G500_HIP(hipGetLastError());
    SortLists hl;
    G500_HIP(hipMemcpyAsync(&hl, L, sizeof(hl), hipMemcpyDeviceToHost, s));
    sync();
    if (failed()) return g;
    g.notisolated = hl.notisolated;
    if (hl.n1)
        sort_rows_wave<<<grid_for((u64)hl.n1 * kWave, 256, (u64)c.num_cus * 32), 256, 0, s>>>(
            g.rowstarts, t1, L, g.column, g.weights);
    if (hl.n2)
        sort_rows_block<<<(unsigned)MIN((u64)hl.n2, (u64)c.num_cus * 16), 256, 0, s>>>(g.rowstarts, t2, L,
                                                                                       g.column, g.weights);
    if (hl.n3) {
        u32 *s_lo = (u32 *)arena_alloc(ws, hl.scratch_total * sizeof(u32), "radix scratch");
        u32 *s_hi = (u32 *)arena_alloc(ws, hl.scratch_total * sizeof(u32), "radix scratch");
        if (failed()) return g;
        int bits = 1;
        while (bits < 32 && (1ull << bits) < n) ++bits;
        const int col_passes = (bits + 7) / 8;
        sort_rows_radix<<<(unsigned)MIN((u64)hl.n3, (u64)c.num_cus * 2), kRadixBlock, 0, s>>>(
            g.rowstarts, t3, t3_off, L, g.column, (u32 *)g.weights, s_lo, s_hi, col_passes);
    }
    G500_HIP(hipGetLastError());
    ws->gpu_arena.used = mark_sort;  // release the scratch
    if (c.relabel) build_traversal_copy(ws, &g);
    timer_stop();
    (void)mark;
    return g;
}

b32 oned_graph_is_vertex_isolated(Worker_State *ws, const Oned_Graph *g, u64 global_vertex) {
    (void)ws;
    if (global_vertex >= g->nglobalverts) {
        fail("Vertex %llu is outside graph vertex range [0, %llu)", (unsigned long long)global_vertex,
             (unsigned long long)g->nglobalverts);
        return 1;
    }
    u64 r[2];
    g500_copy_to_host(r, g->rowstarts + global_vertex, sizeof(r));
    return r[0] == r[1];
}

void oned_graph_free(Oned_Graph *g) {
    if (g) memset(g, 0, sizeof(*g));
}

}  // extern "C"
