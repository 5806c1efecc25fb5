/*
 * CPU oracle for the Graph500 path (test infrastructure; see oracle.h).
 *
 * How each part is pinned (checked by tests/test_oracle*.py, not by
 * retyping the formulas):
 *   - MRG step: Python big-integer recurrence x_n = (107374182 x_{n-1} +
 *     104480 x_{n-5}) mod (2^31-1) (a different arithmetic route).
 *   - Skip powers: Python square-and-multiply of the 5x5 companion matrix for
 *     random exponents; the 9-word table: brute-force single steps for
 *     T[0][k], composition T[b+1][1] = T[b][1]^256, and the SHA-256 digest of
 *     the reference's own table (tests/golden/mrg_skip_table.json).
 *   - make_mrg_seed(2,3) = {3,1,4,1,1} (closed form).
 *   - Generator: every vertex < 2^SCALE; the scramble is a bijection (brute
 *     force, SCALE <= 16); unscrambled endpoints satisfy src <= tgt; level-0
 *     quadrant frequencies 0.57/0.38/0.05; weights are RN_f32(u) * 2^-31.
 *     The scramble values (val0, val1), the scramble itself and whole edges
 *     are restated in Python integers (matrix-power skip, string bit
 *     reversal, products mod 2^64) and compared value by value, with
 *     swapped, dropped or altered constants shown to change the result.
 *   - Roots: the full 64-root loop restated in Python over the stepped MRG
 *     stream (no skip matrices), with degrees from numpy.
 *   - CSR: rowstarts = degree prefix sum; per-row multisets equal the tuple
 *     list's (numpy), no self-loops, symmetric.
 *   - BFS: depths equal scipy.sparse.csgraph unweighted shortest paths.
 *   - Validation: planted mistakes (in the oracle inputs, on the CPU) raise
 *     the error count; valid trees give 0.
 *   - Dijkstra f64: equals scipy.sparse.csgraph.dijkstra exactly (unique FP64
 *     fixed point). Dijkstra f32: the certificate (o_sssp_check, independently
 *     vectorised in numpy in the tests) which fixes the FP32 fixed point
 *     uniquely; delta-stepping emulation equals Dijkstra bit for bit.
 *   - Depth certificate (o_depth_check): zero exactly on the BFS depth array,
 *     checked by brute force over perturbed depth arrays of small graphs.
 *   - Statistics: numpy percentiles / mean / std(ddof=1) on odd and even n.
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define P31 UINT64_C(2147483647)
#define MRG_A1 UINT64_C(107374182) /* multiplier of x_{n-1} */
#define MRG_A5 UINT64_C(104480)    /* multiplier of x_{n-5} */

/* ------------------------------------------------------------------------ */
/* MRG                                                                      */
/* ------------------------------------------------------------------------ */

/* graph_generation.cu:73-85 (cuda_mrg_orig_step / cuda_mrg_get_uint_orig). */
uint32_t o_mrg_next(o_mrg_state *st) {
    uint64_t nz = (MRG_A1 * st->z[0] + MRG_A5 * st->z[4]) % P31;
    st->z[4] = st->z[3];
    st->z[3] = st->z[2];
    st->z[2] = st->z[1];
    st->z[1] = st->z[0];
    st->z[0] = (uint32_t)nz;
    return (uint32_t)nz;
}

typedef struct { uint32_t a[5][5]; } mat5;

static void mat_mul(const mat5 *x, const mat5 *y, mat5 *r) {
    mat5 t;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            uint64_t s = 0;
            for (int k = 0; k < 5; ++k) s += ((uint64_t)x->a[i][k] * y->a[k][j]) % P31;
            t.a[i][j] = (uint32_t)(s % P31);
        }
    *r = t;
}

static void mat_apply(const mat5 *m, o_mrg_state *st) {
    uint32_t out[5];
    for (int i = 0; i < 5; ++i) {
        uint64_t s = 0;
        for (int k = 0; k < 5; ++k) s += ((uint64_t)m->a[i][k] * st->z[k]) % P31;
        out[i] = (uint32_t)(s % P31);
    }
    memcpy(st->z, out, sizeof(out));
}

/* POW[b][v] = M^(v * 256^b), M the one-step companion matrix. */
static mat5 *POW;
static pthread_once_t pow_once = PTHREAD_ONCE_INIT;

static void build_powers(void) {
    POW = (mat5 *)calloc(24 * 256, sizeof(mat5));
    mat5 m1;
    memset(&m1, 0, sizeof(m1));
    m1.a[0][0] = (uint32_t)MRG_A1;
    m1.a[0][4] = (uint32_t)MRG_A5;
    for (int i = 1; i < 5; ++i) m1.a[i][i - 1] = 1;
    for (int b = 0; b < 24; ++b) {
        mat5 *row = POW + b * 256;
        memset(&row[0], 0, sizeof(mat5));
        for (int i = 0; i < 5; ++i) row[0].a[i][i] = 1;
        if (b == 0)
            row[1] = m1;
        else /* M^(256^b) = M^(255 * 256^(b-1)) * M^(256^(b-1)) */
            mat_mul(&POW[(b - 1) * 256 + 255], &POW[(b - 1) * 256 + 1], &row[1]);
        for (int v = 2; v < 256; ++v) mat_mul(&row[v - 1], &row[1], &row[v]);
    }
}

static const mat5 *mrg_pow(int b, int v) {
    pthread_once(&pow_once, build_powers);
    return &POW[b * 256 + v];
}

void o_mrg_power(int b, int v, uint32_t out[25]) {
    memcpy(out, mrg_pow(b, v)->a, 25 * sizeof(uint32_t));
}

/* graph_generation.cu:209-223: one matrix per nonzero byte of the exponent. */
void o_mrg_skip(o_mrg_state *st, uint64_t hi, uint64_t mid, uint64_t lo) {
    uint64_t parts[3] = {lo, mid, hi};
    for (int p = 0; p < 3; ++p)
        for (int k = 0; k < 8; ++k) {
            int byte = (int)((parts[p] >> (8 * k)) & 0xFF);
            if (byte) mat_apply(mrg_pow(8 * p + k, byte), st);
        }
}

/* The compressed 9-word form {s,t,u,v,w,a,b,c,d} of a power of M (the layout
 * of mrg_transitions.cu:14 consumed by graph_generation.cu:56-67): row 5 is
 * (s,t,u,v,w) and the first column is (d,c,b,a,s). */
void o_mrg_skip_table(uint32_t *out) {
    for (int b = 0; b < 24; ++b)
        for (int v = 0; v < 256; ++v) {
            const mat5 *m = mrg_pow(b, v);
            uint32_t *o = out + ((size_t)b * 256 + v) * 9;
            o[0] = m->a[4][0]; o[1] = m->a[4][1]; o[2] = m->a[4][2];
            o[3] = m->a[4][3]; o[4] = m->a[4][4];
            o[5] = m->a[3][0]; o[6] = m->a[2][0]; o[7] = m->a[1][0]; o[8] = m->a[0][0];
        }
}

/* graph_generation.cu:248-254 (Graph500 user seed -> MRG seed). */
void o_make_mrg_seed(uint64_t s1, uint64_t s2, uint32_t seed[5]) {
    seed[0] = (uint32_t)(s1 & 0x3FFFFFFFu) + 1;
    seed[1] = (uint32_t)((s1 >> 30) & 0x3FFFFFFFu) + 1;
    seed[2] = (uint32_t)(s2 & 0x3FFFFFFFu) + 1;
    seed[3] = (uint32_t)((s2 >> 30) & 0x3FFFFFFFu) + 1;
    seed[4] = (uint32_t)((s2 >> 60) << 4) + (uint32_t)(s1 >> 60) + 1;
}

/* 1/(2^31-1) rounded to double: the literal of graph_generation.cu:93. */
static const double INV_P_F64 = 0x1.00000002p-31;

/* graph_generation.cu:256-266. */
void o_make_random_numbers(int64_t nvalues, uint64_t s1, uint64_t s2, int64_t pos, double *out) {
    uint32_t seed[5];
    o_make_mrg_seed(s1, s2, seed);
    o_mrg_state st;
    memcpy(st.z, seed, sizeof(seed));
    o_mrg_skip(&st, 2, 0, 2 * (uint64_t)pos);
    for (int64_t i = 0; i < nvalues; ++i) out[i] = (double)o_mrg_next(&st) * INV_P_F64;
}

/* graph_generation.cu:88-90: RN_f32(u) times 2^-31 (the f32 literal rounds
 * to exactly 2^-31). */
float o_mrg_weight(o_mrg_state *st) {
    uint32_t u = o_mrg_next(st);
    return (float)u * 0x1p-31f;
}

/* ------------------------------------------------------------------------ */
/* Kronecker generator                                                      */
/* ------------------------------------------------------------------------ */

/* graph_generation.cu:290-300 (val0, val1 of generate_kronecker_range). */
void o_scramble_values(const uint32_t seed[5], uint64_t *val0, uint64_t *val1) {
    o_mrg_state st;
    memcpy(st.z, seed, 5 * sizeof(uint32_t));
    o_mrg_skip(&st, 50, 7, 0);
    uint64_t a = o_mrg_next(&st);
    a = a * UINT64_C(0xFFFFFFFF) + o_mrg_next(&st);
    uint64_t b = o_mrg_next(&st);
    b = b * UINT64_C(0xFFFFFFFF) + o_mrg_next(&st);
    *val0 = a;
    *val1 = b;
}

static uint64_t brev64(uint64_t x) {
    uint64_t r = 0;
    for (int i = 0; i < 64; ++i) r |= ((x >> i) & 1u) << (63 - i);
    return r;
}

/* graph_generation.cu:171-179 (cuda_scramble; bit reversal of :145-169). */
uint64_t o_scramble(uint64_t v, int lgN, uint64_t val0, uint64_t val1) {
    v += val0 + val1;
    v *= (val0 | UINT64_C(0x4519840211493211));
    v = brev64(v) >> (64 - lgN);
    v *= (val1 | UINT64_C(0x3050852102C843A5));
    v = brev64(v) >> (64 - lgN);
    return v;
}

/* graph_generation.cu:100-143 with SPK_NOISE_LEVEL 0: reject u below
 * 0x7FFFFFFF mod 10000 = 3647, then u mod 10000 in [0,1900) -> 1 (B),
 * [1900,3800) -> 2 (C), [3800,9500) -> 0 (A), else 3 (D). */
static int quadrant(o_mrg_state *st) {
    uint32_t u;
    do u = o_mrg_next(st); while (u < 3647u);
    u %= 10000u;
    if (u < 1900u) return 1;
    if (u < 3800u) return 2;
    if (u < 9500u) return 0;
    return 3;
}

/* graph_generation.cu:186-207 plus the weight draw of :236. */
static void one_edge(o_mrg_state *st, int lgN, uint64_t val0, uint64_t val1, uint64_t *e, float *w) {
    uint64_t half = (uint64_t)1 << lgN, src = 0, tgt = 0;
    while (half > 1) {
        int sq = quadrant(st);
        int so = sq >> 1, to = sq & 1;
        if (src == tgt && so > to) { int t = so; so = to; to = t; }
        half >>= 1;
        src += half * (uint64_t)so;
        tgt += half * (uint64_t)to;
    }
    e[0] = o_scramble(src, lgN, val0, val1);
    e[1] = o_scramble(tgt, lgN, val0, val1);
    *w = o_mrg_weight(st);
}

/* graph_generation.cu:225-238, :269-310. The state of edge ei is the seed
 * skipped by ei * 2^64; consecutive edges share the skip over the upper bytes
 * of ei (matrix products of powers of M commute), so the state for the
 * 256-aligned block is computed once and the low byte applied per edge. */
void o_generate_kronecker_range(const uint32_t seed[5], int logN, int64_t start, int64_t end,
                                uint64_t *edges, float *weights) {
    uint64_t val0, val1;
    o_scramble_values(seed, &val0, &val1);
    mrg_pow(0, 0);
    int64_t first_blk = start >> 8, last_blk = (end + 255) >> 8;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t blk = first_blk; blk < last_blk; ++blk) {
        o_mrg_state base;
        memcpy(base.z, seed, 5 * sizeof(uint32_t));
        o_mrg_skip(&base, 0, (uint64_t)blk << 8, 0);
        int64_t lo = blk << 8, hi = lo + 256;
        if (lo < start) lo = start;
        if (hi > end) hi = end;
        for (int64_t ei = lo; ei < hi; ++ei) {
            o_mrg_state st = base;
            int byte = (int)(ei & 0xFF);
            if (byte) mat_apply(mrg_pow(8, byte), &st);
            one_edge(&st, logN, val0, val1, edges + 2 * (ei - start), weights + (ei - start));
        }
    }
}

/* ------------------------------------------------------------------------ */
/* CSR                                                                      */
/* ------------------------------------------------------------------------ */

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

static uint32_t fbits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

static float bitsf(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* traversal.cu:15-73, :342-389: self-loops skipped, both directions of every
 * other tuple edge appended with its weight, duplicates kept. Rows are put in
 * the canonical ascending (weight bits, column) order. */
uint64_t o_csr_build(uint64_t n, uint64_t nedges, const uint64_t *edges, const float *weights,
                     uint64_t *rowstarts, uint32_t *column, float *cw) {
    memset(rowstarts, 0, (n + 1) * sizeof(uint64_t));
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nedges; ++i) {
        uint64_t a = edges[2 * i], b = edges[2 * i + 1];
        if (a == b) continue;
        __atomic_fetch_add(&rowstarts[a + 1], 1, __ATOMIC_RELAXED);
        __atomic_fetch_add(&rowstarts[b + 1], 1, __ATOMIC_RELAXED);
    }
    for (uint64_t v = 0; v < n; ++v) rowstarts[v + 1] += rowstarts[v];
    uint64_t m = rowstarts[n];
    if (!column) return m;
    uint64_t *keys = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
    uint64_t *cur = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    memcpy(cur, rowstarts, n * sizeof(uint64_t));
    for (uint64_t i = 0; i < nedges; ++i) {
        uint64_t a = edges[2 * i], b = edges[2 * i + 1];
        if (a == b) continue;
        uint64_t wb = (uint64_t)fbits(weights[i]) << 32;
        keys[cur[a]++] = wb | b;
        keys[cur[b]++] = wb | a;
    }
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint64_t v = 0; v < n; ++v)
        qsort(keys + rowstarts[v], rowstarts[v + 1] - rowstarts[v], sizeof(uint64_t), cmp_u64);
    for (uint64_t k = 0; k < m; ++k) {
        column[k] = (uint32_t)keys[k];
        cw[k] = bitsf((uint32_t)(keys[k] >> 32));
    }
    free(cur);
    free(keys);
    return m;
}

/* ------------------------------------------------------------------------ */
/* BFS                                                                      */
/* ------------------------------------------------------------------------ */

uint64_t o_bfs_serial(uint64_t n, const uint64_t *rs, const uint32_t *col, uint64_t root,
                      int64_t *pred, int64_t *depth) {
    uint32_t *q = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    for (uint64_t v = 0; v < n; ++v) pred[v] = -1;
    if (depth)
        for (uint64_t v = 0; v < n; ++v) depth[v] = -1;
    uint64_t head = 0, tail = 0;
    pred[root] = (int64_t)root;
    if (depth) depth[root] = 0;
    q[tail++] = (uint32_t)root;
    while (head < tail) {
        uint32_t u = q[head++];
        for (uint64_t k = rs[u]; k < rs[u + 1]; ++k) {
            uint32_t v = col[k];
            if (pred[v] != -1) continue;
            pred[v] = u;
            if (depth) depth[v] = depth[u] + 1;
            q[tail++] = v;
        }
    }
    free(q);
    return tail;
}

uint64_t o_bfs_omp(uint64_t n, const uint64_t *rs, const uint32_t *col, uint64_t root,
                   int64_t *pred, int nthreads) {
    uint32_t *cur = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *nxt = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
#pragma omp parallel for num_threads(nthreads)
    for (uint64_t v = 0; v < n; ++v) pred[v] = -1;
    pred[root] = (int64_t)root;
    cur[0] = (uint32_t)root;
    uint64_t ncur = 1, reached = 1;
    while (ncur) {
        uint64_t nnext = 0;
#pragma omp parallel num_threads(nthreads)
        {
            uint32_t local[256];
            int nl = 0;
#pragma omp for schedule(dynamic, 64)
            for (uint64_t i = 0; i < ncur; ++i) {
                uint32_t u = cur[i];
                for (uint64_t k = rs[u]; k < rs[u + 1]; ++k) {
                    uint32_t v = col[k];
                    if (__atomic_load_n(&pred[v], __ATOMIC_RELAXED) != -1) continue;
                    int64_t expect = -1;
                    if (!__atomic_compare_exchange_n(&pred[v], &expect, (int64_t)u, 0, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED))
                        continue;
                    local[nl++] = v;
                    if (nl == 256) {
                        uint64_t at = __atomic_fetch_add(&nnext, 256, __ATOMIC_RELAXED);
                        memcpy(nxt + at, local, sizeof(local));
                        nl = 0;
                    }
                }
            }
            if (nl) {
                uint64_t at = __atomic_fetch_add(&nnext, (uint64_t)nl, __ATOMIC_RELAXED);
                memcpy(nxt + at, local, (size_t)nl * sizeof(uint32_t));
            }
        }
        uint32_t *t = cur;
        cur = nxt;
        nxt = t;
        ncur = nnext;
        reached += nnext;
    }
    free(cur);
    free(nxt);
    return reached;
}

/* Depths of a level-synchronous top-down BFS (OpenMP): the textbook hop
 * counts, -1 unreached (the depth convention of traversal.cu:505-563).
 * Returns the reached count. */
uint64_t o_bfs_depth_omp(uint64_t n, const uint64_t *rs, const uint32_t *col, uint64_t root,
                         int64_t *depth, int nthreads) {
    uint32_t *cur = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *nxt = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
#pragma omp parallel for num_threads(nthreads)
    for (uint64_t v = 0; v < n; ++v) depth[v] = -1;
    depth[root] = 0;
    cur[0] = (uint32_t)root;
    uint64_t ncur = 1, reached = 1;
    for (int64_t level = 1; ncur; ++level) {
        uint64_t nnext = 0;
#pragma omp parallel num_threads(nthreads)
        {
            uint32_t local[256];
            int nl = 0;
#pragma omp for schedule(dynamic, 64)
            for (uint64_t i = 0; i < ncur; ++i) {
                uint32_t u = cur[i];
                for (uint64_t k = rs[u]; k < rs[u + 1]; ++k) {
                    uint32_t v = col[k];
                    if (__atomic_load_n(&depth[v], __ATOMIC_RELAXED) != -1) continue;
                    int64_t expect = -1;
                    if (!__atomic_compare_exchange_n(&depth[v], &expect, level, 0, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED))
                        continue;
                    local[nl++] = v;
                    if (nl == 256) {
                        uint64_t at = __atomic_fetch_add(&nnext, 256, __ATOMIC_RELAXED);
                        memcpy(nxt + at, local, sizeof(local));
                        nl = 0;
                    }
                }
            }
            if (nl) {
                uint64_t at = __atomic_fetch_add(&nnext, (uint64_t)nl, __ATOMIC_RELAXED);
                memcpy(nxt + at, local, (size_t)nl * sizeof(uint32_t));
            }
        }
        uint32_t *t = cur;
        cur = nxt;
        nxt = t;
        ncur = nnext;
        reached += nnext;
    }
    free(cur);
    free(nxt);
    return reached;
}

/* Order-free digest of a multiset of directed entries (u, v, weight bits):
 * a sum mod 2^64 of a 64-bit mix of each entry. */
static uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= UINT64_C(0xbf58476d1ce4e5b9);
    x ^= x >> 27;
    x *= UINT64_C(0x94d049bb133111eb);
    x ^= x >> 31;
    return x;
}
static uint64_t entry_hash(uint64_t u, uint64_t v, uint32_t wb) {
    return mix64(mix64((u << 32) ^ v) ^ ((uint64_t)wb * UINT64_C(0x9e3779b97f4a7c15)));
}

uint64_t o_tuple_digest(uint64_t nedges, const uint64_t *edges, const float *weights, uint64_t *entries) {
    uint64_t sum = 0, cnt = 0;
#pragma omp parallel for schedule(static) reduction(+ : sum, cnt)
    for (uint64_t i = 0; i < nedges; ++i) {
        uint64_t a = edges[2 * i], b = edges[2 * i + 1];
        if (a == b) continue;
        uint32_t wb;
        memcpy(&wb, &weights[i], 4);
        sum += entry_hash(a, b, wb) + entry_hash(b, a, wb);
        cnt += 2;
    }
    if (entries) *entries = cnt;
    return sum;
}

uint64_t o_csr_digest(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *cw) {
    uint64_t sum = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : sum)
    for (uint64_t u = 0; u < n; ++u)
        for (uint64_t k = rs[u]; k < rs[u + 1]; ++k) {
            uint32_t wb;
            memcpy(&wb, &cw[k], 4);
            sum += entry_hash(u, col[k], wb);
        }
    return sum;
}

/* traversal.cu:15-34 over one chunk of a tuple list: both endpoints of every
 * non-loop edge count once. Endpoints outside [0, n) are counted and skipped. */
uint64_t o_degree_add(uint64_t n, uint64_t nedges, const uint64_t *edges, uint64_t *deg) {
    uint64_t bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (uint64_t i = 0; i < nedges; ++i) {
        uint64_t a = edges[2 * i], b = edges[2 * i + 1];
        if (a >= n || b >= n) { ++bad; continue; }
        if (a == b) continue;
        __atomic_fetch_add(&deg[a], 1, __ATOMIC_RELAXED);
        __atomic_fetch_add(&deg[b], 1, __ATOMIC_RELAXED);
    }
    return bad;
}

/* The canonical row order of o_csr_build: rowstarts start at 0 and never
 * decrease; inside a row, keys (weight bits << 32 | column) ascend. Counts the
 * decreasing rowstarts steps plus the rows with a descending adjacent pair. */
uint64_t o_csr_order_errors(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *cw) {
    uint64_t err = rs[0] != 0;
    for (uint64_t u = 0; u < n; ++u) err += rs[u + 1] < rs[u];
    if (err) return err;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : err)
    for (uint64_t u = 0; u < n; ++u) {
        uint64_t prev = 0;
        for (uint64_t k = rs[u]; k < rs[u + 1]; ++k) {
            uint64_t key = ((uint64_t)fbits(cw[k]) << 32) | col[k];
            if (k > rs[u] && key < prev) { ++err; break; }
            prev = key;
        }
    }
    return err;
}

/* traversal.cu:75-112, :505-563: level-by-level, a vertex v gets level L+1
 * when the entry v appears in the row of u = pred[v] and u is at level L. */
void o_dist_from_pred(uint64_t n, const uint64_t *rs, const uint32_t *col, const int64_t *pred,
                      uint64_t root, int64_t *dist) {
    for (uint64_t v = 0; v < n; ++v) dist[v] = pred[v] == -1 ? -1 : INT64_MAX;
    dist[root] = 0;
    for (int64_t level = 0;; ++level) {
        uint64_t added = 0;
        for (uint64_t u = 0; u < n; ++u) {
            if (dist[u] != level) continue;
            for (uint64_t k = rs[u]; k < rs[u + 1]; ++k) {
                uint32_t v = col[k];
                if (pred[v] != (int64_t)u) continue;
                if (dist[v] == INT64_MAX) {
                    dist[v] = level + 1;
                    ++added;
                }
            }
        }
        if (!added) break;
    }
}

static int64_t wrap_inc(int64_t x) { return (int64_t)((uint64_t)x + 1u); }

/* graph500_runner.cu:84-169, :203-255. */
uint64_t o_bfs_validate(uint64_t n, const uint64_t *rs, const uint32_t *col, const int64_t *pred,
                        uint64_t root) {
    int64_t *dist = (int64_t *)malloc((n ? n : 1) * sizeof(int64_t));
    uint8_t *conf = (uint8_t *)calloc(n ? n : 1, 1);
    o_dist_from_pred(n, rs, col, pred, root, dist);
    uint64_t err = 0;
    for (uint64_t v = 0; v < n; ++v) {
        int64_t p = pred[v];
        if (p != -1 && (p < 0 || (uint64_t)p >= n)) ++err;                     /* rule 1 */
        if (v == root) {                                                       /* rule 2 */
            if (p != (int64_t)root || dist[v] != 0) ++err;
            else conf[v] = 1;
        }
    }
    for (uint64_t s = 0; s < n; ++s) {
        int64_t sp = pred[s], sd = dist[s];
        for (uint64_t k = rs[s]; k < rs[s + 1]; ++k) {
            uint64_t d = col[k];
            if (d == s) { ++err; continue; }                                   /* rule 3 */
            int64_t dp = pred[d], dd = dist[d];
            if ((sd == -1 && dd != -1) || (sd != -1 && dd == -1)) ++err;       /* rule 4 */
            if (sd >= 0 && dd >= 0 && (wrap_inc(sd) < dd || wrap_inc(dd) < sd)) ++err; /* 5 */
            if (sp == (int64_t)d) conf[s] = 1;                                 /* rule 6 */
            if (dp == (int64_t)s) conf[d] = 1;
        }
    }
    for (uint64_t v = 0; v < n; ++v) {                                         /* rule 7 */
        if (pred[v] == -1) continue;
        if (dist[v] == INT64_MAX || (v != root && !conf[v])) ++err;
    }
    free(conf);
    free(dist);
    return err;
}

/* graph500_runner.cu:171-182. */
uint64_t o_edge_count(uint64_t n, const uint64_t *rs, const uint32_t *col, const int64_t *pred) {
    uint64_t c = 0;
    for (uint64_t v = 0; v < n; ++v) {
        if (pred[v] == -1) continue;
        for (uint64_t k = rs[v]; k < rs[v + 1]; ++k) c += col[k] <= v;
    }
    return c;
}

/* graph500_runner.cu:43-74. */
int o_generate_roots(uint64_t n, const uint64_t *rs, uint64_t s1, uint64_t s2, uint64_t *roots,
                     int max_roots) {
    uint64_t counter = 0;
    int count = 0;
    for (; count < max_roots; ++count) {
        int found = 0;
        uint64_t root = 0;
        for (;;) {
            double d[2];
            o_make_random_numbers(2, s1, s2, (int64_t)counter, d);
            root = (uint64_t)((d[0] + d[1]) * (double)n) % n;
            counter += 2;
            if (counter > 2 * n) break;
            int dup = 0;
            for (int i = 0; i < count; ++i)
                if (roots[i] == root) { dup = 1; break; }
            if (dup) continue;
            if (rs[root] == rs[root + 1]) continue;
            found = 1;
            break;
        }
        if (!found) break;
        roots[count] = root;
    }
    return count;
}

/* ------------------------------------------------------------------------ */
/* SSSP                                                                     */
/* ------------------------------------------------------------------------ */

typedef struct { double d; uint32_t v; } hent;

typedef struct { hent *a; uint64_t n, cap; } heap;

static void hpush(heap *h, double d, uint32_t v) {
    if (h->n == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 1024;
        h->a = (hent *)realloc(h->a, h->cap * sizeof(hent));
    }
    uint64_t i = h->n++;
    while (i) {
        uint64_t p = (i - 1) / 2;
        if (h->a[p].d <= d) break;
        h->a[i] = h->a[p];
        i = p;
    }
    h->a[i].d = d;
    h->a[i].v = v;
}

static hent hpop(heap *h) {
    hent top = h->a[0], last = h->a[--h->n];
    uint64_t i = 0;
    for (;;) {
        uint64_t c = 2 * i + 1;
        if (c >= h->n) break;
        if (c + 1 < h->n && h->a[c + 1].d < h->a[c].d) ++c;
        if (h->a[c].d >= last.d) break;
        h->a[i] = h->a[c];
        i = c;
    }
    if (h->n) h->a[i] = last;
    return top;
}

void o_dijkstra_f32(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *w,
                    uint64_t root, float *dist, int64_t *pred) {
    for (uint64_t v = 0; v < n; ++v) { dist[v] = INFINITY; pred[v] = -1; }
    heap h = {0, 0, 0};
    dist[root] = 0.0f;
    pred[root] = (int64_t)root;
    hpush(&h, 0.0, (uint32_t)root);
    while (h.n) {
        hent e = hpop(&h);
        float du = dist[e.v];
        if ((float)e.d > du) continue;
        for (uint64_t k = rs[e.v]; k < rs[e.v + 1]; ++k) {
            volatile float cand = du + w[k]; /* one IEEE binary32 add */
            uint32_t v = col[k];
            if (cand < dist[v]) {
                dist[v] = cand;
                pred[v] = e.v;
                hpush(&h, cand, v);
            }
        }
    }
    free(h.a);
    for (uint64_t v = 0; v < n; ++v)
        if (isinf(dist[v])) dist[v] = -1.0f;
}

static double now_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void o_dijkstra_f32_timed(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *w,
                          uint64_t root, double max_seconds, uint64_t *scanned, uint64_t *settled,
                          double *elapsed) {
    float *dist = (float *)malloc((n ? n : 1) * sizeof(float));
    for (uint64_t v = 0; v < n; ++v) dist[v] = INFINITY;
    heap h = {0, 0, 0};
    const double t0 = now_seconds();
    uint64_t sc = 0, st = 0, pops = 0;
    dist[root] = 0.0f;
    hpush(&h, 0.0, (uint32_t)root);
    while (h.n) {
        hent e = hpop(&h);
        float du = dist[e.v];
        if ((float)e.d > du) continue;
        ++st;
        for (uint64_t k = rs[e.v]; k < rs[e.v + 1]; ++k) {
            volatile float cand = du + w[k];
            uint32_t v = col[k];
            if (cand < dist[v]) {
                dist[v] = cand;
                hpush(&h, cand, v);
            }
        }
        sc += rs[e.v + 1] - rs[e.v];
        if ((++pops & 1023) == 0 && now_seconds() - t0 > max_seconds) break;
    }
    *elapsed = now_seconds() - t0;
    *scanned = sc;
    *settled = st;
    free(h.a);
    free(dist);
}

void o_dijkstra_f64(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *w,
                    uint64_t root, double *dist) {
    for (uint64_t v = 0; v < n; ++v) dist[v] = INFINITY;
    heap h = {0, 0, 0};
    dist[root] = 0.0;
    hpush(&h, 0.0, (uint32_t)root);
    while (h.n) {
        hent e = hpop(&h);
        double du = dist[e.v];
        if (e.d > du) continue;
        for (uint64_t k = rs[e.v]; k < rs[e.v + 1]; ++k) {
            double cand = du + (double)w[k];
            uint32_t v = col[k];
            if (cand < dist[v]) {
                dist[v] = cand;
                hpush(&h, cand, v);
            }
        }
    }
    free(h.a);
    for (uint64_t v = 0; v < n; ++v)
        if (isinf(dist[v])) dist[v] = -1.0;
}

/* traversal.cu:638-711 run sequentially: light iterations over the bucket
 * frontier (edges w < delta; a vertex is re-queued when its new distance is
 * below the bucket's upper bound), then one heavy pass over every vertex in
 * [bmin, bmax) (edges w >= delta), bmax += delta by float accumulation, and a
 * full scan for the next bucket while any finite distance >= bmin remains. */
void o_delta_stepping_f32(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *w,
                          uint64_t root, float delta, float *dist, int64_t *pred) {
    uint32_t *cur = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *nxt = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint8_t *seen = (uint8_t *)calloc(n ? n : 1, 1);
    for (uint64_t v = 0; v < n; ++v) { dist[v] = INFINITY; pred[v] = -1; }
    dist[root] = 0.0f;
    pred[root] = (int64_t)root;
    float bmin = 0.0f, bmax = delta;
    uint64_t ncur = 1, pending = 1;
    cur[0] = (uint32_t)root;
    while (pending) {
        while (ncur) {
            memset(seen, 0, n);
            uint64_t nn = 0;
            for (uint64_t i = 0; i < ncur; ++i) {
                uint32_t u = cur[i];
                float du = dist[u];
                for (uint64_t k = rs[u]; k < rs[u + 1]; ++k) {
                    if (!(w[k] < delta)) continue;
                    volatile float cand = du + w[k];
                    uint32_t v = col[k];
                    if (cand < dist[v]) {
                        dist[v] = cand;
                        pred[v] = u;
                        if (cand < bmax && !seen[v]) { seen[v] = 1; nxt[nn++] = v; }
                    }
                }
            }
            uint32_t *t = cur; cur = nxt; nxt = t;
            ncur = nn;
        }
        for (uint64_t u = 0; u < n; ++u) {
            float du = dist[u];
            if (du < bmin || du >= bmax) continue;
            for (uint64_t k = rs[u]; k < rs[u + 1]; ++k) {
                if (w[k] < delta) continue;
                volatile float cand = du + w[k];
                uint32_t v = col[k];
                if (cand < dist[v]) { dist[v] = cand; pred[v] = u; }
            }
        }
        bmin = bmax;
        bmax += delta;
        ncur = 0;
        pending = 0;
        for (uint64_t v = 0; v < n; ++v) {
            if (isinf(dist[v])) continue;
            if (dist[v] >= bmin) {
                ++pending;
                if (dist[v] < bmax) cur[ncur++] = (uint32_t)v;
            }
        }
    }
    for (uint64_t v = 0; v < n; ++v)
        if (isinf(dist[v])) dist[v] = -1.0f;
    free(seen);
    free(cur);
    free(nxt);
}

static int atomic_min_u32(uint32_t *p, uint32_t v) {
    uint32_t old = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v < old)
        if (__atomic_compare_exchange_n(p, &old, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return 1;
    return 0;
}

/* CPU baseline: parallel delta-stepping. Inside a bucket [lo, fl(lo + delta))
 * light edges (w < delta) are relaxed round by round until no distance in the
 * bucket drops, then the heavy edges of the bucket's vertices are relaxed once
 * (they land at or above the bucket's top); the next bucket starts at the
 * smallest open distance. Distances are u32 bits lowered by an atomic min
 * (non-negative floats order like their bits). A last pass takes for each
 * reached vertex a neighbour whose distance plus the entry's weight gives its
 * distance as the parent, discovering vertices breadth-first from the root so
 * that no parent chain closes a cycle. The FP32 fixed point is unique, so the distances
 * equal Dijkstra's (traversal.cu:638-711 schedules the same relaxations). */
void o_delta_stepping_omp(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *w,
                          uint64_t root, float delta, float *dist, int64_t *pred, int nthreads) {
    const uint32_t INF = 0x7f800000u;
    uint32_t *db = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t)); /* distance bits */
    uint32_t *cur = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *nxt = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *members = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *mark = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));   /* last light round that queued v */
    uint32_t *inb = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));    /* last bucket that listed v */
    uint8_t *settled = (uint8_t *)calloc(n ? n : 1, 1);
#pragma omp parallel for num_threads(nthreads)
    for (uint64_t v = 0; v < n; ++v) db[v] = INF;
    db[root] = 0;
    float lo = 0.0f;
    uint32_t round = 0, bucket = 0;
    for (;;) {
        const float hi = lo + delta;
        uint32_t lo_b, hi_b;
        memcpy(&lo_b, &lo, 4);
        memcpy(&hi_b, &hi, 4);
        ++bucket;
        uint64_t ncur = 0, nmem = 0;
        uint32_t next_min = INF;
#pragma omp parallel num_threads(nthreads) reduction(min : next_min)
        {
#pragma omp for schedule(static)
            for (uint64_t v = 0; v < n; ++v) {
                if (settled[v] || db[v] == INF) continue;
                if (db[v] >= lo_b && db[v] < hi_b) {
                    uint64_t at = __atomic_fetch_add(&ncur, 1, __ATOMIC_RELAXED);
                    cur[at] = (uint32_t)v;
                } else if (db[v] < next_min) {
                    next_min = db[v];
                }
            }
        }
        if (!ncur) {
            if (next_min == INF) break;
            memcpy(&lo, &next_min, 4);
            continue;
        }
        while (ncur) {
            ++round;
            uint64_t nn = 0;
#pragma omp parallel num_threads(nthreads)
            {
#pragma omp for schedule(dynamic, 64)
                for (uint64_t i = 0; i < ncur; ++i) {
                    uint32_t u = cur[i];
                    if (__atomic_exchange_n(&inb[u], bucket, __ATOMIC_RELAXED) != bucket) {
                        uint64_t at = __atomic_fetch_add(&nmem, 1, __ATOMIC_RELAXED);
                        members[at] = u;
                    }
                    uint32_t dub = __atomic_load_n(&db[u], __ATOMIC_RELAXED);
                    float du;
                    memcpy(&du, &dub, 4);
                    for (uint64_t k = rs[u]; k < rs[u + 1]; ++k) {
                        if (!(w[k] < delta)) continue;
                        volatile float cand = du + w[k];
                        uint32_t cb;
                        float c = cand;
                        memcpy(&cb, &c, 4);
                        uint32_t v = col[k];
                        if (!atomic_min_u32(&db[v], cb) || !(cb < hi_b)) continue;
                        if (__atomic_exchange_n(&mark[v], round, __ATOMIC_RELAXED) == round) continue;
                        uint64_t at = __atomic_fetch_add(&nn, 1, __ATOMIC_RELAXED);
                        nxt[at] = v;
                    }
                }
            }
            uint32_t *t = cur;
            cur = nxt;
            nxt = t;
            ncur = nn;
        }
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
        for (uint64_t i = 0; i < nmem; ++i) {
            uint32_t u = members[i];
            settled[u] = 1;
            float du;
            memcpy(&du, &db[u], 4);
            for (uint64_t k = rs[u]; k < rs[u + 1]; ++k) {
                if (w[k] < delta) continue;
                volatile float cand = du + w[k];
                uint32_t cb;
                float c = cand;
                memcpy(&cb, &c, 4);
                atomic_min_u32(&db[col[k]], cb);
            }
        }
        lo = hi;
    }
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (uint64_t v = 0; v < n; ++v) {
        float d = -1.0f;
        if (db[v] != INF) memcpy(&d, &db[v], 4);
        dist[v] = d;
    }
    /* parents: a BFS from the root over tight entries (fl(d[u] + w) == d[v]);
     * equal-distance neighbours joined by tiny weights are tight both ways,
     * so a parent is taken on first discovery only, which keeps chains acyclic */
#pragma omp parallel for num_threads(nthreads)
    for (uint64_t v = 0; v < n; ++v) pred[v] = -1;
    pred[root] = (int64_t)root;
    cur[0] = (uint32_t)root;
    uint64_t ncur = 1;
    while (ncur) {
        uint64_t nn = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
        for (uint64_t i = 0; i < ncur; ++i) {
            uint32_t u = cur[i];
            for (uint64_t k = rs[u]; k < rs[u + 1]; ++k) {
                uint32_t v = col[k];
                volatile float cand = dist[u] + w[k];
                if (cand != dist[v] || __atomic_load_n(&pred[v], __ATOMIC_RELAXED) != -1) continue;
                int64_t expect = -1;
                if (!__atomic_compare_exchange_n(&pred[v], &expect, (int64_t)u, 0, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED))
                    continue;
                uint64_t at = __atomic_fetch_add(&nn, 1, __ATOMIC_RELAXED);
                nxt[at] = v;
            }
        }
        uint32_t *t = cur;
        cur = nxt;
        nxt = t;
        ncur = nn;
    }
    free(db);
    free(settled);
    free(inb);
    free(mark);
    free(members);
    free(nxt);
    free(cur);
}

uint64_t o_sssp_check(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *w,
                      uint64_t root, const float *dist, const int64_t *pred) {
    uint64_t err = 0;
    if (dist[root] != 0.0f || pred[root] != (int64_t)root) ++err;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : err)
    for (uint64_t u = 0; u < n; ++u) {
        int reached_u = dist[u] >= 0.0f;
        if (!reached_u) {
            if (dist[u] != -1.0f || pred[u] != -1) ++err;
            continue;
        }
        for (uint64_t k = rs[u]; k < rs[u + 1]; ++k) {
            uint32_t v = col[k];
            volatile float cand = dist[u] + w[k];
            if (dist[v] < 0.0f || dist[v] > cand) ++err; /* (a) */
        }
        if (u == root) continue;
        int64_t p = pred[u];
        if (p < 0 || (uint64_t)p >= n || dist[p] < 0.0f) { ++err; continue; }
        int tight = 0;
        for (uint64_t k = rs[u]; k < rs[u + 1] && !tight; ++k) {
            volatile float cand = dist[p] + w[k];
            if (col[k] == (uint32_t)p && cand == dist[u]) tight = 1;
        }
        if (!tight) ++err; /* (b) */
    }
    /* (c) parent chains reach the root without cycles. */
    uint8_t *state = (uint8_t *)calloc(n ? n : 1, 1); /* 0 new, 1 on path, 2 ok, 3 bad */
    uint32_t *path = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    state[root] = 2;
    for (uint64_t v = 0; v < n; ++v) {
        if (dist[v] < 0.0f || state[v]) continue;
        uint64_t len = 0, x = v;
        uint8_t verdict = 3;
        for (;;) {
            if (state[x] == 2) { verdict = 2; break; }
            if (state[x] == 3 || state[x] == 1) { verdict = 3; break; }
            int64_t p = pred[x];
            if (p < 0 || (uint64_t)p >= n) { verdict = 3; state[x] = 1; path[len++] = (uint32_t)x; break; }
            state[x] = 1;
            path[len++] = (uint32_t)x;
            x = (uint64_t)p;
        }
        for (uint64_t i = 0; i < len; ++i) {
            state[path[i]] = verdict;
            if (verdict == 3 && dist[path[i]] >= 0.0f) ++err; /* reached vertices whose chain fails */
        }
    }
    free(path);
    free(state);
    return err;
}

/* Depth certificate: d is the BFS depth array (hop counts, -1 unreached) of
 * root iff d[root] = 0, every reached v != root has a neighbour at d[v] - 1,
 * and no entry joins a reached vertex to an unreached one or to one more than
 * a level away (a path of length d[v] exists, and along any shortest path d
 * grows by at most one per hop). */
uint64_t o_depth_check(uint64_t n, const uint64_t *rs, const uint32_t *col, const int64_t *depth, uint64_t root) {
    uint64_t err = depth[root] != 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : err)
    for (uint64_t u = 0; u < n; ++u) {
        int64_t du = depth[u];
        if (du < -1) { ++err; continue; }
        if (du == -1) continue;
        int has_parent = u == root;
        for (uint64_t k = rs[u]; k < rs[u + 1]; ++k) {
            int64_t dv = depth[col[k]];
            if (dv < 0 || dv > du + 1 || dv < du - 1) ++err;
            if (dv == du - 1) has_parent = 1;
        }
        if (!has_parent) ++err;
    }
    return err;
}

/* ------------------------------------------------------------------------ */
/* Statistics (graph500_runner.cu:14-41)                                    */
/* ------------------------------------------------------------------------ */

static int cmp_f64(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

void o_statistics(const double *x, int n, double out[7]) {
    double mean = 0.0, var = 0.0;
    for (int i = 0; i < n; ++i) mean += x[i];
    mean /= n;
    for (int i = 0; i < n; ++i) var += (x[i] - mean) * (x[i] - mean);
    var /= n - 1;
    double *s = (double *)malloc((size_t)n * sizeof(double));
    memcpy(s, x, (size_t)n * sizeof(double));
    qsort(s, (size_t)n, sizeof(double), cmp_f64);
    out[0] = s[0];
    out[1] = (s[(n - 1) / 4] + s[n / 4]) * 0.5;
    out[2] = (s[(n - 1) / 2] + s[n / 2]) * 0.5;
    out[3] = (s[n - 1 - (n - 1) / 4] + s[n - 1 - n / 4]) * 0.5;
    out[4] = s[n - 1];
    out[5] = mean;
    out[6] = sqrt(var);
    free(s);
}
