/*
 * CPU oracle for the Graph500 generator / CSR / BFS / SSSP path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (libg500.so and graph500_runner) never links or calls it.
 *
 * Every function restates the reference's semantics in plain C (own code) and
 * cites the reference file:line it follows (paths relative to the reference
 * repository root). How each function is pinned is listed in the header
 * comment of oracle.c.
 */
#ifndef G500_ORACLE_H
#define G500_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- MRG (src/graph_generation.cu:9-94, :209-266) ---------------------- */
typedef struct { uint32_t z[5]; } o_mrg_state;

/* One step of x_n = (107374182 x_{n-1} + 104480 x_{n-5}) mod (2^31-1);
 * returns the new z1 (graph_generation.cu:73-85). */
uint32_t o_mrg_next(o_mrg_state *st);
/* Full 5x5 transition matrix M^e applied to the state, e = hi*2^128 + mid*2^64
 * + lo (graph_generation.cu:209-223; table-free: byte powers of M are built
 * from M by repeated multiplication). */
void o_mrg_skip(o_mrg_state *st, uint64_t hi, uint64_t mid, uint64_t lo);
/* The 24x256 table of compressed 9-word transitions {s,t,u,v,w,a,b,c,d} for
 * M^(v * 256^b) (mrg_transitions.cu:14), derived from the full matrices. */
void o_mrg_skip_table(uint32_t *out /* 24*256*9 */);
/* Full 5x5 power M^(v*256^b), row-major, for tests. */
void o_mrg_power(int b, int v, uint32_t out[25]);
void o_make_mrg_seed(uint64_t s1, uint64_t s2, uint32_t seed[5]);
void o_make_random_numbers(int64_t nvalues, uint64_t s1, uint64_t s2, int64_t pos, double *out);
float o_mrg_weight(o_mrg_state *st);   /* get_float_orig, :88-90 */

/* ---- Kronecker generator (graph_generation.cu:100-310) ----------------- */
void o_scramble_values(const uint32_t seed[5], uint64_t *val0, uint64_t *val1);
uint64_t o_scramble(uint64_t v, int lgN, uint64_t val0, uint64_t val1);
/* edges: 2 u64 per edge (v0, v1); weights: f32 per edge. Edge indices
 * [start, end). OpenMP-parallel. */
void o_generate_kronecker_range(const uint32_t seed[5], int logN, int64_t start,
                                int64_t end, uint64_t *edges, float *weights);

/* ---- CSR (traversal.cu:15-73, :342-389) --------------------------------- */
/* Returns M (directed entries). rowstarts[n+1]; column[M]; cw[M]. Rows sorted
 * ascending by ((float bits of w) << 32 | column). Pass column==NULL to only
 * fill rowstarts and get M. */
uint64_t o_csr_build(uint64_t n, uint64_t nedges, const uint64_t *edges, const float *weights,
                     uint64_t *rowstarts, uint32_t *column, float *cw);

/* ---- BFS (traversal.cu:139-207, :505-563) ------------------------------- */
/* Serial queue BFS. pred: -1 unreached, root -> root. depth may be NULL.
 * Returns number of reached vertices. */
uint64_t o_bfs_serial(uint64_t n, const uint64_t *rs, const uint32_t *col, uint64_t root,
                      int64_t *pred, int64_t *depth);
/* Level-synchronous top-down OpenMP BFS (CPU baseline). Returns reached. */
uint64_t o_bfs_omp(uint64_t n, const uint64_t *rs, const uint32_t *col, uint64_t root,
                   int64_t *pred, int nthreads);
/* Depths (hop counts, -1 unreached) of a level-synchronous OpenMP BFS.
 * Returns the reached count. */
uint64_t o_bfs_depth_omp(uint64_t n, const uint64_t *rs, const uint32_t *col, uint64_t root,
                         int64_t *depth, int nthreads);
/* Order-free digests (sum mod 2^64 of a 64-bit mix per directed entry
 * (u, v, weight bits)) of the entries a tuple list stands for (both
 * directions of each non-loop edge; *entries gets their count) and of a CSR. */
uint64_t o_tuple_digest(uint64_t nedges, const uint64_t *edges, const float *weights, uint64_t *entries);
uint64_t o_csr_digest(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *cw);
/* Degree histogram of one chunk of a tuple list added into deg[n] (both
 * endpoints of each non-loop edge, traversal.cu:15-34); returns the edges with
 * an endpoint outside [0, n), which are skipped. */
uint64_t o_degree_add(uint64_t n, uint64_t nedges, const uint64_t *edges, uint64_t *deg);
/* Violations of the canonical CSR order (rowstarts from 0, non-decreasing;
 * rows ascending by (weight bits << 32 | column)). */
uint64_t o_csr_order_errors(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *cw);
/* bfs_compute_dist_from_pred semantics (traversal.cu:75-112, :505-563):
 * -1 unreached, INT64_MAX reached-but-unresolved, root 0. */
void o_dist_from_pred(uint64_t n, const uint64_t *rs, const uint32_t *col, const int64_t *pred,
                      uint64_t root, int64_t *dist);
/* Graph500 rules 1-7 with the runner's error counting
 * (examples/graph500_runner.cu:84-169, :203-255). Returns the error count. */
uint64_t o_bfs_validate(uint64_t n, const uint64_t *rs, const uint32_t *col, const int64_t *pred,
                        uint64_t root);
/* TEPS numerator (graph500_runner.cu:171-182). */
uint64_t o_edge_count(uint64_t n, const uint64_t *rs, const uint32_t *col, const int64_t *pred);
/* Root selection (graph500_runner.cu:43-74). Returns the root count. */
int o_generate_roots(uint64_t n, const uint64_t *rs, uint64_t s1, uint64_t s2, uint64_t *roots,
                     int max_roots);

/* ---- SSSP (traversal.cu:209-340, :638-711) ------------------------------ */
/* Binary-heap Dijkstra; dist -1.0f for unreached, pred root->root, -1 unreached. */
void o_dijkstra_f32(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *w,
                    uint64_t root, float *dist, int64_t *pred);
void o_dijkstra_f64(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *w,
                    uint64_t root, double *dist);
/* CPU-baseline sample: binary-heap FP32 Dijkstra that stops after
 * max_seconds; reports CSR entries scanned and vertices settled so far and
 * the elapsed seconds (whole run when it finishes first). */
void o_dijkstra_f32_timed(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *w,
                          uint64_t root, double max_seconds, uint64_t *scanned, uint64_t *settled,
                          double *elapsed);
/* The reference's delta-stepping bucket schedule run sequentially (same
 * light/heavy split and bucket bounds as traversal.cu:638-711). */
void o_delta_stepping_f32(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *w,
                          uint64_t root, float delta, float *dist, int64_t *pred);
/* CPU baseline: OpenMP delta-stepping (bucketed light rounds, heavy edges
 * once per bucket, empty buckets skipped), then a parent per reached vertex.
 * Same distances as Dijkstra; unreached -1.0f / -1, root -> root. */
void o_delta_stepping_omp(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *w,
                          uint64_t root, float delta, float *dist, int64_t *pred, int nthreads);
/* Parent certificate for FP32 SSSP output: counts violations of
 * (a) dist[v] <= fl(dist[u]+w) on every edge, (b) tight parent edge,
 * (c) acyclic parent chains reaching the root (one count per reached vertex
 * whose chain does not), (d) unreached consistency. */
uint64_t o_sssp_check(uint64_t n, const uint64_t *rs, const uint32_t *col, const float *w,
                      uint64_t root, const float *dist, const int64_t *pred);

/* BFS depth certificate: violations of d[root] = 0, a neighbour at d[v] - 1
 * for every reached v != root, and |d[u] - d[v]| <= 1 with both reached on
 * every entry of a reached u (values below -1 count too). Zero exactly for
 * the BFS depths of root. */
uint64_t o_depth_check(uint64_t n, const uint64_t *rs, const uint32_t *col, const int64_t *depth, uint64_t root);

/* ---- Statistics (graph500_runner.cu:14-41) ------------------------------ */
/* out: min, q1, median, q3, max, mean, stddev */
void o_statistics(const double *x, int n, double out[7]);

#ifdef __cplusplus
}
#endif
#endif
