/*
 * Library entry points beyond the reference headers (src/ *.h).
 *
 * The first group restates helpers that are `static` in the reference's
 * runner (examples/graph500_runner.cu) so that any host program can validate
 * and count like the runner does. The second group is library control
 * (error mode, timing, host<->device copies, tuning knobs) used by the
 * runner, the tests and the benchmark.
 */
#ifndef G500_EXT_H
#define G500_EXT_H
#include "traversal.h"

#ifdef __cplusplus
extern "C" {
#endif

/* graph500_runner.cu:43-74: up to max_roots distinct non-isolated roots drawn
 * from make_random_numbers(2, seed1, seed2, counter). Returns the count. */
i32 generate_bfs_roots(Worker_State *ws, const Oned_Graph *g, u64 seed1, u64 seed2, u64 *roots,
                       i32 max_roots);
/* graph500_runner.cu:171-201: sum over vertices with pred != -1 of the CSR
 * entries u <= v (the TEPS numerator). pred is a device pointer. */
u64 oned_graph_edge_count(Worker_State *ws, const Oned_Graph *g, const i64 *pred);
/* graph500_runner.cu:84-169, :203-255: Graph500 rules 1-7 on bfs->pred
 * (depths derived with bfs_compute_dist_from_pred). Returns true when no
 * error is found; *errors_out (may be NULL) receives the error count. */
b32 oned_graph_bfs_validate(Worker_State *ws, const Oned_Graph *g, Bfs_State *bfs, u64 root,
                            u64 *errors_out);
/* No counterpart in the reference. The same validation for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): it calls no hipMalloc, hipFree or
 * arena push. Inputs: the graph and a device parent array pred[n]; outputs:
 * depth[n] (device, bfs_compute_dist_from_pred's values) and *errors_out.
 * Its scratch (confirmation bits and counters, 16 MiB + 256 B) is a static
 * device array that each GPU holds until the process ends, so graphs are
 * limited to 2^27 vertices (larger ones fail with an error). Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU race. Blocking;
 * needs worker_init first. */
b32 oned_graph_bfs_validate_noalloc(const Oned_Graph *g, const i64 *pred, u64 root, i64 *depth,
                                    u64 *errors_out);
/* Device-side certificates, no counterpart in the reference (its runner
 * validates BFS parent trees only, graph500_runner.cu:84-169); used to check
 * batched results, which carry depths or (distance, parent) pairs. Both return
 * the violation count (0: the result is the exact one), arrays on the device.
 * Depths: d[root] = 0, every reached v != root has a neighbour at d[v] - 1,
 * every entry of a reached vertex joins a reached vertex at most one level
 * away (values below -1 are violations); this holds exactly for the BFS
 * depths. SSSP (FP32 distances, -1.0f unreached; parents, -1 unreached, root
 * -> root): d[v] <= fl(d[u] + w) on every entry of a reached u, a tight
 * reached parent entry for every reached v != root, parent chains that end at
 * the root; this holds exactly for the FP32 shortest-path distances. */
u64 oned_graph_bfs_depth_check(Worker_State *ws, const Oned_Graph *g, const i64 *depth, u64 root);
u64 oned_graph_sssp_check(Worker_State *ws, const Oned_Graph *g, const f32 *dist, const i64 *pred, u64 root);
/* graph500_runner.cu:22-41: min, q1, median, q3, max, mean, stddev. */
void get_statistics(const f64 *x, i32 n, f64 r[7]);

/* ---- batched multi-source BFS ------------------------------------------- */
/* Up to 64 roots traversed together in one persistent kernel: per-vertex u64
 * masks (bit r = reached from root r), top-down and bottom-up levels over all
 * roots at once. No counterpart in the reference (its runner runs one root at
 * a time, graph500_runner.cu:306-336); the depths per root equal the plain
 * BFS's. The result is a discovery log: for every level, the vertices that
 * some root reached first at that level and the mask of those roots. */
typedef struct Bfs_Batch_State Bfs_Batch_State;
struct Bfs_Batch_State {
    u64 *seen;         /* [n] roots that reached each vertex */
    u64 *fmask[2];     /* [n] per-level frontier masks (bottom-up levels) */
    u64 *fbits[2];     /* [ceil(n/64)+1] frontier bitmaps matching fmask */
    u64 *next;         /* [n] top-down accumulation of the next level */
    u32 *touched;      /* [n] vertices touched in a top-down level */
    u32 *log_vertex;   /* [log_capacity] discovery log: vertex */
    u64 *log_mask;     /* [log_capacity] discovery log: newly reaching roots */
    u64 *level_start;  /* [n + 2] first log record of each level */
    u32 *hubs;         /* [3n] hub lists: records of two consecutive top-down
                        * levels, then the heavy vertices of bottom-up levels */
    u64 *hub_offsets;  /* [3n] */
    u32 *hub_lengths;  /* [3n] */
    void *ctl;
    u64 log_capacity;
    u64 max_nlocalverts;
    i32 nroots;        /* roots of the last run */
    i32 levels;        /* levels of the last run (records exist for 0..levels-1) */
};
Bfs_Batch_State oned_graph_bfs_batch_create(Worker_State *ws, const Oned_Graph *g);
/* roots: host array of nroots (1..64) vertex ids (repeats allowed). */
void oned_graph_bfs_batch_run(Worker_State *ws, const Oned_Graph *g, Bfs_Batch_State *b, const u64 *roots,
                              i32 nroots);
/* Depth of every vertex from root index r of the last run into the device
 * array depth[n] (-1: unreached), the convention of bfs_compute_dist_from_pred. */
void oned_graph_bfs_batch_depth(Worker_State *ws, const Oned_Graph *g, const Bfs_Batch_State *b, i32 r,
                                i64 *depth);
void oned_graph_bfs_batch_destroy(Bfs_Batch_State *b);
/* Per-level log of the last batched run: mode (0 top-down, 1 bottom-up),
 * record count, degree sum of the records and device milliseconds (any
 * pointer may be NULL). Returns the number of levels. */
i32 g500_bfs_batch_level_log(const Bfs_Batch_State *b, i32 *modes, u64 *records, u64 *edges, f64 *ms,
                             i32 max_levels);

/* ---- batched delta-stepping SSSP ---------------------------------------- */
/* Launch grouping for batched SSSP, no counterpart in the reference: writes to
 * order[0..nroots) the permutation that sorts roots[] (host array) by their
 * distance from the graph's highest-degree vertex (unreached roots first,
 * ties in input order). Shortest paths of Kronecker graphs run through the
 * hubs, so roots at similar hub distances reach every vertex at similar
 * distances; consecutive roots in this order, launched together, settle each
 * vertex in the same buckets and share more work. Runs one single-root SSSP
 * on temporary arena memory. Returns 0, or -1 after a failure. */
i32 oned_graph_sssp_root_order(Worker_State *ws, const Oned_Graph *g, const u64 *roots, i32 nroots, i32 *order);
/* No counterpart in the reference: out[i] = FP32 distance of roots[i] (host
 * array) from the graph's highest-degree vertex, -1.0f when unreached (one
 * single-root SSSP on temporary arena memory). These are the offsets that
 * oned_graph_sssp_batch_run_offsets turns into bucket lags. Returns 0, or -1
 * after a failure. */
i32 oned_graph_sssp_hub_distances(Worker_State *ws, const Oned_Graph *g, const u64 *roots, i32 nroots, f32 *out);
/* Up to 32 roots in one persistent kernel with one bucket sequence shared by
 * all roots; items are (vertex, root mask) pairs and the 32 distance probes of
 * a vertex share one 128-byte line. Edges with w >= mid_weight are relaxed in
 * deferred batches, in three classes: mid (mid_weight <= w < long_weight),
 * long (long_weight <= w < far_weight) and far (w >= far_weight). A class-c
 * relaxation lands at least W_c above the
 * bucket that settled its source, so the kernel collects the roots a vertex
 * settles over a W_c-wide window and relaxes the class's edges of the vertex
 * once for all of them.
 * No counterpart in the reference (its runner solves one root at a time,
 * graph500_runner.cu:393-416); each root's distances equal the single-root
 * oned_graph_sssp_run's (the FP32 fixed point is unique), and its parents
 * satisfy the same certificate. */
typedef struct Sssp_Batch_State Sssp_Batch_State;
struct Sssp_Batch_State {
    f32 *dist;          /* [32n] root-major: root index r, vertex v at r * n + v; -1.0f unreached
                           (vertex-major distance probes while a run is in flight) */
    i64 *pred;          /* [32n] root-major parents; -1 unreached; root -> root */
    u64 *labels;        /* [32n] vertex-major working labels (v * 32 + r) */
    u32 *pending;       /* [2n] pending root masks, two steps */
    u32 *frontier;      /* [6n] light queues: 2 buffers x 3 work tiers */
    u64 *heavy;         /* [2n] heavy items (mask << 32 | vertex), two work tiers */
    u32 *hub_vertices;  /* [n] heavy hub tier */
    u64 *hub_offsets;   /* [3n] hub work offsets (2 light queues, heavy) */
    u32 *hub_lengths;   /* [3n] hub segment lengths */
    u32 *hub_aux;       /* [3n] light: finished chunks per hub; heavy: root masks */
    u32 *split;         /* [n] light entries (w < split_delta) per row */
    u32 *root_masks;    /* [2n] per vertex: roots whose label is final, then the smallest distance
                           bits among the roots not yet final */
    u64 *open_words;    /* [ceil(n/64)] vertices with a reached root not yet final */
    u32 *mid_split;     /* [n] entries with w < W_mid per row (W_mid = mid_weight clamped to [delta, W_long]) */
    u32 *mid_work;      /* [6n] deferred mid edges: pending root masks, queue tiers 0-2, hub lengths,
                           hub chunk counts */
    u64 *mid_offsets;   /* [n] hub work offsets of the deferred mid queue */
    u32 *far_split;     /* [n] entries with w < W_far per row (W_far = max(far_weight, delta)) */
    u32 *far_work;      /* [6n] deferred far edges, laid out as mid_work */
    u64 *far_offsets;   /* [n] hub work offsets of the deferred far queue */
    u32 *long_split;    /* [n] entries with w < W_long per row (W_long = long_weight clamped to [delta, W_far]) */
    u32 *long_work;     /* [6n] deferred long edges: pending root masks, queue tiers 0-2, hub lengths,
                           hub chunk counts */
    u64 *long_offsets;  /* [n] hub work offsets of the deferred queue */
    void *ctl;          /* device control block */
    const void *split_graph; /* row offsets that `split` / `long_split` were computed for */
    f32 delta;          /* bucket width (default 0.1f) */
    f32 split_delta;
    f32 long_weight;    /* long-edge threshold (default 0.25f; <= delta: no deferral window) */
    f32 long_split_weight;
    f32 mid_weight;     /* mid-edge threshold (default 0.075f; <= delta: no mid class) */
    f32 mid_split_weight;
    f32 far_weight;     /* far-edge threshold (default 0.5f; <= long: no far class) */
    f32 far_split_weight;
    i32 nroots;         /* roots of the last run */
    i32 reserved;
    u64 max_nlocalverts;
    void *row_info;     /* [n] 32-byte row records: start, light, mid, long and far splits, end */
};
Sssp_Batch_State oned_graph_sssp_batch_create(Worker_State *ws, const Oned_Graph *g);
/* roots: host array of nroots (1..32, g500_get_param("sssp_batch_max")) vertex ids (repeats allowed). */
void oned_graph_sssp_batch_run(Worker_State *ws, const Oned_Graph *g, Sssp_Batch_State *s, const u64 *roots,
                               i32 nroots);
/* No counterpart in the reference: oned_graph_sssp_batch_run with per-root
 * bucket lags. offsets: host array of nroots distance offsets (typically
 * oned_graph_sssp_hub_distances of the roots; a negative or non-finite entry
 * means none). Root r runs the usual bucket sequence lagging the shared bucket
 * counter by round((max offset - offsets[r]) / delta) buckets, so roots whose
 * distances differ by about their offsets settle each vertex in the same
 * buckets and share the row passes. Distances and the parent certificate are
 * those of oned_graph_sssp_batch_run for any offsets: every label-correcting
 * schedule reaches the same FP32 fixed point. */
void oned_graph_sssp_batch_run_offsets(Worker_State *ws, const Oned_Graph *g, Sssp_Batch_State *s, const u64 *roots,
                                       const f32 *offsets, i32 nroots);
void oned_graph_sssp_batch_destroy(Sssp_Batch_State *s);
/* Bucket / light-step / label-improvement counts of the last batched run
 * (improvements summed over the roots). */
void g500_sssp_batch_run_stats(const Sssp_Batch_State *s, u64 *buckets, u64 *light_iterations,
                               u64 *relaxations);
/* Device milliseconds of the last batched run: light steps, dense steps
 * (scan + medium + flush), min + collect steps, whole kernel; ms[4] = dense
 * passes, ms[5] = deferred-queue flushes; ms[6..8] = the dense steps' scan,
 * medium and flush parts. */
void g500_sssp_batch_phase_ms(const Sssp_Batch_State *s, f64 ms[9]);
/* Work of the last batched run per row segment k (0 light, 1 medium, 3 mid,
 * 4 long, 5 far): out[3k + 0] items, out[3k + 1] row entries visited,
 * out[3k + 2] (entry, root) relaxations; out[6..8] of the dense scans: open
 * 64-vertex words walked, vertices that read their probe line, vertices
 * settling a root. */
void g500_sssp_batch_work(const Sssp_Batch_State *s, u64 out[18]);
/* Pull-form far-class flushes of the last batched run (g500_set_param("batch_pull", share): a flush whose
 * window items are at least share x the vertices with entries is pulled by the targets): out[0] pulled
 * flushes, out[1] live (vertex, root) pairs, out[2] (vertex, entry) scan steps, out[3] vertices finished by
 * whole waves. */
void g500_sssp_batch_pull_work(const Sssp_Batch_State *s, u64 out[4]);
/* Per-phase log of the last batched run (first 4096 phases): kind (0 light
 * step, 1 dense scan, 2 medium relaxations, 3 deferred flush, 4 min + collect
 * steps), items processed and device milliseconds. Returns the entries. */
i32 g500_sssp_batch_step_log(const Sssp_Batch_State *s, i32 *kind, u64 *items, f64 *ms, i32 max_steps);

/* ---- serving loop -------------------------------------------------------- */
/* Resident graphs, (graph, root, BFS | SSSP) requests and continuous batching.
 * No counterpart in the reference (its runner is a batch program). Execution
 * slots each own a HIP stream and batched BFS / SSSP states; a free slot takes
 * the oldest pending request's (graph, kind) plus further pending requests of
 * that pair (up to 64 BFS / 32 SSSP roots) that the device result pool admits
 * (footprint per request: BFS n * 8 bytes of depths; SSSP n * 4 bytes of
 * distances + n * 8 bytes of parents when requested). Slots run concurrently,
 * each with its share of the device. Results are those of the batched entry
 * points, so a request's result does not depend on what it was batched with. */
typedef struct G500_Server G500_Server;
enum { G500_BFS = 0, G500_SSSP = 1 };
typedef struct G500_Completion G500_Completion;
struct G500_Completion {
    u64 id;            /* from g500_server_submit */
    i32 graph, kind;
    u64 root;
    i32 status;        /* 0: ok; 1..63: the kernel's error bits; 65536: the launch or a result copy
                          failed on the host side (no results are copied) */
    i32 batch_size;    /* requests in the launch that served it */
    u64 reached;       /* vertices reached from the root */
    f64 sum;           /* BFS: sum of depths; SSSP: sum of distances (2^-24 fixed point) */
    f64 latency_ms;    /* submit -> results in the caller's buffers */
    f64 queue_ms;      /* submit -> launch */
    f64 run_ms;        /* launch -> results staged on the device */
};
/* slots: 1..8 concurrent launches; result_pool_bytes: device memory for staged
 * results (requests wait for admission when it is full). */
G500_Server *g500_server_create(Worker_State *ws, i32 slots, u64 result_pool_bytes);
/* Registers a resident graph (before start); returns its id. */
i32 g500_server_add_graph(G500_Server *s, const Oned_Graph *g);
/* Allocates the slot states (sized for the largest graph) and starts the
 * threads; delta is the SSSP bucket width of batches of up to 16 roots
 * (larger batches run at delta / 2). Returns 0, or -1 on error. */
i32 g500_server_start(G500_Server *s, f32 delta);
/* Queues a request; thread-safe. out0: host buffer for BFS depths (i64[n]) or
 * SSSP distances (f32[n]); out1: SSSP parents (i64[n]); either may be NULL
 * (then only the summary is returned). The buffers must stay valid until the
 * completion is polled. Returns the request id, 0 on error. */
u64 g500_server_submit(G500_Server *s, i32 graph, i32 kind, u64 root, void *out0, void *out1);
/* Submits `count` requests at the given offsets (ms after the call), from the
 * calling thread, without result buffers; returns when all are submitted. */
void g500_server_replay(G500_Server *s, u64 count, const f64 *at_ms, const i32 *graph, const i32 *kind,
                        const u64 *root);
/* Up to `max` completions, waiting up to wait_ms for the first. Returns the count. */
i32 g500_server_poll(G500_Server *s, G500_Completion *out, i32 max, f64 wait_ms);
/* Returns when every submitted request has completed. */
void g500_server_drain(G500_Server *s);
void g500_server_stats(const G500_Server *s, u64 *submitted, u64 *completed, u64 *launches, u64 *pool_used);
/* Drains, stops the threads and frees the host side (device memory returns
 * with worker_finalize). */
void g500_server_destroy(G500_Server *s);

/* ---- library control --------------------------------------------------- */
/* G500_ERRORS_ABORT (default): fatal errors print and exit like the
 * reference. G500_ERRORS_RETURN: they set a status and the entry point
 * returns early (used by tests and serving code). */
#define G500_ERRORS_ABORT 0
#define G500_ERRORS_RETURN 1
void g500_set_error_mode(i32 mode);
i32 g500_status(void);
const char *g500_last_error(void);
void g500_clear_error(void);

/* Device time in milliseconds of the last oned_graph_bfs_run /
 * oned_graph_sssp_run / generate_kronecker_range /
 * oned_graph_from_tuple_graph, measured with HIP events on the library's
 * stream around its kernels. */
f64 g500_last_run_ms(void);

/* Blocking copies between host memory and the library's device memory. */
void g500_copy_to_host(void *dst, const void *src_device, u64 bytes);
void g500_copy_to_device(void *dst_device, const void *src, u64 bytes);
/* Device-to-device copy with the library's own streaming kernel (16-byte
 * vectors when both pointers and the size allow it); also the reference
 * workload for calibrating HBM byte counters. */
void g500_copy_device(void *dst_device, const void *src_device, u64 bytes);

/* Tuning knobs (names: "bfs_alpha", "bfs_beta", "bfs_hub_degree", "batch_alpha",
 * "batch_beta", "batch_heavy_degree", "bfs_block" and "sssp_block" (threads
 * per block of the persistent BFS / SSSP kernels, single and batched: 128,
 * 256 or 512), "gen_blocks", "coop_launch", "relabel"; read-only "num_cus", "sssp_batch_max").
 * Returns 0 on success, -1 for an unknown name or a value out of range. */
i32 g500_set_param(const char *name, f64 value);
i32 g500_get_param(const char *name, f64 *value);

/* Per-level log of the last BFS run: mode (0 top-down, 1 bottom-up),
 * frontier size and device milliseconds of each level (any pointer may be
 * NULL). Returns the number of levels. */
i32 g500_bfs_level_log(const Bfs_State *bfs, i32 *modes, u64 *frontier, f64 *ms, i32 max_levels);
/* Per-step log of the last SSSP run (first 1023 steps): kind (0 light step,
 * 1 fused heavy step, 2 min step, 3 collect step), queued vertices, padded
 * hub-tier entries and device milliseconds. Returns the number of steps. */
i32 g500_sssp_step_log(const Sssp_State *sssp, i32 *kind, u64 *items, u64 *hub_entries, f64 *ms, i32 max_steps);
/* Bucket / light-iteration / label-improvement counts of the last SSSP run. */
void g500_sssp_run_stats(const Sssp_State *sssp, u64 *buckets, u64 *light_iterations,
                         u64 *relaxations);
/* Device milliseconds of the last SSSP run spent in light iterations,
 * heavy passes, bucket scans and the whole kernel; ms[4] = number of scans. */
void g500_sssp_phase_ms(const Sssp_State *sssp, f64 ms[5]);

#ifdef __cplusplus
}
#endif
#endif /* G500_EXT_H */
