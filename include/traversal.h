/*
 * CSR graph, BFS and delta-stepping SSSP of the MI355X Graph500 library.
 *
 * Replaces src/traversal.h:8-54 and src/traversal.cu:1-721 (single rank).
 * All pointers inside the structs are device pointers into the worker's
 * device arena; *_destroy and oned_graph_free only zero the struct
 * (traversal.cu:565-568, :713-721); memory returns with worker_finalize.
 * Every entry point is blocking: it returns after the device work is done.
 *
 * Layout deltas from the reference (the fields the reference runner reads
 * keep their names and meaning):
 *   - Oned_Graph.column is u32 (32-bit vertex ids, SCALE <= 32) instead of
 *     u64; row offsets stay u64. Each row is stored in ascending
 *     (weight, column) order, so the CSR is deterministic (the reference's
 *     in-row order depends on atomic arrival, traversal.cu:51).
 *   - Oned_Graph also carries a degree-ordered traversal copy of the CSR
 *     (fields after `weights`) and an SSSP copy of it without two-hop-
 *     dominated entries; results are reported in the original ids.
 *   - Bfs_State keeps pred/dist/max_nlocalverts; the frontier counters and
 *     grid-barrier words (traversal.h:23-26) live in the device control block
 *     `ctl`; frontiers are u32 queues plus bitmaps for bottom-up levels.
 *   - Sssp_State keeps pred/dist/max_nlocalverts; the working labels are one
 *     packed u64 per vertex (f32 distance bits << 32 | parent), lowered by a
 *     64-bit atomic min once a u32 distance probe (the dist array while the
 *     kernel runs) is strictly lowered, instead of dist_bits + pred_lock
 *     (traversal.cu:211-240); delta is a field, 0.1f by default
 *     (traversal.cu:645).
 */
#ifndef G500_TRAVERSAL_H
#define G500_TRAVERSAL_H
#include "base.h"
#include "graph_generation.h"
#include "worker.h"

#ifdef __cplusplus
extern "C" {
#endif

/* traversal.h:8-16 */
typedef struct Oned_Graph Oned_Graph;
struct Oned_Graph {
    u64 nlocalverts;
    u64 nlocaledges;
    u64 nglobalverts, notisolated;
    u64 *rowstarts; /* [nlocalverts + 1] */
    u32 *column;    /* [nlocaledges] */
    f32 *weights;   /* [nlocaledges] */
    /* Traversal copy built with the CSR: vertices renumbered by decreasing
     * degree (so the most-probed labels share cache lines); same rows and
     * entry order. BFS/SSSP run on it and report in the original ids. All
     * NULL when the copy is switched off (g500_set_param("relabel", 0)): the
     * traversals then run on the arrays above. */
    u32 *perm;          /* [n] original id -> traversal id */
    u32 *iperm;         /* [n] traversal id -> original id */
    u64 *rowstarts_t;   /* [n + 1] */
    u32 *column_t;      /* [nlocaledges] traversal ids */
    f32 *weights_t;     /* [nlocaledges] */
    /* SSSP copy, built after the traversal copy (all NULL without it, or with
     * g500_set_param("sssp_prune", 0)): the traversal rows, same order,
     * without each entry (u, v, w) that a two-hop path u-x-v through one of
     * the 32 highest-degree vertices beats by a margin,
     * w(u,x) + w(x,v) + 2^-10 < w (w < 2). No such entry is needed by any shortest path while distances
     * stay below 2048, so both SSSP kernels relax this copy (and report error
     * 32 on a distance at or above 2048); BFS, validation and the TEPS counts
     * use the full rows. */
    u64 *rowstarts_s;   /* [n + 1] */
    u32 *column_s;      /* [rowstarts_s[n]] traversal ids */
    f32 *weights_s;     /* [rowstarts_s[n]] */
};

/* traversal.h:18-27 */
typedef struct Bfs_State Bfs_State;
struct Bfs_State {
    i64 *pred;             /* [n] parent; -1 unreached; root -> root */
    i64 *dist;             /* [n] depth, allocated by bfs_compute_dist_from_pred */
    u32 *frontier[2];      /* [n] vertex queues */
    u64 *frontier_bits[2]; /* [ceil(n/64)] frontier bitmaps */
    u64 *visited;          /* [ceil(n/64)] */
    u32 *hubs;             /* [2n] queues of high-degree frontier vertices */
    u64 *hub_offsets;      /* [2n] their offsets in the flattened edge work space */
    u32 *hub_lengths;      /* [2n] their degrees */
    void *ctl;             /* device control block */
    u64 max_nlocalverts;
};

/* traversal.h:30-41 */
typedef struct Sssp_State Sssp_State;
struct Sssp_State {
    i64 *pred;          /* [n] parent; -1 unreached; root -> root */
    f32 *dist;          /* [n] distance; -1.0f unreached */
    u64 *dist_pred;     /* [n] packed working labels */
    u32 *frontier;      /* [6n] light-phase queues: 2 buffers x 3 work tiers */
    u32 *heavy;         /* [2n] heavy-pass lists: wave tier, hub tier */
    u64 *hub_offsets;   /* [3n] hub-tier work offsets (2 light queues, 1 heavy list) */
    u32 *hub_lengths;   /* [3n] hub-tier row-segment lengths */
    u32 *split;         /* [n] number of light entries (w < split_delta) per row */
    void *rows;         /* [n] 16-byte row records for the same split: row start (u64), light
                           entries (u32), row length (u32) */
    u32 *stamp;         /* [n] light-queue de-duplication */
    u64 *settled;       /* [ceil(n/64)] vertices whose distance is final */
    u64 *reached;       /* [ceil(n/64)] vertices with a finite label */
    void *ctl;          /* device control block */
    const void *split_graph; /* row offsets that `split` was computed for */
    f32 delta;          /* bucket width (default 0.1f) */
    f32 split_delta;    /* delta that `split` was computed for (0: none yet) */
    u64 max_nlocalverts;
};

/* traversal.h:43 */
Oned_Graph oned_graph_from_tuple_graph(Worker_State *ws, Tuple_Graph *tg, u64 nglobalverts);
/* traversal.h:44 */
b32 oned_graph_is_vertex_isolated(Worker_State *ws, const Oned_Graph *g, u64 global_vertex);
/* traversal.h:45-49 */
Bfs_State oned_graph_bfs_create(Worker_State *ws, const Oned_Graph *g);
void oned_graph_bfs_clear(Worker_State *ws, Bfs_State *bfs);
void oned_graph_bfs_run(Worker_State *ws, const Oned_Graph *g, Bfs_State *bfs, u64 root);
i64 *bfs_compute_dist_from_pred(Worker_State *ws, const Oned_Graph *g, Bfs_State *bfs, u64 root);
void oned_graph_bfs_destroy(Bfs_State *bfs);
/* traversal.h:50-54 */
Sssp_State oned_graph_sssp_create(Worker_State *ws, const Oned_Graph *g);
void oned_graph_sssp_clear(Worker_State *ws, Sssp_State *sssp);
void oned_graph_sssp_run(Worker_State *ws, const Oned_Graph *g, Sssp_State *sssp, u64 root);
void oned_graph_sssp_destroy(Sssp_State *sssp);
void oned_graph_free(Oned_Graph *g);

#ifdef __cplusplus
}
#endif
#endif /* G500_TRAVERSAL_H */
