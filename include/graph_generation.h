/*
 * Graph500 Kronecker edge generator with the splittable MRG.
 *
 * Replaces src/graph_generation.h:6-43 (types and entry points) and the code
 * of src/graph_generation.cu / src/mrg_transitions.cu. The skip table is
 * regenerated from the MRG recurrence at library start (same values as
 * mrg_transitions.cu:14) and uploaded to the device.
 */
#ifndef G500_GRAPH_GENERATION_H
#define G500_GRAPH_GENERATION_H
#include "base.h"
#include "worker.h"

#ifdef __cplusplus
extern "C" {
#endif

/* graph_generation.h:6-10 */
typedef struct Packed_Edge Packed_Edge;
struct Packed_Edge {
    u64 v0;
    u64 v1;
};

/* graph_generation.h:12-19. edges/weights are device pointers. */
typedef struct Tuple_Graph Tuple_Graph;
struct Tuple_Graph {
    Packed_Edge *edges;
    u64 edges_size;
    u64 max_edges_size;
    u64 nglobaledges;
    f32 *weights;
};

/* graph_generation.h:21-25 */
typedef struct Mrg_Transition_Matrix Mrg_Transition_Matrix;
struct Mrg_Transition_Matrix {
    u32 s, t, u, v, w;
    u32 a, b, c, d;
};

/* graph_generation.h:27-30 */
typedef struct Mrg_State Mrg_State;
struct Mrg_State {
    u32 z1, z2, z3, z4, z5;
};

/* graph_generation.h:33-37: edges [start_edge, end_edge) of the graph with
 * 2^logN vertices; allocates tg->edges / tg->weights from the device arena,
 * launches the generator and waits for it. */
void generate_kronecker_range(Worker_State *st, u32 seed[5], u32 logN, i64 start_edge, i64 end_edge,
                              Tuple_Graph *tg);
/* graph_generation.h:39 */
void make_mrg_seed(u64 userseed1, u64 userseed2, u32 *seed);
/* graph_generation.h:40 */
void make_random_numbers(i64 nvalues, u64 userseed1, u64 userseed2, i64 position, f64 *result);

/* graph_generation.h:42-43: raw little-endian files graph-refactor.bin
 * (Packed_Edge x E) and graph-refactor.bin.weights (f32 x E), written at the
 * byte offset of start_edge (graph_generation.cu:6-7, :314-396). The file
 * names can be overridden with G500_DUMP_PATH (".weights" is appended). */
void tuple_graph_dump(const Tuple_Graph *tg, i64 start_edge);
void tuple_graph_load(Worker_State *ws, Tuple_Graph *tg, i64 start_edge, i64 edge_count);

#ifdef __cplusplus
}
#endif
#endif /* G500_GRAPH_GENERATION_H */
