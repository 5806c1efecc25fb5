/*
 * Single-rank runtime of the MI355X Graph500 library.
 *
 * Replaces src/worker.h:5-22 and src/worker.cu:6-109 with the single-rank
 * semantics: rank = 0, rank_size = 1, pes_on_device = 1. There is no MPI and
 * no NVSHMEM: worker_init picks HIP device 0, sizes the device arena exactly
 * as the reference does (free/2 with at least 512 MiB reserved when
 * gpu_memory_size == 0; override with G500_ARENA_BYTES), creates the
 * library's stream and allocates the arenas.
 *
 * worker_kernel_block_count / worker_kernel_launch (worker.h:17-18) take a
 * device-kernel pointer; they stay internal to the library.
 */
#ifndef G500_WORKER_H
#define G500_WORKER_H
#include "base.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct Worker_State Worker_State;
struct Worker_State {
    Arena cpu_arena;
    Arena gpu_arena;
    i32 rank;
    i32 rank_size;
    i32 pes_on_device;
};

/* worker.h:14 */
Worker_State worker_init(i32 argc, char **argv, u64 gpu_memory_size);
/* worker.h:15 */
void worker_finalize(Worker_State *ws);
/* worker.h:16 — releases the arenas and exits with return_code. */
void worker_abort(Worker_State *ws, i32 return_code);
/* worker.h:19 — one rank: *max_count_out = local_count. Device allocations
 * are 256-byte aligned at least. */
void *_worker_arena_push_array(Worker_State *ws, u64 local_count, u64 elem_size, u64 alignment,
                               u64 *max_count_out);

#ifdef __cplusplus
}
#endif

#define worker_arena_push_array(ws, count, type, max_count_out) \
    ((type *)_worker_arena_push_array((ws), (count), sizeof(type), G500_ALIGNOF(type), (max_count_out)))

#endif /* G500_WORKER_H */
