/*
 * Base types, fatal-error macros and the bump arena of the MI355X Graph500
 * library (libg500.so). C-ABI drop-in for the reference's src/base.h.
 *
 * Replaces: src/base.h:11-25 (typedefs), :40-46 (ERROR / ABORT),
 * :56-80 (mem_align_forward, Arena, arena_*), src/base.cu:1-78.
 *
 * Deltas from the reference header:
 *   - CHECK_CUDA is not exported (the library checks HIP calls internally);
 *   - `true`/`false` are not redefined for C++ callers;
 *   - arena_create_gpu allocates device memory with hipMalloc (one
 *     allocation), not NVSHMEM's symmetric heap.
 */
#ifndef G500_BASE_H
#define G500_BASE_H
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef int8_t i8;
typedef int16_t i16;
typedef int32_t i32;
typedef int64_t i64;

typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;

typedef int32_t b32;
typedef int8_t b8;

typedef float f32;
typedef double f64;

#ifndef __cplusplus
#ifndef true
#define true 1
#define false 0
#endif
#endif

#define ABS(x) ((x) >= 0 ? (x) : -(x))
#define MAX(x, y) ((x) >= (y) ? (x) : (y))
#define MIN(x, y) ((x) <= (y) ? (x) : (y))

#define GIGABYTE(x) ((x)*1024L * 1024L * 1024L)
#define MEGABYTE(x) ((x)*1024L * 1024L)
#define KILOBYTE(x) ((x)*1024L)

#define INT_CEIL(a, b) (((a) + (b)-1) / (b))

#define ERROR(...)                    \
    do {                              \
        fprintf(stderr, __VA_ARGS__); \
        fputc('\n', stderr);          \
    } while (0)

#define ABORT() exit(EXIT_FAILURE)

#define ARENA_DEFAULT_ALIGNMENT 16

#ifdef __cplusplus
extern "C" {
#endif

typedef struct Arena Arena;
struct Arena {
    u64 reserved;
    u64 used;
    void *base;
};

/* base.h:56 */
u64 mem_align_forward(u64 size, u64 alignment);
/* base.h:68-72 */
Arena arena_create_cpu(u64 capacity);
Arena arena_create_gpu(u64 capacity);
Arena arena_from_arena(Arena *arena, u64 capacity);
void arena_release_cpu(Arena *arena);
void arena_release_gpu(Arena *arena);
/* base.h:74 — returns NULL (after printing) when the arena is exhausted. */
void *_arena_push(Arena *arena, u64 size, u64 alignment);
/* base.h:80 */
void arena_clear(Arena *arena);

#ifdef __cplusplus
}
#define G500_ALIGNOF(type) alignof(type)
#else
#define G500_ALIGNOF(type) _Alignof(type)
#endif

#define arena_push(arena, size) _arena_push(arena, size, ARENA_DEFAULT_ALIGNMENT)
#define arena_push_struct(arena, type) ((type *)_arena_push(arena, sizeof(type), G500_ALIGNOF(type)))
#define arena_push_array(arena, count, type) \
    ((type *)_arena_push(arena, (count) * sizeof(type), G500_ALIGNOF(type)))

#endif /* G500_BASE_H */
