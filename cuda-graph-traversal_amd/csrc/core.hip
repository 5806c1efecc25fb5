// Library context, errors, arenas, single-rank worker and control entry points.
// Reference: src/base.cu:1-78, src/worker.cu:6-109.
#include <cstring>
#include <mutex>
#include <string>

#include "common.hpp"

namespace g500 {

Context &ctx() {
    static Context c;
    return c;
}

thread_local ErrorSink *t_error_sink = nullptr;

static std::mutex &error_mutex() {
    static std::mutex m;
    return m;
}

void fail(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    if (ErrorSink *sink = t_error_sink) {
        if (!sink->status) vsnprintf(sink->msg, sizeof(sink->msg), fmt, ap);
        va_end(ap);
        sink->status = 1;
        return;
    }
    Context &c = ctx();
    std::lock_guard<std::mutex> lk(error_mutex());
    vsnprintf(c.msg, sizeof(c.msg), fmt, ap);
    va_end(ap);
    if (c.error_mode == G500_ERRORS_ABORT) {
        ERROR("%s", c.msg);
        ABORT();
    }
    if (!c.status) c.status = 1;
}

void release_library_resources() {
    Context &c = ctx();
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    if (c.d_counters) (void)hipFree(c.d_counters);
    if (c.d_mrg_table) (void)hipFree((void *)c.d_mrg_table);
    if (c.ev_start) (void)hipEventDestroy(c.ev_start);
    if (c.ev_stop) (void)hipEventDestroy(c.ev_stop);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    c.d_counters = nullptr;
    c.d_mrg_table = nullptr;
    c.ev_start = c.ev_stop = nullptr;
    c.stream = nullptr;
}

void *arena_alloc(Worker_State *ws, u64 bytes, const char *what) {
    void *p = _arena_push(&ws->gpu_arena, bytes ? bytes : 1, 256);
    if (!p) fail("device arena exhausted allocating %llu bytes for %s (used %llu of %llu)",
                 (unsigned long long)bytes, what, (unsigned long long)ws->gpu_arena.used,
                 (unsigned long long)ws->gpu_arena.reserved);
    return p;
}

void timer_start() { G500_HIP(hipEventRecord(ctx().ev_start, ctx().stream)); }

void timer_stop() {
    Context &c = ctx();
    G500_HIP(hipEventRecord(c.ev_stop, c.stream));
    G500_HIP(hipEventSynchronize(c.ev_stop));
    float ms = 0.f;
    G500_HIP(hipEventElapsedTime(&ms, c.ev_start, c.ev_stop));
    c.last_ms = ms;
}

int cooperative_blocks(const void *kernel, int block, size_t dyn_lds) {
    int per_cu = 0;
    G500_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, dyn_lds));
    hipFuncAttributes attr{};
    G500_HIP(hipFuncGetAttributes(&attr, kernel));
    int lds_cu = 0;
    G500_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, ctx().device));
    const size_t lds = attr.sharedSizeBytes + dyn_lds;
    constexpr size_t kLdsRound = 2048;
    if (lds && lds_cu > 0) per_cu = std::min(per_cu, (int)((size_t)lds_cu / ((lds + kLdsRound - 1) / kLdsRound * kLdsRound)));
    if (per_cu <= 0) {
        fail("kernel cannot be resident with %d threads per block", block);
        return 1;
    }
    return per_cu * ctx().num_cus;
}

void launch_persistent(const void *kernel, int grid, int block, void **args, hipStream_t stream,
                       bool cooperative, size_t dyn_lds) {
    Context &c = ctx();
    if (!stream) stream = c.stream;
    if (cooperative && c.coop_launch) {
        hipError_t e = hipLaunchCooperativeKernel(kernel, dim3(grid), dim3(block), args, dyn_lds, stream);
        if (e == hipSuccess) return;
        (void)hipGetLastError();
        static bool warned = false;
        if (!warned) {
            fprintf(stderr, "libg500: cooperative launch unavailable (%s); using plain launches of the "
                            "co-resident grid\n", hipGetErrorString(e));
            warned = true;
        }
    }
    G500_HIP(hipLaunchKernel(kernel, dim3(grid), dim3(block), args, dyn_lds, stream));
}

void init_mrg_table();  // generator.hip

__global__ __launch_bounds__(256) void copy16_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, u64 n) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void copy1_kernel(u8 *__restrict__ dst, const u8 *__restrict__ src, u64 n) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) dst[i] = src[i];
}

}  // namespace g500

using namespace g500;

extern "C" {

// ------------------------------------------------------------------ arenas
u64 mem_align_forward(u64 size, u64 alignment) { return (size + alignment - 1) & ~(alignment - 1); }

void *_arena_push(Arena *arena, u64 size, u64 alignment) {
    if (!arena->base) {
        ERROR("Arena was not initialized");
        return nullptr;
    }
    u64 at = mem_align_forward(arena->used, alignment);
    if (at + size > arena->reserved) {
        ERROR("Arena out of space");
        return nullptr;
    }
    arena->used = at + size;
    return (u8 *)arena->base + at;
}

Arena arena_create_cpu(u64 capacity) {
    Arena a;
    a.reserved = mem_align_forward(capacity, ARENA_DEFAULT_ALIGNMENT);
    a.used = 0;
    a.base = malloc(a.reserved);
    return a;
}

Arena arena_create_gpu(u64 capacity) {
    Arena a;
    a.reserved = mem_align_forward(capacity, 256);
    a.used = 0;
    a.base = nullptr;
    if (a.reserved) {
        hipError_t e = hipMalloc(&a.base, a.reserved);
        if (e != hipSuccess || !a.base) {
            a.base = nullptr;
            fail("Failed to allocate a %llu-byte device arena: %s", (unsigned long long)a.reserved,
                 hipGetErrorString(e));
            a.reserved = 0;
        }
    }
    return a;
}

Arena arena_from_arena(Arena *arena, u64 capacity) {
    Arena a;
    a.reserved = mem_align_forward(capacity, 256);
    a.used = 0;
    a.base = _arena_push(arena, a.reserved, 256);
    return a;
}

void arena_release_cpu(Arena *arena) {
    if (arena->base) {
        free(arena->base);
        *arena = Arena{0, 0, nullptr};
    }
}

void arena_release_gpu(Arena *arena) {
    if (arena->base) {
        (void)hipFree(arena->base);
        *arena = Arena{0, 0, nullptr};
    }
}

void arena_clear(Arena *arena) { arena->used = 0; }

// ------------------------------------------------------------------ worker
Worker_State worker_init(i32 argc, char **argv, u64 gpu_memory_size) {
    (void)argc;
    (void)argv;
    Context &c = ctx();
    Worker_State ws;
    memset(&ws, 0, sizeof(ws));
    ws.rank = 0;
    ws.rank_size = 1;
    ws.pes_on_device = 1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        fail("Rank 0: no HIP devices available");
        return ws;
    }
    c.device = 0;
    G500_HIP(hipSetDevice(c.device));
    hipDeviceProp_t prop;
    G500_HIP(hipGetDeviceProperties(&prop, c.device));
    c.num_cus = prop.multiProcessorCount;
    int wc = 0;
    if (hipDeviceGetAttribute(&wc, hipDeviceAttributeWallClockRate, c.device) == hipSuccess && wc > 0)
        c.wall_clock_khz = wc;
    if (!c.stream) {
        G500_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        G500_HIP(hipEventCreate(&c.ev_start));
        G500_HIP(hipEventCreate(&c.ev_stop));
    }
    if (const char *env = getenv("G500_ARENA_BYTES")) gpu_memory_size = strtoull(env, nullptr, 0);
    if (gpu_memory_size == 0) {  // worker.cu:29-38 with one PE on the device
        size_t freeb = 0, total = 0;
        G500_HIP(hipMemGetInfo(&freeb, &total));
        u64 f = freeb;
        u64 reserve = MAX(f / 2, (u64)MEGABYTE(512));
        gpu_memory_size = f > reserve ? f - reserve : f / 2;
    }
    if (failed()) return ws;
    init_mrg_table();
    resolve_noalloc_scratch();
    ws.cpu_arena = arena_create_cpu(MEGABYTE(256));
    ws.gpu_arena = arena_create_gpu(gpu_memory_size);
    return ws;
}

void worker_finalize(Worker_State *ws) {
    // worker.cu:57-62 releases the arenas; the library's stream, events, skip
    // table and counters go too, so nothing is left for exit-time teardown
    if (ctx().stream) (void)hipStreamSynchronize(ctx().stream);
    arena_release_gpu(&ws->gpu_arena);
    arena_release_cpu(&ws->cpu_arena);
    release_library_resources();
}

void worker_abort(Worker_State *ws, i32 return_code) {
    worker_finalize(ws);
    exit(return_code);
}

void *_worker_arena_push_array(Worker_State *ws, u64 local_count, u64 elem_size, u64 alignment,
                               u64 *max_count_out) {
    if (max_count_out) *max_count_out = local_count;
    return _arena_push(&ws->gpu_arena, local_count * elem_size, MAX(alignment, (u64)256));
}

// ------------------------------------------------------------------ control
void g500_set_error_mode(i32 mode) { ctx().error_mode = mode; }
i32 g500_status(void) { return ctx().status; }
const char *g500_last_error(void) { return ctx().msg; }
void g500_clear_error(void) {
    std::lock_guard<std::mutex> lk(error_mutex());
    ctx().status = 0;
    ctx().msg[0] = 0;
}
f64 g500_last_run_ms(void) { return ctx().last_ms; }

void g500_copy_to_host(void *dst, const void *src_device, u64 bytes) {
    if (!bytes) return;
    G500_HIP(hipMemcpyAsync(dst, src_device, bytes, hipMemcpyDeviceToHost, ctx().stream));
    sync();
}

void g500_copy_to_device(void *dst_device, const void *src, u64 bytes) {
    if (!bytes) return;
    G500_HIP(hipMemcpyAsync(dst_device, src, bytes, hipMemcpyHostToDevice, ctx().stream));
    sync();
}

void g500_copy_device(void *dst_device, const void *src_device, u64 bytes) {
    Context &c = ctx();
    if (!bytes) return;
    timer_start();
    if (((uintptr_t)dst_device | (uintptr_t)src_device | bytes) % 16 == 0)
        copy16_kernel<<<grid_for(bytes / 16, 256, (u64)c.num_cus * 16), 256, 0, c.stream>>>(
            (uint4 *)dst_device, (const uint4 *)src_device, bytes / 16);
    else
        copy1_kernel<<<grid_for(bytes, 256, (u64)c.num_cus * 16), 256, 0, c.stream>>>((u8 *)dst_device,
                                                                                     (const u8 *)src_device, bytes);
    G500_HIP(hipGetLastError());
    timer_stop();
}

i32 g500_set_param(const char *name, f64 value) {
    Context &c = ctx();
    std::string n(name);
    if (n == "bfs_alpha") c.bfs_alpha = value;
    else if (n == "bfs_beta") c.bfs_beta = value;
    else if (n == "batch_alpha") c.batch_alpha = value;
    else if (n == "batch_beta") c.batch_beta = value;
    else if (n == "batch_heavy_degree") c.batch_heavy_degree = (int)value;
    else if (n == "bfs_block" || n == "sssp_block" || n == "batch_bfs_block" || n == "batch_sssp_block") {
        // persistent kernels are built for at most 512 threads per block (wave multiples)
        const int v = (int)value;
        if (v != 128 && v != 256 && v != 512) return -1;
        (n == "bfs_block" ? c.bfs_block : n == "sssp_block" ? c.sssp_block
                                        : n == "batch_bfs_block" ? c.batch_bfs_block : c.batch_sssp_block) = v;
    }
    else if (n == "bfs_hub_degree") c.bfs_hub_degree = (int)value;
    else if (n == "gen_blocks") c.gen_blocks = (int)value;
    else if (n == "coop_launch") c.coop_launch = (int)value;
    else if (n == "relabel") c.relabel = (int)value;
    else if (n == "sssp_prune") c.sssp_prune = (int)value;
    else if (n == "sssp_screen") c.sssp_screen = (int)value;
    else if (n == "sssp_pull") c.sssp_pull = value;
    else if (n == "batch_pull") c.batch_pull = value;
    else return -1;
    return 0;
}

i32 g500_get_param(const char *name, f64 *value) {
    Context &c = ctx();
    std::string n(name);
    if (n == "bfs_alpha") *value = c.bfs_alpha;
    else if (n == "bfs_beta") *value = c.bfs_beta;
    else if (n == "batch_alpha") *value = c.batch_alpha;
    else if (n == "batch_beta") *value = c.batch_beta;
    else if (n == "batch_heavy_degree") *value = c.batch_heavy_degree;
    else if (n == "bfs_block") *value = c.bfs_block;
    else if (n == "bfs_hub_degree") *value = c.bfs_hub_degree;
    else if (n == "sssp_block") *value = c.sssp_block;
    else if (n == "batch_bfs_block") *value = c.batch_bfs_block;
    else if (n == "batch_sssp_block") *value = c.batch_sssp_block;
    else if (n == "gen_blocks") *value = c.gen_blocks;
    else if (n == "coop_launch") *value = c.coop_launch;
    else if (n == "relabel") *value = c.relabel;
    else if (n == "sssp_prune") *value = c.sssp_prune;
    else if (n == "sssp_screen") *value = c.sssp_screen;
    else if (n == "sssp_pull") *value = c.sssp_pull;
    else if (n == "batch_pull") *value = c.batch_pull;
    else if (n == "num_cus") *value = c.num_cus;
    else if (n == "sssp_batch_max") *value = kSsspBatchRoots;
    else return -1;
    return 0;
}

}  // extern "C"
