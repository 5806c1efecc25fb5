// Graph500 Kronecker generator: MRG skip table, host RNG entry points, the
// gfx950 edge kernel and the raw tuple dump/load format.
// Reference: src/graph_generation.cu:6-7, :100-396, src/mrg_transitions.cu:10-14.
#include <fcntl.h>
#include <unistd.h>

#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "mrg.hpp"

namespace g500 {

// Host copy of the skip table: kSkip[b][v] = M^(v * 256^b) in the 9-word form,
// M the one-step matrix of the MRG. Rebuilt from the recurrence (the same
// values as the table of mrg_transitions.cu:14).
static std::vector<Mrg_Transition_Matrix> kSkip;

namespace {
struct Mat5 {
    u64 a[5][5];
};

Mat5 mul(const Mat5 &x, const Mat5 &y) {
    Mat5 r;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            u64 s = 0;
            for (int k = 0; k < 5; ++k) s = (s + x.a[i][k] * y.a[k][j] % kP) % kP;
            r.a[i][j] = s;
        }
    return r;
}

Mrg_Transition_Matrix compress(const Mat5 &m) {
    Mrg_Transition_Matrix t;
    t.s = (u32)m.a[4][0];
    t.t = (u32)m.a[4][1];
    t.u = (u32)m.a[4][2];
    t.v = (u32)m.a[4][3];
    t.w = (u32)m.a[4][4];
    t.a = (u32)m.a[3][0];
    t.b = (u32)m.a[2][0];
    t.c = (u32)m.a[1][0];
    t.d = (u32)m.a[0][0];
    return t;
}

// Inverse of compress(): the full matrix the 9 words stand for (used to check
// that every power really has the compressed structure).
bool structured(const Mat5 &m) {
    Mrg_Transition_Matrix t = compress(m);
    const u64 Y = kA5;
    u64 full[5][5] = {{t.d, Y * t.s % kP, Y * t.a % kP, Y * t.b % kP, Y * t.c % kP},
                      {t.c, t.w, Y * t.s % kP, Y * t.a % kP, Y * t.b % kP},
                      {t.b, t.v, t.w, Y * t.s % kP, Y * t.a % kP},
                      {t.a, t.u, t.v, t.w, Y * t.s % kP},
                      {t.s, t.t, t.u, t.v, t.w}};
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j)
            if (full[i][j] != m.a[i][j]) return false;
    return true;
}
}  // namespace

// Host table only (no device work): used by the host RNG entry points.
static void build_host_table() {
    if (kSkip.empty()) {
        kSkip.resize(24 * 256);
        Mat5 m1 = {};
        m1.a[0][0] = kA1;
        m1.a[0][4] = kA5;
        for (int i = 1; i < 5; ++i) m1.a[i][i - 1] = 1;
        Mat5 ident = {};
        for (int i = 0; i < 5; ++i) ident.a[i][i] = 1;
        Mat5 step = m1;  // M^(256^b)
        for (int b = 0; b < 24; ++b) {
            Mat5 p = ident;
            for (int v = 0; v < 256; ++v) {
                if (!structured(p)) {
                    fail("MRG power %d/%d lost the companion structure", b, v);
                    return;
                }
                kSkip[b * 256 + v] = compress(p);
                p = mul(p, step);
            }
            step = p;  // M^(256 * 256^b)
        }
    }
}

void init_mrg_table() {
    Context &c = ctx();
    build_host_table();
    if (!kSkip.empty() && !c.d_mrg_table) {
        void *d = nullptr;
        G500_HIP(hipMalloc(&d, kSkip.size() * sizeof(Mrg_Transition_Matrix)));
        G500_HIP(hipMemcpy(d, kSkip.data(), kSkip.size() * sizeof(Mrg_Transition_Matrix),
                           hipMemcpyHostToDevice));
        c.d_mrg_table = (const Mrg_Transition_Matrix *)d;
    }
}

// graph_generation.cu:209-223: skip hi*2^128 + mid*2^64 + lo steps.
static void host_skip(Mrg_State &st, u64 hi, u64 mid, u64 lo) {
    build_host_table();
    const u64 parts[3] = {lo, mid, hi};
    for (int p = 0; p < 3; ++p)
        for (int k = 0; k < 8 && (parts[p] >> (8 * k)); ++k) {
            u32 byte = (u32)(parts[p] >> (8 * k)) & 0xFF;
            if (byte) mrg_apply(kSkip[(8 * p + k) * 256 + byte], st);
        }
}

G500_HD u64 scramble(u64 v, int lgN, u64 sum01, u64 mul0, u64 mul1) {
    v += sum01;
    v *= mul0;
    v = __builtin_bitreverse64(v) >> (64 - lgN);
    v *= mul1;
    v = __builtin_bitreverse64(v) >> (64 - lgN);
    return v;
}

// One block per 256-edge tile whose first edge index is a multiple of 256:
// the state skip over bytes 1..7 of the edge index is uniform across the block
// (scalar), each lane then applies the power for its low byte.
__global__ __launch_bounds__(256) void kronecker_kernel(
    Mrg_State seed, const Mrg_Transition_Matrix *__restrict__ T, int logN, i64 start, i64 end,
    u64 sum01, u64 mul0, u64 mul1, Packed_Edge *__restrict__ edges, f32 *__restrict__ weights,
    i64 tile_first, i64 tile_count) {
    for (i64 t = blockIdx.x; t < tile_count; t += gridDim.x) {
        const i64 tile = tile_first + t;
        Mrg_State st = seed;
        u64 hi = (u64)tile;
        for (int b = 9; hi; ++b, hi >>= 8) {
            const u32 byte = (u32)hi & 0xFF;
            if (byte) mrg_apply(T[b * 256 + byte], st);
        }
        const i64 ei = tile * 256 + threadIdx.x;
        if (ei < start || ei >= end) continue;
        if (threadIdx.x) mrg_apply(T[8 * 256 + threadIdx.x], st);
        // graph_generation.cu:114-143, :186-207: one quadrant per level,
        // rejecting u < 0x7FFFFFFF mod 10000.
        u64 src = 0, tgt = 0;
        for (int lvl = logN - 1; lvl >= 0; --lvl) {
            u32 u = mrg_next(st);
            while (u < 3647u) u = mrg_next(st);
            u %= 10000u;
            u32 so, to;
            if (u < 1900u) { so = 0; to = 1; }         // B
            else if (u < 3800u) { so = 1; to = 0; }    // C
            else if (u < 9500u) { so = 0; to = 0; }    // A
            else { so = 1; to = 1; }                   // D
            if (src == tgt && so > to) { so = 0; to = 1; }
            src |= (u64)so << lvl;
            tgt |= (u64)to << lvl;
        }
        Packed_Edge e;
        e.v0 = scramble(src, logN, sum01, mul0, mul1);
        e.v1 = scramble(tgt, logN, sum01, mul0, mul1);
        edges[ei - start] = e;
        weights[ei - start] = (f32)mrg_next(st) * 0x1p-31f;  // graph_generation.cu:88-90
    }
}

}  // namespace g500

using namespace g500;

extern "C" {

void make_mrg_seed(u64 userseed1, u64 userseed2, u32 *seed) {
    // graph_generation.cu:248-254 (Graph500 user seed mapping).
    seed[0] = (u32)(userseed1 & 0x3FFFFFFFu) + 1;
    seed[1] = (u32)((userseed1 >> 30) & 0x3FFFFFFFu) + 1;
    seed[2] = (u32)(userseed2 & 0x3FFFFFFFu) + 1;
    seed[3] = (u32)((userseed2 >> 30) & 0x3FFFFFFFu) + 1;
    seed[4] = (u32)((userseed2 >> 60) << 4) + (u32)(userseed1 >> 60) + 1;
}

void make_random_numbers(i64 nvalues, u64 userseed1, u64 userseed2, i64 position, f64 *result) {
    u32 seed[5];
    make_mrg_seed(userseed1, userseed2, seed);
    Mrg_State st = {seed[0], seed[1], seed[2], seed[3], seed[4]};
    host_skip(st, 2, 0, 2 * (u64)position);
    for (i64 i = 0; i < nvalues; ++i) result[i] = (f64)mrg_next(st) * 0x1.00000002p-31;  // 1/(2^31-1)
}

void generate_kronecker_range(Worker_State *ws, u32 seed[5], u32 logN, i64 start_edge, i64 end_edge,
                              Tuple_Graph *tg) {
    Context &c = ctx();
    i64 count = end_edge - start_edge;
    u64 local = count > 0 ? (u64)count : 0;
    tg->edges_size = local;
    tg->max_edges_size = local;
    tg->edges = (Packed_Edge *)arena_alloc(ws, local * sizeof(Packed_Edge), "tuple edges");
    tg->weights = (f32 *)arena_alloc(ws, local * sizeof(f32), "tuple weights");
    if (failed() || count <= 0) return;
    if (logN < 1 || logN > 32) {
        fail("generate_kronecker_range: logN %u outside [1, 32]", logN);
        return;
    }
    Mrg_State st = {seed[0], seed[1], seed[2], seed[3], seed[4]};
    // graph_generation.cu:290-300: scramble values from the state skipped by
    // 50 * 2^128 + 7 * 2^64.
    Mrg_State s2 = st;
    host_skip(s2, 50, 7, 0);
    u64 val0 = mrg_next(s2);
    val0 = val0 * 0xFFFFFFFFull + mrg_next(s2);
    u64 val1 = mrg_next(s2);
    val1 = val1 * 0xFFFFFFFFull + mrg_next(s2);

    const i64 tile_first = start_edge >> 8;
    const i64 tile_count = ((end_edge - 1) >> 8) - tile_first + 1;
    i64 grid = c.gen_blocks > 0 ? c.gen_blocks : MIN(tile_count, (i64)c.num_cus * 2048);
    timer_start();
    kronecker_kernel<<<(unsigned)grid, 256, 0, c.stream>>>(
        st, c.d_mrg_table, (int)logN, start_edge, end_edge, val0 + val1, val0 | 0x4519840211493211ull,
        val1 | 0x3050852102C843A5ull, tg->edges, tg->weights, tile_first, tile_count);
    G500_HIP(hipGetLastError());
    timer_stop();
}

// ------------------------------------------------------------- dump / load
static std::string dump_path(bool weights) {
    const char *env = getenv("G500_DUMP_PATH");
    std::string base = env ? env : "graph-refactor.bin";
    return weights ? base + ".weights" : base;
}

static bool transfer_file(int fd, void *dev, u64 bytes, u64 offset, bool write) {
    const u64 chunk = 64ull << 20;
    std::vector<u8> buf(MIN(bytes, chunk));
    for (u64 done = 0; done < bytes;) {
        u64 n = MIN(chunk, bytes - done);
        if (write) {
            g500_copy_to_host(buf.data(), (u8 *)dev + done, n);
            for (u64 k = 0; k < n;) {
                ssize_t r = pwrite(fd, buf.data() + k, n - k, (off_t)(offset + done + k));
                if (r <= 0) return false;
                k += (u64)r;
            }
        } else {
            for (u64 k = 0; k < n;) {
                ssize_t r = pread(fd, buf.data() + k, n - k, (off_t)(offset + done + k));
                if (r <= 0) return false;
                k += (u64)r;
            }
            g500_copy_to_device((u8 *)dev + done, buf.data(), n);
        }
        done += n;
    }
    return true;
}

void tuple_graph_dump(const Tuple_Graph *tg, i64 start_edge) {
    const char *paths[2] = {nullptr, nullptr};
    std::string pe = dump_path(false), pw = dump_path(true);
    paths[0] = pe.c_str();
    paths[1] = pw.c_str();
    const u64 esz[2] = {sizeof(Packed_Edge), sizeof(f32)};
    void *src[2] = {tg->edges, tg->weights};
    for (int f = 0; f < 2; ++f) {
        int fd = open(paths[f], O_CREAT | O_WRONLY, 0644);
        if (fd < 0) {
            fail("tuple_graph_dump: cannot open %s", paths[f]);
            return;
        }
        bool ok = ftruncate(fd, (off_t)(tg->nglobaledges * esz[f])) == 0 &&
                  transfer_file(fd, src[f], tg->edges_size * esz[f], (u64)start_edge * esz[f], true);
        close(fd);
        if (!ok) {
            fail("tuple_graph_dump: write to %s failed", paths[f]);
            return;
        }
    }
}

void tuple_graph_load(Worker_State *ws, Tuple_Graph *tg, i64 start_edge, i64 edge_count) {
    u64 local = edge_count > 0 ? (u64)edge_count : 0;
    tg->edges_size = local;
    tg->max_edges_size = local;
    tg->edges = (Packed_Edge *)arena_alloc(ws, local * sizeof(Packed_Edge), "tuple edges");
    tg->weights = (f32 *)arena_alloc(ws, local * sizeof(f32), "tuple weights");
    if (failed() || !local) return;
    std::string pe = dump_path(false), pw = dump_path(true);
    const char *paths[2] = {pe.c_str(), pw.c_str()};
    const u64 esz[2] = {sizeof(Packed_Edge), sizeof(f32)};
    void *dst[2] = {tg->edges, tg->weights};
    for (int f = 0; f < 2; ++f) {
        int fd = open(paths[f], O_RDONLY);
        if (fd < 0) {
            fail("tuple_graph_load: cannot open %s", paths[f]);
            return;
        }
        bool ok = transfer_file(fd, dst[f], local * esz[f], (u64)start_edge * esz[f], false);
        close(fd);
        if (!ok) {
            fail("tuple_graph_load: read of %s failed (file shorter than requested range?)", paths[f]);
            return;
        }
    }
}

}  // extern "C"
