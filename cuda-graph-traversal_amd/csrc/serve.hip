// Serving loop: resident graphs, (graph, root, BFS | SSSP) requests and
// continuous batching. No counterpart in the reference (its runner is a batch
// program over 64 roots, examples/graph500_runner.cu:257-441).
//
// Requests enter a pending queue. Each execution slot (a host thread, a HIP
// stream, and one batched BFS and one batched SSSP state sized for the largest
// graph) repeatedly takes the oldest pending request's (graph, kind) and as
// many further pending requests of that pair as a launch holds (64 BFS roots,
// 32 SSSP roots) and as the result pool admits: each request reserves its
// result footprint (BFS: i64 depths; SSSP: f32 distances + i64 parents) in a
// device pool, and a request that does not fit waits until completions free
// memory. Requests that arrive while launches run join the next launch of a
// free slot. The slot launches the batch on its stream with its share of the
// co-resident grid, so several batches run on the device at once; when the
// batch ends it stages every request's results in the pool (device copies),
// computes per-request summaries, and hands the requests to a copy thread,
// which copies the results to the callers' host buffers, frees the pool
// memory and completes the requests.
//
// Co-residency: a slot's persistent kernels synchronize their whole grid, so
// every block of a launch must be resident at once. With 2 or 4 slots each
// slot's stream carries a CU mask that gives it its own 1/slots of the CUs of
// every XCD (mask bits i with (i / 8) % slots == k), and its launches are
// sized to fill exactly those CUs; kernels of different slots then never
// compete for a CU. Other slot counts share all CUs with launches of 1/slots
// of the co-resident grid (residency then relies on the dispatcher spreading
// blocks evenly; a grid that is not resident ends with a barrier timeout,
// reported as the batch's status).
//
// Row splits and row records of the batched SSSP depend on the graph and on
// the bucket width; the server computes them once per graph for both widths
// at start, and every slot reads them from there.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "common.hpp"

namespace g500 {

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    u64 id;
    i32 graph, kind;
    u64 root;
    void *out0, *out1;
    u64 off0, off1;  // pool offsets of the staged results (~0: none)
    u64 bytes0, bytes1;
    double t_submit, t_launch, t_staged;
    i32 batch;
    i32 status;
    u64 reached;
    f64 sum;
};

// First-fit allocator over the device result pool (256-byte granules).
class Pool {
public:
    void init(u64 bytes) {
        size_ = bytes;
        free_.clear();
        if (bytes) free_[0] = bytes;
    }
    u64 alloc(u64 bytes) {
        bytes = (bytes + 255) & ~255ull;
        if (!bytes) return 0;
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes) {
                const u64 off = it->first, rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used_ += bytes;
                return off;
            }
        return ~0ull;
    }
    void release(u64 off, u64 bytes) {
        bytes = (bytes + 255) & ~255ull;
        if (!bytes || off == ~0ull) return;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    u64 size() const { return size_; }
    u64 used() const { return used_; }

private:
    std::map<u64, u64> free_;
    u64 size_ = 0, used_ = 0;
};

}  // namespace

}  // namespace g500

using namespace g500;

struct G500_Server {
    Worker_State *ws = nullptr;
    int nslots = 0;
    std::vector<Oned_Graph> graphs;
    u8 *pool_base = nullptr;
    Pool pool;
    std::mutex mu;
    std::condition_variable cv_work, cv_copy, cv_done;
    std::deque<Request *> pending, copying;
    std::deque<G500_Completion> done;
    u64 next_id = 1, submitted = 0, completed = 0, in_system = 0;
    bool started = false, stopping = false;
    struct Slot {
        hipStream_t stream = nullptr;
        Bfs_Batch_State bfs;
        Sssp_Batch_State sssp;
        u64 *summary = nullptr;  // device [128]
        int grid_bfs = 0, grid_sssp = 0;
        int cus = 0;             // CUs of the slot's mask (0: no mask)
        std::thread th;
    };
    std::vector<Slot> slots;
    std::thread copier;
    hipStream_t copy_stream = nullptr;
    f32 delta = 0.05f;
    // per graph: SSSP row splits and row records for batches of up to 16 roots
    // (delta) and of more (delta / 2)
    std::vector<std::array<Sssp_Batch_State, 2>> splits;
    // per graph: distance of every vertex from the graph's highest-degree
    // vertex (host copy), the key that groups SSSP requests into launches
    std::vector<std::vector<f32>> hubd;
    u64 launches = 0, launched_requests = 0;
};

namespace {

u64 footprint0(const G500_Server *s, const Request *r) {
    const u64 n = s->graphs[r->graph].nlocalverts;
    return r->kind == G500_SSSP ? n * sizeof(f32) : n * sizeof(i64);
}
u64 footprint1(const G500_Server *s, const Request *r) {
    return r->kind == G500_SSSP && r->out1 ? s->graphs[r->graph].nlocalverts * sizeof(i64) : 0;
}

// Takes the next batch under the lock: the oldest request fixes (graph, kind);
// later requests of the pair join while the launch has room and the pool
// admits their footprint. When more SSSP requests of the pair wait than a
// launch holds, those whose roots lie closest to the oldest one's in hub
// distance are taken (the oldest first): roots at similar hub distances share
// more work in one launch.
std::vector<Request *> take_batch(G500_Server *s) {
    std::vector<Request *> batch;
    if (s->pending.empty()) return batch;
    const i32 graph = s->pending.front()->graph, kind = s->pending.front()->kind;
    const size_t cap = kind == G500_BFS ? 64 : kSsspBatchRoots;
    std::vector<Request *> cand;
    for (Request *r : s->pending)
        if (r->graph == graph && r->kind == kind) cand.push_back(r);
    if (kind == G500_SSSP && cand.size() > cap && !s->hubd[graph].empty()) {
        const std::vector<f32> &hd = s->hubd[graph];
        const f32 d0 = hd[cand[0]->root];
        std::stable_sort(cand.begin() + 1, cand.end(), [&](const Request *x, const Request *y) {
            return std::fabs(hd[x->root] - d0) < std::fabs(hd[y->root] - d0);
        });
    }
    for (Request *r : cand) {
        if (batch.size() == cap) break;
        r->bytes0 = r->out0 ? footprint0(s, r) : 0;
        r->bytes1 = footprint1(s, r);
        r->off0 = r->bytes0 ? s->pool.alloc(r->bytes0) : ~0ull;
        if (r->bytes0 && r->off0 == ~0ull) break;  // pool full: admit later
        r->off1 = r->bytes1 ? s->pool.alloc(r->bytes1) : ~0ull;
        if (r->bytes1 && r->off1 == ~0ull) {
            s->pool.release(r->off0, r->bytes0);
            break;
        }
        batch.push_back(r);
    }
    if (!batch.empty()) {
        std::deque<Request *> rest;
        std::vector<Request *> taken(batch);
        std::sort(taken.begin(), taken.end());
        for (Request *r : s->pending)
            if (!std::binary_search(taken.begin(), taken.end(), r)) rest.push_back(r);
        s->pending.swap(rest);
    }
    return batch;
}

// Status of a request whose batch failed on the host side (launch refused,
// copy failed); kernel errors are reported with their own bits (1..63).
constexpr i32 kServeHostError = 1 << 16;

// Points a slot's SSSP state at the server's splits of one graph and width.
void use_splits(Sssp_Batch_State *st, const Sssp_Batch_State &c) {
    st->delta = c.delta;
    st->mid_weight = c.mid_weight;
    st->long_weight = c.long_weight;
    st->far_weight = c.far_weight;
    st->split = c.split;
    st->mid_split = c.mid_split;
    st->long_split = c.long_split;
    st->far_split = c.far_split;
    st->row_info = c.row_info;
    st->split_delta = c.split_delta;
    st->mid_split_weight = c.mid_split_weight;
    st->long_split_weight = c.long_split_weight;
    st->far_split_weight = c.far_split_weight;
    st->split_graph = c.split_graph;
}

void slot_loop(G500_Server *s, int k) {
    ErrorSink sink;
    t_error_sink = &sink;
    G500_HIP(hipSetDevice(ctx().device));
    G500_Server::Slot &slot = s->slots[k];
    std::vector<u64> roots;
    std::vector<f32> offsets;
    u64 summary[128];
    for (;;) {
        std::vector<Request *> batch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stopping) return;
                batch = take_batch(s);
                if (!batch.empty()) break;
                s->cv_work.wait(lk);
            }
            ++s->launches;
            s->launched_requests += batch.size();
        }
        const double t_launch = now_ms();
        const Oned_Graph *g = &s->graphs[batch[0]->graph];
        const u64 n = g->nlocalverts;
        const bool bfs = batch[0]->kind == G500_BFS;
        roots.clear();
        for (Request *r : batch) roots.push_back(r->root);
        sink.status = 0;
        sink.msg[0] = 0;
        u32 err = 0;
        bool ok;
        if (bfs) {
            ok = bfs_batch_enqueue(g, &slot.bfs, roots.data(), (i32)roots.size(), slot.stream, slot.grid_bfs);
            if (ok) err = bfs_batch_error(&slot.bfs, slot.stream);
            if (ok && !err) {
                for (size_t i = 0; i < batch.size(); ++i)
                    if (batch[i]->bytes0)
                        bfs_batch_depth_enqueue(g, &slot.bfs, (i32)i, (i64 *)(s->pool_base + batch[i]->off0),
                                                slot.stream);
                bfs_batch_summary_enqueue(&slot.bfs, slot.summary, slot.stream);
            }
        } else {
            // block shape and grid share were fixed at start; batches of more than
            // 16 roots share more work per bucket and run at half the server's delta
            use_splits(&slot.sssp, s->splits[batch[0]->graph][roots.size() > 16 ? 1 : 0]);
            // bucket lags from the roots' hub distances, so that the roots settle
            // each vertex in the same buckets (scheduling only)
            const std::vector<f32> &hd = s->hubd[batch[0]->graph];
            offsets.clear();
            for (u64 r : roots) offsets.push_back(hd.empty() ? -1.0f : hd[r]);
            ok = sssp_batch_enqueue(g, &slot.sssp, roots.data(), offsets.data(), (i32)roots.size(), slot.stream,
                                    slot.grid_sssp);
            if (ok) err = sssp_batch_error(&slot.sssp, slot.stream);
            if (ok && !err) {
                for (size_t i = 0; i < batch.size(); ++i) {
                    Request *r = batch[i];
                    if (r->bytes0)
                        G500_HIP(hipMemcpyAsync(s->pool_base + r->off0, slot.sssp.dist + i * n, n * sizeof(f32),
                                                hipMemcpyDeviceToDevice, slot.stream));
                    if (r->bytes1)
                        G500_HIP(hipMemcpyAsync(s->pool_base + r->off1, slot.sssp.pred + i * n, n * sizeof(i64),
                                                hipMemcpyDeviceToDevice, slot.stream));
                }
                sssp_batch_summary_enqueue(g, &slot.sssp, slot.summary, slot.stream);
            }
        }
        if (ok && !err) {
            G500_HIP(hipMemcpyAsync(summary, slot.summary, sizeof(summary), hipMemcpyDeviceToHost, slot.stream));
            G500_HIP(hipStreamSynchronize(slot.stream));
        }
        // a failed launch or copy leaves nothing valid to stage: the requests
        // complete with an error status and no results
        const i32 status = failed() ? kServeHostError : (i32)err;
        if (status) {
            fprintf(stderr, "libg500 server: slot %d batch of %zu %s requests on graph %d failed (status %d%s%s)\n", k,
                    batch.size(), bfs ? "BFS" : "SSSP", batch[0]->graph, status, sink.status ? ": " : "", sink.msg);
            (void)hipStreamSynchronize(slot.stream);
        }
        const double t_staged = now_ms();
        const int R = bfs ? 64 : kSsspBatchRoots;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (size_t i = 0; i < batch.size(); ++i) {
                Request *r = batch[i];
                r->t_launch = t_launch;
                r->t_staged = t_staged;
                r->batch = (i32)batch.size();
                r->status = status;
                r->reached = status ? 0 : summary[i];
                r->sum = status ? 0.0 : bfs ? (f64)summary[R + i] : (f64)summary[R + i] / 16777216.0;
                if (status) {  // nothing to copy out
                    s->pool.release(r->off0, r->bytes0);
                    s->pool.release(r->off1, r->bytes1);
                    r->bytes0 = r->bytes1 = 0;
                    r->off0 = r->off1 = ~0ull;
                }
                s->copying.push_back(r);
            }
        }
        s->cv_copy.notify_one();
    }
}

void copy_loop(G500_Server *s) {
    ErrorSink sink;
    t_error_sink = &sink;
    G500_HIP(hipSetDevice(ctx().device));
    for (;;) {
        Request *r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_copy.wait(lk, [&] { return s->stopping || !s->copying.empty(); });
            if (s->copying.empty()) return;  // stopping
            r = s->copying.front();
            s->copying.pop_front();
        }
        sink.status = 0;
        if (r->bytes0)
            G500_HIP(hipMemcpyAsync(r->out0, s->pool_base + r->off0, r->bytes0, hipMemcpyDeviceToHost, s->copy_stream));
        if (r->bytes1)
            G500_HIP(hipMemcpyAsync(r->out1, s->pool_base + r->off1, r->bytes1, hipMemcpyDeviceToHost, s->copy_stream));
        G500_HIP(hipStreamSynchronize(s->copy_stream));
        if (sink.status) {
            fprintf(stderr, "libg500 server: result copy of request %llu failed: %s\n", (unsigned long long)r->id,
                    sink.msg);
            if (!r->status) r->status = kServeHostError;
        }
        const double t_done = now_ms();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->pool.release(r->off0, r->bytes0);
            s->pool.release(r->off1, r->bytes1);
            G500_Completion c;
            c.id = r->id;
            c.graph = r->graph;
            c.kind = r->kind;
            c.root = r->root;
            c.status = r->status;
            c.batch_size = r->batch;
            c.reached = r->reached;
            c.sum = r->sum;
            c.latency_ms = t_done - r->t_submit;
            c.queue_ms = r->t_launch - r->t_submit;
            c.run_ms = r->t_staged - r->t_launch;
            s->done.push_back(c);
            ++s->completed;
            --s->in_system;
            delete r;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();  // pool memory freed: waiting requests may be admitted
    }
}

}  // namespace

extern "C" {

G500_Server *g500_server_create(Worker_State *ws, i32 slots, u64 result_pool_bytes) {
    if (slots < 1 || slots > 8) {
        fail("server slots must be 1..8 (got %d)", slots);
        return nullptr;
    }
    G500_Server *s = new G500_Server;
    s->ws = ws;
    s->nslots = slots;
    s->pool_base = (u8 *)arena_alloc(ws, result_pool_bytes ? result_pool_bytes : 256, "server result pool");
    if (failed()) {
        delete s;
        return nullptr;
    }
    s->pool.init(result_pool_bytes);
    return s;
}

i32 g500_server_add_graph(G500_Server *s, const Oned_Graph *g) {
    if (s->started) {
        fail("graphs are added before g500_server_start");
        return -1;
    }
    s->graphs.push_back(*g);
    return (i32)s->graphs.size() - 1;
}

i32 g500_server_start(G500_Server *s, f32 delta) {
    if (s->started || s->graphs.empty()) {
        fail("g500_server_start needs at least one graph and runs once");
        return -1;
    }
    if (!(delta > 0.0f) || !(delta < 3.0e38f)) {
        fail("SSSP delta must be positive and finite (got %g)", (double)delta);
        return -1;
    }
    s->delta = delta;
    const Oned_Graph *big = &s->graphs[0];
    for (const Oned_Graph &g : s->graphs)
        if (g.nlocalverts > big->nlocalverts) big = &g;
    const int cus = ctx().num_cus;
    const int bfs_per_cu = std::max(1, bfs_batch_grid() / cus), sssp_per_cu = std::max(1, sssp_batch_grid() / cus);
    const bool masked = s->nslots == 2 || s->nslots == 4;
    s->slots.resize(s->nslots);
    for (int k = 0; k < s->nslots; ++k) {
        G500_Server::Slot &slot = s->slots[k];
        if (masked) {
            // the slot's own CUs: mask bits i with (i / 8) % slots == k, an equal
            // share of every XCD (the runtime spreads mask bits round-robin over
            // the XCDs); its launches fill exactly these CUs
            std::vector<u32> mask((cus + 31) / 32, 0u);
            for (int i = 0; i < cus; ++i)
                if ((i / 8) % s->nslots == k) {
                    mask[i / 32] |= 1u << (i % 32);
                    ++slot.cus;
                }
            G500_HIP(hipExtStreamCreateWithCUMask(&slot.stream, (u32)mask.size(), mask.data()));
            slot.grid_bfs = bfs_per_cu * slot.cus;
            slot.grid_sssp = sssp_per_cu * slot.cus;
        } else {
            // an equal share of the co-resident grid for each slot
            G500_HIP(hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking));
            slot.grid_bfs = std::max(1, bfs_per_cu * cus / s->nslots);
            slot.grid_sssp = std::max(1, sssp_per_cu * cus / s->nslots);
        }
        slot.bfs = oned_graph_bfs_batch_create(s->ws, big);
        slot.sssp = oned_graph_sssp_batch_create(s->ws, big);
        slot.summary = (u64 *)arena_alloc(s->ws, 128 * sizeof(u64), "server summaries");
    }
    G500_HIP(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
    if (failed()) return -1;
    // SSSP row splits and row records of every graph for both bucket widths
    const Sssp_Batch_State &proto = s->slots[0].sssp;
    s->splits.resize(s->graphs.size());
    for (size_t gi = 0; gi < s->graphs.size(); ++gi)
        for (int w = 0; w < 2; ++w) {
            const u64 n = s->graphs[gi].nlocalverts;
            Sssp_Batch_State &c = s->splits[gi][w];
            memset(&c, 0, sizeof(c));
            c.max_nlocalverts = n;
            c.delta = w ? 0.5f * delta : delta;
            c.mid_weight = proto.mid_weight;
            c.long_weight = proto.long_weight;
            c.far_weight = proto.far_weight;
            c.split = (u32 *)arena_alloc(s->ws, n * sizeof(u32), "server SSSP light splits");
            c.mid_split = (u32 *)arena_alloc(s->ws, n * sizeof(u32), "server SSSP mid splits");
            c.long_split = (u32 *)arena_alloc(s->ws, n * sizeof(u32), "server SSSP long splits");
            c.far_split = (u32 *)arena_alloc(s->ws, n * sizeof(u32), "server SSSP far splits");
            c.row_info = arena_alloc(s->ws, n * 32, "server SSSP row records");
            if (failed()) return -1;
            sssp_batch_prepare(&s->graphs[gi], &c, ctx().stream, nullptr, nullptr, nullptr);
        }
    sync();
    if (failed()) return -1;
    // hub distances of every graph (host copies) for grouping SSSP requests
    s->hubd.resize(s->graphs.size());
    for (size_t gi = 0; gi < s->graphs.size(); ++gi) {
        const u64 n = s->graphs[gi].nlocalverts;
        const u64 mark = s->ws->gpu_arena.used;
        f32 *dh = (f32 *)arena_alloc(s->ws, n * sizeof(f32), "server hub distances");
        if (!failed() && hub_distances(s->ws, &s->graphs[gi], dh)) {
            s->hubd[gi].resize(n);
            g500_copy_to_host(s->hubd[gi].data(), dh, n * sizeof(f32));
        }
        s->ws->gpu_arena.used = mark;
        if (failed()) return -1;
    }
    s->started = true;
    for (int k = 0; k < s->nslots; ++k) s->slots[k].th = std::thread(slot_loop, s, k);
    s->copier = std::thread(copy_loop, s);
    return 0;
}

u64 g500_server_submit(G500_Server *s, i32 graph, i32 kind, u64 root, void *out0, void *out1) {
    if (!s->started || graph < 0 || graph >= (i32)s->graphs.size() || (kind != G500_BFS && kind != G500_SSSP) ||
        root >= s->graphs[graph].nglobalverts) {
        fail("bad server request (graph %d kind %d root %llu)", graph, kind, (unsigned long long)root);
        return 0;
    }
    const u64 n = s->graphs[graph].nlocalverts;
    const u64 need = (out0 ? (kind == G500_SSSP ? n * sizeof(f32) : n * sizeof(i64)) : 0) +
                     (kind == G500_SSSP && out1 ? n * sizeof(i64) : 0);
    if (need + 512 > s->pool.size() && need) {
        fail("request results (%llu bytes) exceed the server result pool (%llu bytes)", (unsigned long long)need,
             (unsigned long long)s->pool.size());
        return 0;
    }
    Request *r = new Request();
    r->graph = graph;
    r->kind = kind;
    r->root = root;
    r->out0 = out0;
    r->out1 = kind == G500_SSSP ? out1 : nullptr;
    r->off0 = r->off1 = ~0ull;
    r->t_submit = now_ms();
    u64 id;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        id = r->id = s->next_id++;
        s->pending.push_back(r);
        ++s->submitted;
        ++s->in_system;
    }
    s->cv_work.notify_one();
    return id;
}

void g500_server_replay(G500_Server *s, u64 count, const f64 *at_ms, const i32 *graph, const i32 *kind,
                        const u64 *root) {
    const auto t0 = std::chrono::steady_clock::now();
    for (u64 i = 0; i < count; ++i) {
        const auto due = t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                  std::chrono::duration<double, std::milli>(at_ms[i]));
        std::this_thread::sleep_until(due);
        if (!g500_server_submit(s, graph[i], kind[i], root[i], nullptr, nullptr)) return;
    }
}

i32 g500_server_poll(G500_Server *s, G500_Completion *out, i32 max, f64 wait_ms) {
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(wait_ms),
                            [&] { return !s->done.empty(); });
    i32 k = 0;
    while (k < max && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    return k;
}

void g500_server_drain(G500_Server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->in_system == 0; });
}

void g500_server_stats(const G500_Server *s, u64 *submitted, u64 *completed, u64 *launches, u64 *pool_used) {
    G500_Server *m = const_cast<G500_Server *>(s);
    std::lock_guard<std::mutex> lk(m->mu);
    if (submitted) *submitted = s->submitted;
    if (completed) *completed = s->completed;
    if (launches) *launches = s->launches;
    if (pool_used) *pool_used = s->pool.used();
}

void g500_server_destroy(G500_Server *s) {
    if (!s) return;
    if (s->started) {
        g500_server_drain(s);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stopping = true;
        }
        s->cv_work.notify_all();
        s->cv_copy.notify_all();
        for (auto &slot : s->slots)
            if (slot.th.joinable()) slot.th.join();
        if (s->copier.joinable()) s->copier.join();
        for (auto &slot : s->slots) (void)hipStreamDestroy(slot.stream);
        (void)hipStreamDestroy(s->copy_stream);
    }
    for (Request *r : s->pending) delete r;
    delete s;
}

}  // extern "C"
