"""ctypes binding of libg500.so, the MI355X Graph500 library.

The C-ABI (include/*.h) is the product boundary; this module mirrors its
structs and declares its entry points so that tests and the benchmark can call
the library exactly as a C caller would. There is no Python compute path: if
libg500.so is missing, load() raises.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libg500.so")
RUNNER_PATH = os.path.join(PKG_DIR, "graph500_runner")

u64 = C.c_uint64
i64 = C.c_int64
u32 = C.c_uint32
i32 = C.c_int32
f32 = C.c_float
f64 = C.c_double
vp = C.c_void_p


class Arena(C.Structure):
    _fields_ = [("reserved", u64), ("used", u64), ("base", vp)]


class WorkerState(C.Structure):
    _fields_ = [("cpu_arena", Arena), ("gpu_arena", Arena), ("rank", i32), ("rank_size", i32),
                ("pes_on_device", i32)]


class TupleGraph(C.Structure):
    _fields_ = [("edges", vp), ("edges_size", u64), ("max_edges_size", u64), ("nglobaledges", u64),
                ("weights", vp)]


class OnedGraph(C.Structure):
    _fields_ = [("nlocalverts", u64), ("nlocaledges", u64), ("nglobalverts", u64), ("notisolated", u64),
                ("rowstarts", vp), ("column", vp), ("weights", vp), ("perm", vp), ("iperm", vp),
                ("rowstarts_t", vp), ("column_t", vp), ("weights_t", vp),
                ("rowstarts_s", vp), ("column_s", vp), ("weights_s", vp)]


class BfsState(C.Structure):
    _fields_ = [("pred", vp), ("dist", vp), ("frontier", vp * 2), ("frontier_bits", vp * 2),
                ("visited", vp), ("hubs", vp), ("hub_offsets", vp), ("hub_lengths", vp), ("ctl", vp),
                ("max_nlocalverts", u64)]


class SsspState(C.Structure):
    _fields_ = [("pred", vp), ("dist", vp), ("dist_pred", vp), ("frontier", vp), ("heavy", vp),
                ("hub_offsets", vp), ("hub_lengths", vp), ("split", vp), ("rows", vp), ("stamp", vp), ("settled", vp),
                ("reached", vp),
                ("ctl", vp), ("split_graph", vp), ("delta", f32), ("split_delta", f32), ("max_nlocalverts", u64)]


class BfsBatchState(C.Structure):
    _fields_ = [("seen", vp), ("fmask", vp * 2), ("fbits", vp * 2), ("next", vp), ("touched", vp), ("log_vertex", vp),
                ("log_mask", vp), ("level_start", vp), ("hubs", vp), ("hub_offsets", vp), ("hub_lengths", vp),
                ("ctl", vp), ("log_capacity", u64), ("max_nlocalverts", u64), ("nroots", i32), ("levels", i32)]


class SsspBatchState(C.Structure):
    _fields_ = [("dist", vp), ("pred", vp), ("labels", vp), ("pending", vp), ("frontier", vp), ("heavy", vp),
                ("hub_vertices", vp), ("hub_offsets", vp), ("hub_lengths", vp), ("hub_aux", vp), ("split", vp),
                ("root_masks", vp), ("open_words", vp), ("mid_split", vp), ("mid_work", vp), ("mid_offsets", vp),
                ("far_split", vp), ("far_work", vp), ("far_offsets", vp),
                ("long_split", vp), ("long_work", vp), ("long_offsets", vp),
                ("ctl", vp), ("split_graph", vp), ("delta", f32), ("split_delta", f32), ("long_weight", f32),
                ("long_split_weight", f32), ("mid_weight", f32), ("mid_split_weight", f32), ("far_weight", f32),
                ("far_split_weight", f32), ("nroots", i32), ("reserved", i32),
                ("max_nlocalverts", u64), ("row_info", vp)]


class Completion(C.Structure):
    _fields_ = [("id", u64), ("graph", i32), ("kind", i32), ("root", u64), ("status", i32), ("batch_size", i32),
                ("reached", u64), ("sum", f64), ("latency_ms", f64), ("queue_ms", f64), ("run_ms", f64)]


BFS, SSSP = 0, 1


P = C.POINTER
_SIGS = {
    # base.h
    "mem_align_forward": (u64, [u64, u64]),
    "arena_create_cpu": (Arena, [u64]),
    "arena_create_gpu": (Arena, [u64]),
    "arena_from_arena": (Arena, [P(Arena), u64]),
    "arena_release_cpu": (None, [P(Arena)]),
    "arena_release_gpu": (None, [P(Arena)]),
    "_arena_push": (vp, [P(Arena), u64, u64]),
    "arena_clear": (None, [P(Arena)]),
    # worker.h
    "worker_init": (WorkerState, [i32, vp, u64]),
    "worker_finalize": (None, [P(WorkerState)]),
    "worker_abort": (None, [P(WorkerState), i32]),
    "_worker_arena_push_array": (vp, [P(WorkerState), u64, u64, u64, P(u64)]),
    # graph_generation.h
    "generate_kronecker_range": (None, [P(WorkerState), P(u32), u32, i64, i64, P(TupleGraph)]),
    "make_mrg_seed": (None, [u64, u64, P(u32)]),
    "make_random_numbers": (None, [i64, u64, u64, i64, P(f64)]),
    "tuple_graph_dump": (None, [P(TupleGraph), i64]),
    "tuple_graph_load": (None, [P(WorkerState), P(TupleGraph), i64, i64]),
    # traversal.h
    "oned_graph_from_tuple_graph": (OnedGraph, [P(WorkerState), P(TupleGraph), u64]),
    "oned_graph_is_vertex_isolated": (i32, [P(WorkerState), P(OnedGraph), u64]),
    "oned_graph_bfs_create": (BfsState, [P(WorkerState), P(OnedGraph)]),
    "oned_graph_bfs_clear": (None, [P(WorkerState), P(BfsState)]),
    "oned_graph_bfs_run": (None, [P(WorkerState), P(OnedGraph), P(BfsState), u64]),
    "bfs_compute_dist_from_pred": (vp, [P(WorkerState), P(OnedGraph), P(BfsState), u64]),
    "oned_graph_bfs_destroy": (None, [P(BfsState)]),
    "oned_graph_sssp_create": (SsspState, [P(WorkerState), P(OnedGraph)]),
    "oned_graph_sssp_clear": (None, [P(WorkerState), P(SsspState)]),
    "oned_graph_sssp_run": (None, [P(WorkerState), P(OnedGraph), P(SsspState), u64]),
    "oned_graph_sssp_destroy": (None, [P(SsspState)]),
    "oned_graph_free": (None, [P(OnedGraph)]),
    # g500_ext.h
    "generate_bfs_roots": (i32, [P(WorkerState), P(OnedGraph), u64, u64, P(u64), i32]),
    "oned_graph_edge_count": (u64, [P(WorkerState), P(OnedGraph), vp]),
    "oned_graph_bfs_validate": (i32, [P(WorkerState), P(OnedGraph), P(BfsState), u64, P(u64)]),
    "oned_graph_bfs_validate_noalloc": (i32, [P(OnedGraph), vp, u64, vp, P(u64)]),
    "oned_graph_bfs_depth_check": (u64, [P(WorkerState), P(OnedGraph), vp, u64]),
    "oned_graph_sssp_check": (u64, [P(WorkerState), P(OnedGraph), vp, vp, u64]),
    "get_statistics": (None, [P(f64), i32, P(f64)]),
    "oned_graph_bfs_batch_create": (BfsBatchState, [P(WorkerState), P(OnedGraph)]),
    "oned_graph_bfs_batch_run": (None, [P(WorkerState), P(OnedGraph), P(BfsBatchState), P(u64), i32]),
    "oned_graph_bfs_batch_depth": (None, [P(WorkerState), P(OnedGraph), P(BfsBatchState), i32, vp]),
    "oned_graph_bfs_batch_destroy": (None, [P(BfsBatchState)]),
    "g500_bfs_batch_level_log": (i32, [P(BfsBatchState), P(i32), P(u64), P(u64), P(f64), i32]),
    "oned_graph_sssp_root_order": (i32, [P(WorkerState), P(OnedGraph), P(u64), i32, P(i32)]),
    "oned_graph_sssp_hub_distances": (i32, [P(WorkerState), P(OnedGraph), P(u64), i32, P(f32)]),
    "oned_graph_sssp_batch_create": (SsspBatchState, [P(WorkerState), P(OnedGraph)]),
    "oned_graph_sssp_batch_run": (None, [P(WorkerState), P(OnedGraph), P(SsspBatchState), P(u64), i32]),
    "oned_graph_sssp_batch_run_offsets": (None, [P(WorkerState), P(OnedGraph), P(SsspBatchState), P(u64), P(f32),
                                                 i32]),
    "oned_graph_sssp_batch_destroy": (None, [P(SsspBatchState)]),
    "g500_sssp_batch_run_stats": (None, [P(SsspBatchState), P(u64), P(u64), P(u64)]),
    "g500_sssp_batch_phase_ms": (None, [P(SsspBatchState), P(f64)]),
    "g500_sssp_batch_work": (None, [P(SsspBatchState), P(u64)]),
    "g500_sssp_batch_pull_work": (None, [P(SsspBatchState), P(u64)]),
    "g500_sssp_batch_step_log": (i32, [P(SsspBatchState), P(i32), P(u64), P(f64), i32]),
    "g500_server_create": (vp, [P(WorkerState), i32, u64]),
    "g500_server_add_graph": (i32, [vp, P(OnedGraph)]),
    "g500_server_start": (i32, [vp, f32]),
    "g500_server_submit": (u64, [vp, i32, i32, u64, vp, vp]),
    "g500_server_replay": (None, [vp, u64, P(f64), P(i32), P(i32), P(u64)]),
    "g500_server_poll": (i32, [vp, P(Completion), i32, f64]),
    "g500_server_drain": (None, [vp]),
    "g500_server_stats": (None, [vp, P(u64), P(u64), P(u64), P(u64)]),
    "g500_server_destroy": (None, [vp]),
    "g500_set_error_mode": (None, [i32]),
    "g500_status": (i32, []),
    "g500_last_error": (C.c_char_p, []),
    "g500_clear_error": (None, []),
    "g500_last_run_ms": (f64, []),
    "g500_copy_to_host": (None, [vp, vp, u64]),
    "g500_copy_to_device": (None, [vp, vp, u64]),
    "g500_copy_device": (None, [vp, vp, u64]),
    "g500_set_param": (i32, [C.c_char_p, f64]),
    "g500_get_param": (i32, [C.c_char_p, P(f64)]),
    "g500_bfs_level_log": (i32, [P(BfsState), P(i32), P(u64), P(f64), i32]),
    "g500_sssp_run_stats": (None, [P(SsspState), P(u64), P(u64), P(u64)]),
    "g500_sssp_phase_ms": (None, [P(SsspState), P(f64)]),
    "g500_sssp_step_log": (i32, [P(SsspState), P(i32), P(u64), P(u64), P(f64), i32]),
}

_LIB = None


def load():
    """Loads libg500.so (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def exported_names():
    return sorted(_SIGS)


class LibraryError(RuntimeError):
    pass


def _check(lib):
    if lib.g500_status():
        msg = lib.g500_last_error().decode()
        lib.g500_clear_error()
        raise LibraryError(msg)


def _ptr(a):
    return a.ctypes.data_as(vp)


class Session:
    """One worker (device 0) with its device arena. Errors raise LibraryError."""

    def __init__(self, arena_bytes=0):
        self.lib = load()
        self.lib.g500_set_error_mode(1)
        self.ws = self.lib.worker_init(0, None, arena_bytes)
        _check(self.lib)

    # ---- helpers
    def check(self):
        _check(self.lib)

    def to_host(self, dev_ptr, count, dtype):
        out = np.empty(count, dtype)
        if count:
            self.lib.g500_copy_to_host(_ptr(out), dev_ptr, out.nbytes)
            self.check()
        return out

    def to_device(self, dev_ptr, arr):
        arr = np.ascontiguousarray(arr)
        if arr.size:
            self.lib.g500_copy_to_device(dev_ptr, _ptr(arr), arr.nbytes)
            self.check()

    def mark(self):
        return self.ws.gpu_arena.used

    def release(self, mark):
        self.ws.gpu_arena.used = mark

    # ---- pipeline
    def generate(self, scale, edgefactor=16, s1=2, s2=3, start=0, end=None):
        seed = (u32 * 5)()
        self.lib.make_mrg_seed(s1, s2, seed)
        tg = TupleGraph()
        tg.nglobaledges = edgefactor << scale
        if end is None:
            end = tg.nglobaledges
        self.lib.generate_kronecker_range(C.byref(self.ws), seed, scale, start, end, C.byref(tg))
        self.check()
        return tg

    def tuples_to_host(self, tg):
        e = self.to_host(tg.edges, 2 * tg.edges_size, np.uint64).reshape(-1, 2)
        w = self.to_host(tg.weights, tg.edges_size, np.float32)
        return e, w

    def upload_tuples(self, edges, weights, nglobaledges=None):
        """Tuple graph in device memory from host arrays (for CSR tests on
        hand-made inputs)."""
        edges = np.ascontiguousarray(edges, np.uint64).reshape(-1, 2)
        weights = np.ascontiguousarray(weights, np.float32)
        tg = TupleGraph()
        ne = weights.shape[0]
        tg.edges_size = tg.max_edges_size = ne
        tg.nglobaledges = ne if nglobaledges is None else nglobaledges
        cnt = u64()
        tg.edges = self.lib._worker_arena_push_array(C.byref(self.ws), ne, 16, 16, C.byref(cnt))
        tg.weights = self.lib._worker_arena_push_array(C.byref(self.ws), ne, 4, 4, C.byref(cnt))
        if ne and (not tg.edges or not tg.weights):
            raise LibraryError("device arena exhausted")
        self.to_device(tg.edges, edges)
        self.to_device(tg.weights, weights)
        return tg

    def csr(self, tg, n):
        g = self.lib.oned_graph_from_tuple_graph(C.byref(self.ws), C.byref(tg), n)
        self.check()
        return g

    def csr_to_host(self, g):
        rs = self.to_host(g.rowstarts, g.nlocalverts + 1, np.uint64)
        col = self.to_host(g.column, g.nlocaledges, np.uint32)
        w = self.to_host(g.weights, g.nlocaledges, np.float32)
        return rs, col, w

    def roots(self, g, s1=2, s2=3, max_roots=64):
        out = (u64 * max_roots)()
        k = self.lib.generate_bfs_roots(C.byref(self.ws), C.byref(g), s1, s2, out, max_roots)
        self.check()
        return np.array(out[:k], np.uint64)

    def bfs_create(self, g):
        b = self.lib.oned_graph_bfs_create(C.byref(self.ws), C.byref(g))
        self.check()
        return b

    def bfs(self, g, b, root):
        self.lib.oned_graph_bfs_clear(C.byref(self.ws), C.byref(b))
        self.lib.oned_graph_bfs_run(C.byref(self.ws), C.byref(g), C.byref(b), int(root))
        self.check()
        return self.lib.g500_last_run_ms()

    def bfs_pred(self, b):
        return self.to_host(b.pred, b.max_nlocalverts, np.int64)

    def bfs_depth(self, g, b, root):
        p = self.lib.bfs_compute_dist_from_pred(C.byref(self.ws), C.byref(g), C.byref(b), int(root))
        self.check()
        return self.to_host(p, b.max_nlocalverts, np.int64)

    def bfs_validate(self, g, b, root):
        err = u64()
        ok = self.lib.oned_graph_bfs_validate(C.byref(self.ws), C.byref(g), C.byref(b), int(root), C.byref(err))
        self.check()
        return bool(ok), err.value

    def level_log(self, b, max_levels=256, with_ms=False):
        modes = (i32 * max_levels)()
        sizes = (u64 * max_levels)()
        ms = (f64 * max_levels)()
        n = self.lib.g500_bfs_level_log(C.byref(b), modes, sizes, ms, max_levels)
        if with_ms:
            return list(modes[:n]), list(sizes[:n]), list(ms[:n])
        return list(modes[:n]), list(sizes[:n])

    def batch_create(self, g):
        b = self.lib.oned_graph_bfs_batch_create(C.byref(self.ws), C.byref(g))
        self.check()
        return b

    def batch(self, g, b, roots):
        r = np.ascontiguousarray(roots, np.uint64)
        self.lib.oned_graph_bfs_batch_run(C.byref(self.ws), C.byref(g), C.byref(b),
                                          r.ctypes.data_as(P(u64)), int(r.size))
        self.check()
        return self.lib.g500_last_run_ms()

    def batch_depth(self, g, b, r):
        """Depths from root index r of the last batched run (host array)."""
        n = g.nlocalverts
        if getattr(self, "_depth_buf", (None, 0))[1] != n:
            cnt = u64()
            p = self.lib._worker_arena_push_array(C.byref(self.ws), n, 8, 8, C.byref(cnt))
            if not p:
                raise LibraryError("device arena exhausted")
            self._depth_buf = (p, n)
        p = self._depth_buf[0]
        self.lib.oned_graph_bfs_batch_depth(C.byref(self.ws), C.byref(g), C.byref(b), int(r), p)
        self.check()
        return self.to_host(p, n, np.int64)

    def batch_level_log(self, b, max_levels=256, with_edges=False):
        modes = (i32 * max_levels)()
        recs = (u64 * max_levels)()
        edges = (u64 * max_levels)()
        ms = (f64 * max_levels)()
        n = self.lib.g500_bfs_batch_level_log(C.byref(b), modes, recs, edges, ms, max_levels)
        if with_edges:
            return list(modes[:n]), list(recs[:n]), list(edges[:n]), list(ms[:n])
        return list(modes[:n]), list(recs[:n]), list(ms[:n])

    def sssp_root_order(self, g, roots):
        """Permutation of roots sorted by distance from the highest-degree vertex."""
        r = np.ascontiguousarray(roots, np.uint64)
        out = np.zeros(r.size, np.int32)
        self.lib.oned_graph_sssp_root_order(C.byref(self.ws), C.byref(g), r.ctypes.data_as(P(u64)), int(r.size),
                                            out.ctypes.data_as(P(i32)))
        self.check()
        return out

    def sssp_hub_distances(self, g, roots):
        """FP32 distances of the roots from the highest-degree vertex (-1 unreached)."""
        r = np.ascontiguousarray(roots, np.uint64)
        out = np.zeros(r.size, np.float32)
        self.lib.oned_graph_sssp_hub_distances(C.byref(self.ws), C.byref(g), r.ctypes.data_as(P(u64)), int(r.size),
                                               out.ctypes.data_as(P(f32)))
        self.check()
        return out

    def sssp_batch_create(self, g, delta=0.1):
        s = self.lib.oned_graph_sssp_batch_create(C.byref(self.ws), C.byref(g))
        self.check()
        s.delta = delta
        return s

    def sssp_batch(self, g, s, roots, offsets=None):
        """Batched SSSP from up to 32 roots; `offsets` (one per root, e.g. the
        roots' hub distances) set per-root bucket lags."""
        r = np.ascontiguousarray(roots, np.uint64)
        if offsets is None:
            self.lib.oned_graph_sssp_batch_run(C.byref(self.ws), C.byref(g), C.byref(s), r.ctypes.data_as(P(u64)),
                                               int(r.size))
        else:
            o = np.ascontiguousarray(offsets, np.float32)
            if o.size != r.size:
                raise ValueError("one offset per root")
            self.lib.oned_graph_sssp_batch_run_offsets(C.byref(self.ws), C.byref(g), C.byref(s),
                                                       r.ctypes.data_as(P(u64)), o.ctypes.data_as(P(f32)),
                                                       int(r.size))
        self.check()
        return self.lib.g500_last_run_ms()

    def sssp_batch_result(self, g, s, i):
        """Distances and parents from root index i of the last batched run."""
        n = g.nlocalverts
        return (self.to_host(s.dist + 4 * n * i, n, np.float32),
                self.to_host(s.pred + 8 * n * i, n, np.int64))

    def sssp_batch_phase_ms(self, s):
        out = (f64 * 9)()
        self.lib.g500_sssp_batch_phase_ms(C.byref(s), out)
        return list(out)

    def sssp_batch_work(self, s):
        out = (u64 * 18)()
        self.lib.g500_sssp_batch_work(C.byref(s), out)
        return {k: tuple(int(x) for x in out[3 * i:3 * i + 3])
                for i, k in enumerate(("light", "medium", "scan", "mid", "long", "far"))}

    def sssp_batch_pull_work(self, s):
        out = (u64 * 4)()
        self.lib.g500_sssp_batch_pull_work(C.byref(s), out)
        return dict(zip(("pulls", "live_pairs", "scan_steps", "long_vertices"), (int(x) for x in out)))

    def sssp_batch_stats(self, s):
        b, li, r = u64(), u64(), u64()
        self.lib.g500_sssp_batch_run_stats(C.byref(s), C.byref(b), C.byref(li), C.byref(r))
        return b.value, li.value, r.value

    def sssp_batch_step_log(self, s, max_steps=4096):
        kind = (i32 * max_steps)()
        items = (u64 * max_steps)()
        ms = (f64 * max_steps)()
        k = self.lib.g500_sssp_batch_step_log(C.byref(s), kind, items, ms, max_steps)
        return list(kind[:k]), list(items[:k]), list(ms[:k])

    def sssp_step_log(self, s, max_steps=1023):
        kind = (i32 * max_steps)()
        items = (u64 * max_steps)()
        hubs = (u64 * max_steps)()
        ms = (f64 * max_steps)()
        k = self.lib.g500_sssp_step_log(C.byref(s), kind, items, hubs, ms, max_steps)
        return list(kind[:k]), list(items[:k]), list(hubs[:k]), list(ms[:k])

    def sssp_phase_ms(self, s):
        out = (f64 * 5)()
        self.lib.g500_sssp_phase_ms(C.byref(s), out)
        return list(out)

    def depth_check(self, g, depth_dev, root):
        """Violations of the BFS depth certificate (device depths)."""
        c = self.lib.oned_graph_bfs_depth_check(C.byref(self.ws), C.byref(g), depth_dev, int(root))
        self.check()
        return c

    def sssp_check(self, g, dist_dev, pred_dev, root):
        """Violations of the SSSP certificate (device distances and parents)."""
        c = self.lib.oned_graph_sssp_check(C.byref(self.ws), C.byref(g), dist_dev, pred_dev, int(root))
        self.check()
        return c

    def edge_count(self, g, pred_dev):
        c = self.lib.oned_graph_edge_count(C.byref(self.ws), C.byref(g), pred_dev)
        self.check()
        return c

    def sssp_create(self, g, delta=0.1):
        s = self.lib.oned_graph_sssp_create(C.byref(self.ws), C.byref(g))
        self.check()
        s.delta = delta
        return s

    def sssp(self, g, s, root):
        self.lib.oned_graph_sssp_clear(C.byref(self.ws), C.byref(s))
        self.lib.oned_graph_sssp_run(C.byref(self.ws), C.byref(g), C.byref(s), int(root))
        self.check()
        return self.lib.g500_last_run_ms()

    def sssp_result(self, s):
        return (self.to_host(s.dist, s.max_nlocalverts, np.float32),
                self.to_host(s.pred, s.max_nlocalverts, np.int64))

    def sssp_stats(self, s):
        b, li, r = u64(), u64(), u64()
        self.lib.g500_sssp_run_stats(C.byref(s), C.byref(b), C.byref(li), C.byref(r))
        return b.value, li.value, r.value

    # ---- serving
    def server(self, graphs, slots=2, pool_bytes=1 << 30, delta=0.05):
        """Starts a server over resident graphs (OnedGraph structs); returns its handle.
        delta 0.05 (0.025 for batches above 16 roots) gave the most requests per
        second and the lowest p99 over 8 graphs of SCALE 16-24 at 1k-50k req/s."""
        h = self.lib.g500_server_create(C.byref(self.ws), slots, pool_bytes)
        self.check()
        for g in graphs:
            self.lib.g500_server_add_graph(h, C.byref(g))
            self.check()
        self.lib.g500_server_start(h, delta)
        self.check()
        return h

    def server_submit(self, h, graph, kind, root, out0=None, out1=None):
        rid = self.lib.g500_server_submit(h, graph, kind, int(root), None if out0 is None else _ptr(out0),
                                          None if out1 is None else _ptr(out1))
        self.check()
        return rid

    def server_poll(self, h, max_items=4096, wait_ms=0.0):
        buf = (Completion * max_items)()
        k = self.lib.g500_server_poll(h, buf, max_items, wait_ms)
        return [buf[i] for i in range(k)]

    def server_stats(self, h):
        a, b, c, d = u64(), u64(), u64(), u64()
        self.lib.g500_server_stats(h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return {"submitted": a.value, "completed": b.value, "launches": c.value, "pool_used": d.value}

    def set_param(self, name, value):
        if self.lib.g500_set_param(name.encode(), float(value)):
            raise KeyError(name)

    def get_param(self, name):
        v = f64()
        if self.lib.g500_get_param(name.encode(), C.byref(v)):
            raise KeyError(name)
        return v.value

    def close(self):
        self.lib.worker_finalize(C.byref(self.ws))
