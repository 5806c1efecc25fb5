"""numpy wrappers around oracle/liboracle.so (the CPU oracle).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class MrgState(C.Structure):
    _fields_ = [("z", C.c_uint32 * 5)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        L.o_mrg_next.argtypes = [C.POINTER(MrgState)]
        L.o_mrg_next.restype = C.c_uint32
        L.o_mrg_skip.argtypes = [C.POINTER(MrgState), C.c_uint64, C.c_uint64, C.c_uint64]
        L.o_mrg_skip_table.argtypes = [u32p]
        L.o_mrg_power.argtypes = [C.c_int, C.c_int, u32p]
        L.o_make_mrg_seed.argtypes = [C.c_uint64, C.c_uint64, u32p]
        L.o_make_random_numbers.argtypes = [C.c_int64, C.c_uint64, C.c_uint64, C.c_int64, f64p]
        L.o_mrg_weight.argtypes = [C.POINTER(MrgState)]
        L.o_mrg_weight.restype = C.c_float
        L.o_scramble_values.argtypes = [u32p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.o_scramble.argtypes = [C.c_uint64, C.c_int, C.c_uint64, C.c_uint64]
        L.o_scramble.restype = C.c_uint64
        L.o_generate_kronecker_range.argtypes = [u32p, C.c_int, C.c_int64, C.c_int64, u64p, f32p]
        L.o_csr_build.argtypes = [C.c_uint64, C.c_uint64, u64p, f32p, u64p, C.c_void_p, C.c_void_p]
        L.o_csr_build.restype = C.c_uint64
        L.o_bfs_serial.argtypes = [C.c_uint64, u64p, u32p, C.c_uint64, i64p, C.c_void_p]
        L.o_bfs_serial.restype = C.c_uint64
        L.o_bfs_omp.argtypes = [C.c_uint64, u64p, u32p, C.c_uint64, i64p, C.c_int]
        L.o_bfs_omp.restype = C.c_uint64
        L.o_dist_from_pred.argtypes = [C.c_uint64, u64p, u32p, i64p, C.c_uint64, i64p]
        L.o_bfs_validate.argtypes = [C.c_uint64, u64p, u32p, i64p, C.c_uint64]
        L.o_bfs_validate.restype = C.c_uint64
        L.o_edge_count.argtypes = [C.c_uint64, u64p, u32p, i64p]
        L.o_edge_count.restype = C.c_uint64
        L.o_generate_roots.argtypes = [C.c_uint64, u64p, C.c_uint64, C.c_uint64, u64p, C.c_int]
        L.o_generate_roots.restype = C.c_int
        L.o_dijkstra_f32.argtypes = [C.c_uint64, u64p, u32p, f32p, C.c_uint64, f32p, i64p]
        L.o_dijkstra_f32_timed.argtypes = [C.c_uint64, u64p, u32p, f32p, C.c_uint64, C.c_double,
                                           C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                           C.POINTER(C.c_double)]
        L.o_dijkstra_f64.argtypes = [C.c_uint64, u64p, u32p, f32p, C.c_uint64, f64p]
        L.o_delta_stepping_f32.argtypes = [C.c_uint64, u64p, u32p, f32p, C.c_uint64, C.c_float,
                                           f32p, i64p]
        L.o_sssp_check.argtypes = [C.c_uint64, u64p, u32p, f32p, C.c_uint64, f32p, i64p]
        L.o_sssp_check.restype = C.c_uint64
        L.o_statistics.argtypes = [f64p, C.c_int, f64p]
        L.o_bfs_depth_omp.argtypes = [C.c_uint64, u64p, u32p, C.c_uint64, i64p, C.c_int]
        L.o_bfs_depth_omp.restype = C.c_uint64
        L.o_tuple_digest.argtypes = [C.c_uint64, u64p, f32p, C.POINTER(C.c_uint64)]
        L.o_tuple_digest.restype = C.c_uint64
        L.o_csr_digest.argtypes = [C.c_uint64, u64p, u32p, f32p]
        L.o_csr_digest.restype = C.c_uint64
        L.o_depth_check.argtypes = [C.c_uint64, u64p, u32p, i64p, C.c_uint64]
        L.o_depth_check.restype = C.c_uint64
        L.o_degree_add.argtypes = [C.c_uint64, C.c_uint64, u64p, u64p]
        L.o_degree_add.restype = C.c_uint64
        L.o_csr_order_errors.argtypes = [C.c_uint64, u64p, u32p, f32p]
        L.o_csr_order_errors.restype = C.c_uint64
        L.o_delta_stepping_omp.argtypes = [C.c_uint64, u64p, u32p, f32p, C.c_uint64, C.c_float, f32p, i64p,
                                           C.c_int]
        _LIB = L
    return _LIB


def mrg_seed(s1=2, s2=3):
    out = np.zeros(5, np.uint32)
    lib().o_make_mrg_seed(s1, s2, out)
    return out


def mrg_state(seed):
    st = MrgState()
    for i in range(5):
        st.z[i] = int(seed[i])
    return st


def skip_table():
    out = np.zeros(24 * 256 * 9, np.uint32)
    lib().o_mrg_skip_table(out)
    return out.reshape(24, 256, 9)


def mrg_power(b, v):
    out = np.zeros(25, np.uint32)
    lib().o_mrg_power(b, v, out)
    return out.reshape(5, 5)


def random_numbers(nvalues, s1, s2, pos):
    out = np.zeros(nvalues, np.float64)
    lib().o_make_random_numbers(nvalues, s1, s2, pos, out)
    return out


def scramble_values(seed):
    a, b = C.c_uint64(), C.c_uint64()
    lib().o_scramble_values(np.ascontiguousarray(seed, np.uint32), C.byref(a), C.byref(b))
    return a.value, b.value


def kronecker(scale, edgefactor=16, s1=2, s2=3, start=0, end=None):
    """Tuple list (edges[E,2] u64, weights[E] f32) for edge indices [start, end)."""
    if end is None:
        end = edgefactor << scale
    cnt = end - start
    edges = np.zeros(2 * cnt, np.uint64)
    w = np.zeros(cnt, np.float32)
    lib().o_generate_kronecker_range(mrg_seed(s1, s2), scale, start, end, edges, w)
    return edges.reshape(cnt, 2), w


class Csr:
    def __init__(self, n, rowstarts, column, weights):
        self.n = n
        self.rowstarts = rowstarts
        self.column = column
        self.weights = weights

    @property
    def m(self):
        return int(self.rowstarts[-1])


def csr(n, edges, weights):
    edges = np.ascontiguousarray(edges, np.uint64).reshape(-1)
    weights = np.ascontiguousarray(weights, np.float32)
    ne = weights.shape[0]
    rs = np.zeros(n + 1, np.uint64)
    m = lib().o_csr_build(n, ne, edges, weights, rs, None, None)
    col = np.zeros(max(m, 1), np.uint32)
    cw = np.zeros(max(m, 1), np.float32)
    lib().o_csr_build(n, ne, edges, weights, rs, col.ctypes.data, cw.ctypes.data)
    return Csr(n, rs, col[:m], cw[:m])


def _c(a, dt):
    return np.ascontiguousarray(a, dt)


def bfs(g, root, with_depth=True):
    pred = np.zeros(g.n, np.int64)
    depth = np.zeros(g.n, np.int64) if with_depth else None
    lib().o_bfs_serial(g.n, g.rowstarts, _c(g.column, np.uint32), root, pred,
                       depth.ctypes.data if with_depth else None)
    return pred, depth


def bfs_omp(g, root, nthreads):
    pred = np.zeros(g.n, np.int64)
    lib().o_bfs_omp(g.n, g.rowstarts, _c(g.column, np.uint32), root, pred, nthreads)
    return pred


def bfs_depth_omp(g, root, nthreads=16):
    """Hop counts (-1 unreached) from a parallel level-synchronous BFS."""
    depth = np.zeros(g.n, np.int64)
    lib().o_bfs_depth_omp(g.n, g.rowstarts, _c(g.column, np.uint32), root, depth, nthreads)
    return depth


def tuple_digest(edges, weights):
    """(digest, entries) of the directed entries a tuple list stands for."""
    edges = np.ascontiguousarray(edges, np.uint64).reshape(-1)
    weights = np.ascontiguousarray(weights, np.float32)
    cnt = C.c_uint64()
    d = lib().o_tuple_digest(weights.shape[0], edges, weights, C.byref(cnt))
    return d, cnt.value


def csr_digest(g):
    return lib().o_csr_digest(g.n, g.rowstarts, _c(g.column, np.uint32), _c(g.weights, np.float32))


def degree_add(n, edges, deg):
    """Adds the degrees of a chunk of tuples into deg (u64[n]); returns the
    edges with an endpoint outside [0, n)."""
    edges = np.ascontiguousarray(edges, np.uint64).reshape(-1)
    return int(lib().o_degree_add(n, edges.shape[0] // 2, edges, deg))


def csr_order_errors(g):
    return int(lib().o_csr_order_errors(g.n, g.rowstarts, _c(g.column, np.uint32), _c(g.weights, np.float32)))


def rowstarts_of(n, edges):
    """Row offsets of the CSR a tuple list gives (degree prefix sum, loops excluded)."""
    edges = np.ascontiguousarray(edges, np.uint64).reshape(-1)
    rs = np.zeros(n + 1, np.uint64)
    lib().o_csr_build(n, edges.shape[0] // 2, edges, np.zeros(1, np.float32), rs, None, None)
    return rs


def delta_stepping_omp(g, root, delta=0.1, nthreads=16):
    dist = np.zeros(g.n, np.float32)
    pred = np.zeros(g.n, np.int64)
    lib().o_delta_stepping_omp(g.n, g.rowstarts, _c(g.column, np.uint32), _c(g.weights, np.float32), root,
                               delta, dist, pred, nthreads)
    return dist, pred


def dist_from_pred(g, pred, root):
    dist = np.zeros(g.n, np.int64)
    lib().o_dist_from_pred(g.n, g.rowstarts, _c(g.column, np.uint32), _c(pred, np.int64), root, dist)
    return dist


def bfs_validate(g, pred, root):
    return int(lib().o_bfs_validate(g.n, g.rowstarts, _c(g.column, np.uint32), _c(pred, np.int64), root))


def edge_count(g, pred):
    return int(lib().o_edge_count(g.n, g.rowstarts, _c(g.column, np.uint32), _c(pred, np.int64)))


def roots(g, s1=2, s2=3, max_roots=64):
    out = np.zeros(max_roots, np.uint64)
    k = lib().o_generate_roots(g.n, g.rowstarts, s1, s2, out, max_roots)
    return out[:k]


def dijkstra_f32(g, root):
    dist = np.zeros(g.n, np.float32)
    pred = np.zeros(g.n, np.int64)
    lib().o_dijkstra_f32(g.n, g.rowstarts, _c(g.column, np.uint32), _c(g.weights, np.float32), root,
                         dist, pred)
    return dist, pred


def dijkstra_f32_timed(g, root, max_seconds):
    """(entries scanned, vertices settled, seconds) of a time-bounded Dijkstra."""
    sc, st, el = C.c_uint64(), C.c_uint64(), C.c_double()
    lib().o_dijkstra_f32_timed(g.n, g.rowstarts, _c(g.column, np.uint32), _c(g.weights, np.float32), root,
                               max_seconds, C.byref(sc), C.byref(st), C.byref(el))
    return sc.value, st.value, el.value


def dijkstra_f64(g, root):
    dist = np.zeros(g.n, np.float64)
    lib().o_dijkstra_f64(g.n, g.rowstarts, _c(g.column, np.uint32), _c(g.weights, np.float32), root,
                         dist)
    return dist


def delta_stepping_f32(g, root, delta=0.1):
    dist = np.zeros(g.n, np.float32)
    pred = np.zeros(g.n, np.int64)
    lib().o_delta_stepping_f32(g.n, g.rowstarts, _c(g.column, np.uint32), _c(g.weights, np.float32),
                               root, delta, dist, pred)
    return dist, pred


def sssp_check(g, root, dist, pred):
    return int(lib().o_sssp_check(g.n, g.rowstarts, _c(g.column, np.uint32), _c(g.weights, np.float32),
                                  root, _c(dist, np.float32), _c(pred, np.int64)))


def depth_check(g, depth, root):
    return int(lib().o_depth_check(g.n, g.rowstarts, _c(g.column, np.uint32), _c(depth, np.int64), root))


def statistics(x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros(7, np.float64)
    lib().o_statistics(x, len(x), out)
    return out
