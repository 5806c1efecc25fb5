"""Serving workload driver over the library's request loop (g500_server_*).

Resident graphs: Kronecker graphs of SCALE 16, 18, 20, 22, 24 with user seed
(2, 3) plus SCALE 16, 18, 20 with user seed (5, 7), edgefactor 16. Requests are
(graph uniform, root uniform over the graph's non-isolated vertices,
BFS / SSSP 50/50) with Poisson arrivals from a fixed numpy seed, submitted at
their arrival times by the library's native replay loop (no result buffers:
each completion carries the reached count and depth / distance sum).

TEPS per request: m_c of the root's component. Every giant-component root has
the same m_c (measured once per graph with the single-root BFS and the TEPS
edge count); a request whose reached count differs from the giant component's
is counted with reached - 1 edges (a lower bound; such components are tiny).
"""
import ctypes as C

import numpy as np

from . import BFS, SSSP, P, f64, i32, u64

GRAPHS = [(16, 2, 3), (18, 2, 3), (20, 2, 3), (22, 2, 3), (24, 2, 3), (16, 5, 7), (18, 5, 7), (20, 5, 7)]


class Resident:
    """A graph held in HBM for serving, with its root pool and giant-component size."""

    def __init__(self, sess, scale, s1, s2, edgefactor=16):
        self.scale, self.seed = scale, (s1, s2)
        tg = sess.generate(scale, edgefactor, s1, s2)
        self.g = sess.csr(tg, 1 << scale)
        deg = np.diff(sess.to_host(self.g.rowstarts, self.g.nlocalverts + 1, np.uint64).astype(np.int64))
        self.candidates = np.flatnonzero(deg > 0).astype(np.uint64)
        # giant component: the component of the highest-degree vertex
        b = sess.bfs_create(self.g)
        sess.bfs(self.g, b, int(np.argmax(deg)))
        self.m_giant = sess.edge_count(self.g, b.pred)
        self.n_giant = int(sum(sess.level_log(b)[1]))


def load_graphs(sess, spec=GRAPHS, edgefactor=16):
    return [Resident(sess, sc, s1, s2, edgefactor) for sc, s1, s2 in spec]


def poisson_run(sess, graphs, rate, count, slots=4, delta=0.05, seed=12345):
    """Submits `count` requests at `rate` per second; returns latency and
    throughput figures of the completed requests."""
    rng = np.random.default_rng(seed)
    at = np.cumsum(rng.exponential(1e3 / rate, count))
    gi = rng.integers(0, len(graphs), count).astype(np.int32)
    kind = np.where(rng.random(count) < 0.5, BFS, SSSP).astype(np.int32)
    roots = np.array([graphs[k].candidates[rng.integers(0, graphs[k].candidates.size)] for k in gi], np.uint64)
    h = sess.server([r.g for r in graphs], slots=slots, pool_bytes=1 << 20, delta=delta)
    lib = sess.lib
    try:
        import time
        t0 = time.perf_counter()
        lib.g500_server_replay(h, count, at.ctypes.data_as(P(f64)), gi.ctypes.data_as(P(i32)),
                               kind.ctypes.data_as(P(i32)), roots.ctypes.data_as(P(u64)))
        lib.g500_server_drain(h)
        wall = time.perf_counter() - t0
        done = sess.server_poll(h, max_items=count + 16)
        stats = sess.server_stats(h)
    finally:
        lib.g500_server_destroy(h)
    lat = np.array([c.latency_ms for c in done])
    edges = 0.0
    for c in done:
        r = graphs[c.graph]
        edges += r.m_giant if c.reached == r.n_giant else max(int(c.reached) - 1, 0)
    bs = np.array([c.batch_size for c in done])
    return {
        "graphs": [f"S{r.scale}s{r.seed[0]}{r.seed[1]}" for r in graphs],
        "rate_offered": rate, "requests": count, "completed": len(done),
        "errors": int(sum(1 for c in done if c.status)),
        "slots": slots, "delta": delta, "launches": stats["launches"], "mean_batch": round(float(bs.mean()), 2),
        "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "wall_s": round(wall, 3), "req_per_s": round(len(done) / wall, 1),
        # arrivals span at[-1]; the rest of the wall time drains the backlog left at the last arrival
        "arrival_s": round(float(at[-1]) * 1e-3, 3), "drain_s": round(wall - float(at[-1]) * 1e-3, 3),
        "gteps": round(edges / wall / 1e9, 4),
    }
