"""Serving on the configured resident graphs (Kronecker SCALE 16, 18, 20, 22,
24 with user seed (2, 3) and SCALE 16, 18, 20 with user seed (5, 7),
edgefactor 16), through the C-ABI:

- 48 requests (3 roots per graph, BFS and SSSP each) give the same depths and
  distances bit for bit alone and in a concurrent mix submitted from four
  threads; BFS depths equal the oracle's BFS, SSSP results pass the
  certificate, and on the SCALE 16 graphs equal the oracle's FP32 Dijkstra;
- replayed Poisson arrivals at 5k and 20k requests/s over the same
  (graph, kind, root) triples complete without error, with the summaries
  (reached count, depth / distance sum) of the alone runs.
"""
import threading

import numpy as np
import pytest

from bigcheck import nthreads, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served():
    import g500
    from g500 import serving
    s = g500.Session(arena_bytes=64 << 30)
    res = serving.load_graphs(s)
    hosts = []
    import pyoracle
    for r in res:
        rs, col, w = s.csr_to_host(r.g)
        hosts.append(pyoracle.Csr(int(r.g.nlocalverts), rs, col, w))
    say("resident graphs:", [f"S{r.scale}s{r.seed[0]}{r.seed[1]}" for r in res])
    yield s, res, hosts
    s.close()


def buffers(n, kind):
    import g500
    if kind == g500.BFS:
        return np.full(n, -7, np.int64), None
    return np.full(n, -7, np.float32), np.full(n, -7, np.int64)


def requests(res):
    import g500
    rng = np.random.default_rng(11)
    out = []
    for gi, r in enumerate(res):
        for root in rng.choice(r.candidates, 3, replace=False):
            out.append((gi, g500.BFS, int(root)))
            out.append((gi, g500.SSSP, int(root)))
    return out


def test_serving_alone_vs_mix(served, oracle):
    import g500
    s, res, hosts = served
    reqs = requests(res)
    m = s.mark()
    h = s.server([r.g for r in res], slots=2, pool_bytes=2 << 30)
    try:
        alone = []
        for gi, kind, root in reqs:
            b = buffers(int(res[gi].g.nlocalverts), kind)
            s.server_submit(h, gi, kind, root, *b)
            s.lib.g500_server_drain(h)
            (c,) = s.server_poll(h)
            assert c.status == 0 and c.batch_size == 1
            alone.append((b, c.reached, c.sum))
        say("alone:", len(reqs), "requests")
        bufs = [buffers(int(res[gi].g.nlocalverts), kind) for gi, kind, _ in reqs]
        ids = [0] * len(reqs)

        def submit(part):
            for i in part:
                gi, kind, root = reqs[i]
                ids[i] = s.lib.g500_server_submit(h, gi, kind, root, bufs[i][0].ctypes.data,
                                                  None if bufs[i][1] is None else bufs[i][1].ctypes.data)

        order = np.random.default_rng(5).permutation(len(reqs))
        ths = [threading.Thread(target=submit, args=(order[k::4],)) for k in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        s.lib.g500_server_drain(h)
        done = {c.id: c for c in s.server_poll(h)}
        assert len(done) == len(reqs) and all(i in done for i in ids)
        say("mix: largest batch", max(c.batch_size for c in done.values()))
        for i, (gi, kind, root) in enumerate(reqs):
            c = done[ids[i]]
            assert c.status == 0, reqs[i]
            (a0, a1), reached, ssum = alone[i]
            b0, b1 = bufs[i]
            assert np.array_equal(b0.view(np.uint8), a0.view(np.uint8)), reqs[i]
            assert (c.reached, c.sum) == (reached, ssum), reqs[i]
            ref = hosts[gi]
            if kind == g500.BFS:
                assert np.array_equal(b0, oracle.bfs_depth_omp(ref, root, nthreads())), reqs[i]
            else:
                assert oracle.sssp_check(ref, root, b0, b1) == 0, reqs[i]
                assert oracle.sssp_check(ref, root, a0, a1) == 0, reqs[i]
                if res[gi].scale == 16:
                    d_ref, _ = oracle.dijkstra_f32(ref, root)
                    assert np.array_equal(b0.view(np.uint32), d_ref.view(np.uint32)), reqs[i]
        served_alone = {reqs[i]: (alone[i][1], alone[i][2]) for i in range(len(reqs))}
    finally:
        s.lib.g500_server_destroy(h)
        s.release(m)

    # replayed arrivals at the configured rates: summaries equal the alone runs'
    keys = list(served_alone)
    for rate in (5000.0, 20000.0):
        rng = np.random.default_rng(int(rate))
        cnt = 256
        pick = rng.integers(0, len(keys), cnt)
        gi = np.array([keys[k][0] for k in pick], np.int32)
        kind = np.array([keys[k][1] for k in pick], np.int32)
        roots = np.array([keys[k][2] for k in pick], np.uint64)
        at = np.cumsum(rng.exponential(1e3 / rate, cnt))
        m = s.mark()
        h = s.server([r.g for r in res], slots=2, pool_bytes=1 << 20)
        try:
            s.lib.g500_server_replay(h, cnt, at.ctypes.data_as(g500.P(g500.f64)),
                                     gi.ctypes.data_as(g500.P(g500.i32)), kind.ctypes.data_as(g500.P(g500.i32)),
                                     roots.ctypes.data_as(g500.P(g500.u64)))
            s.lib.g500_server_drain(h)
            done = s.server_poll(h, max_items=cnt + 16)
            assert len(done) == cnt
            lat = np.array([c.latency_ms for c in done])
            say(f"replay at {rate:g}/s: p50 {np.percentile(lat, 50):.1f} ms, p99 {np.percentile(lat, 99):.1f} ms, "
                f"mean batch {np.mean([c.batch_size for c in done]):.1f}")
            for c in done:
                assert c.status == 0
                assert (c.reached, c.sum) == served_alone[(c.graph, c.kind, int(c.root))]
        finally:
            s.lib.g500_server_destroy(h)
            s.release(m)
