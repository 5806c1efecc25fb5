"""Serving loop through the C-ABI: every result is the same alone and in a
mix of concurrent requests (and equal to the oracle), requests wait for result
pool admission, oversized requests are refused."""
import threading

import numpy as np
import pytest

from conftest import oracle_graph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sess():
    import g500
    s = g500.Session(arena_bytes=16 << 30)
    yield s
    s.close()


def say(*a):
    print(*a, flush=True)


def make_graphs(sess, scales):
    out = []
    for sc in scales:
        e, w, ref = oracle_graph(sc)
        out.append((sess.csr(sess.upload_tuples(e, w), 1 << sc), ref))
    return out


def buffers(n, kind):
    import g500
    if kind == g500.BFS:
        return np.full(n, -7, np.int64), None
    return np.full(n, -7, np.float32), np.full(n, -7, np.int64)


def test_serve_alone_vs_mix(sess, oracle):
    import g500
    graphs = make_graphs(sess, [10, 12, 13])
    reqs = []
    for gi, (g, ref) in enumerate(graphs):
        for r in oracle.roots(ref)[:12]:
            reqs.append((gi, g500.BFS, int(r)))
            reqs.append((gi, g500.SSSP, int(r)))
    m = sess.mark()
    h = sess.server([g for g, _ in graphs], slots=2, pool_bytes=256 << 20)
    try:
        # alone: one request in the system at a time
        alone = []
        for gi, kind, root in reqs:
            b = buffers(graphs[gi][0].nlocalverts, kind)
            sess.server_submit(h, gi, kind, root, *b)
            sess.lib.g500_server_drain(h)
            (c,) = sess.server_poll(h)
            assert c.status == 0 and c.batch_size == 1
            alone.append((b, c.reached, c.sum))
        say("alone done", len(reqs))
        # mix: all requests at once from four threads
        bufs = [buffers(graphs[gi][0].nlocalverts, kind) for gi, kind, _ in reqs]
        ids = [0] * len(reqs)

        def submit(part):
            for i in part:
                gi, kind, root = reqs[i]
                ids[i] = sess.lib.g500_server_submit(h, gi, kind, root, bufs[i][0].ctypes.data,
                                                     None if bufs[i][1] is None else bufs[i][1].ctypes.data)

        order = np.random.default_rng(5).permutation(len(reqs))
        ths = [threading.Thread(target=submit, args=(order[k::4],)) for k in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        sess.lib.g500_server_drain(h)
        done = {c.id: c for c in sess.server_poll(h)}
        assert len(done) == len(reqs) and all(i in done for i in ids)
        assert max(c.batch_size for c in done.values()) > 1  # requests were batched together
        for i, (gi, kind, root) in enumerate(reqs):
            c = done[ids[i]]
            assert c.status == 0
            (a0, a1), reached, ssum = alone[i]
            b0, b1 = bufs[i]
            # depths and distances are equal bit for bit; SSSP parents may differ
            # between runs only where two parents give the same distance (the
            # first strict decrease wins), so they are checked by the certificate
            assert np.array_equal(b0.view(np.uint8), a0.view(np.uint8)), reqs[i]
            assert (c.reached, c.sum) == (reached, ssum)
            ref = graphs[gi][1]
            if kind == g500.BFS:
                _, depth_ref = oracle.bfs(ref, root)
                assert np.array_equal(b0, depth_ref)
                assert c.reached == int((depth_ref >= 0).sum()) and c.sum == float(depth_ref[depth_ref >= 0].sum())
            else:
                d_ref, _ = oracle.dijkstra_f32(ref, root)
                assert np.array_equal(b0.view(np.uint32), d_ref.view(np.uint32))
                assert oracle.sssp_check(ref, root, b0, b1) == 0
        st = sess.server_stats(h)
        say("server stats", st)
        assert st["completed"] == 2 * len(reqs) and st["pool_used"] == 0
    finally:
        sess.lib.g500_server_destroy(h)
        sess.release(m)


def test_serve_pool_admission(sess, oracle):
    """A pool that holds two SSSP results: every request still completes, and
    no launch holds more requests than the pool admits."""
    import g500
    (g, ref), = make_graphs(sess, [12])
    n = g.nlocalverts
    per = 12 * n  # f32 distances + i64 parents
    m = sess.mark()
    h = sess.server([g], slots=2, pool_bytes=2 * per + 1024)
    try:
        roots = oracle.roots(ref)[:10]
        bufs = [buffers(n, g500.SSSP) for _ in roots]
        ids = [sess.server_submit(h, 0, g500.SSSP, int(r), *b) for r, b in zip(roots, bufs)]
        sess.lib.g500_server_drain(h)
        done = {c.id: c for c in sess.server_poll(h)}
        assert sorted(done) == sorted(ids)
        assert max(c.batch_size for c in done.values()) <= 2
        for r, (d, p) in zip(roots, bufs):
            d_ref, _ = oracle.dijkstra_f32(ref, int(r))
            assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32))
    finally:
        sess.lib.g500_server_destroy(h)
        sess.release(m)
    m = sess.mark()
    h = sess.server([g], slots=1, pool_bytes=per // 2)
    try:
        with pytest.raises(g500.LibraryError):  # results larger than the whole pool
            sess.server_submit(h, 0, g500.SSSP, int(roots[0]), *buffers(n, g500.SSSP))
        # a request without result buffers needs no pool memory
        rid = sess.server_submit(h, 0, g500.BFS, int(roots[0]))
        sess.lib.g500_server_drain(h)
        (c,) = sess.server_poll(h)
        assert c.id == rid and c.status == 0
    finally:
        sess.lib.g500_server_destroy(h)
        sess.release(m)


def test_serve_replay_summaries(sess, oracle):
    """Timed submissions without result buffers: summaries equal the oracle's."""
    import g500
    graphs = make_graphs(sess, [11, 13])
    m = sess.mark()
    h = sess.server([g for g, _ in graphs], slots=3, pool_bytes=1 << 20)
    try:
        rng = np.random.default_rng(7)
        cnt = 200
        gi = rng.integers(0, 2, cnt).astype(np.int32)
        kind = rng.integers(0, 2, cnt).astype(np.int32)
        roots = np.array([int(rng.choice(oracle.roots(graphs[k][1]))) for k in gi], np.uint64)
        at = np.cumsum(rng.exponential(0.05, cnt))  # 20k requests/s
        sess.lib.g500_server_replay(h, cnt, at.ctypes.data_as(g500.P(g500.f64)), gi.ctypes.data_as(g500.P(g500.i32)),
                                    kind.ctypes.data_as(g500.P(g500.i32)), roots.ctypes.data_as(g500.P(g500.u64)))
        sess.lib.g500_server_drain(h)
        done = sess.server_poll(h)
        assert len(done) == cnt
        lat = np.array([c.latency_ms for c in done])
        say("replay latency ms p50 %.3f p99 %.3f" % (np.percentile(lat, 50), np.percentile(lat, 99)))
        for c in done:
            ref = graphs[c.graph][1]
            if c.kind == g500.BFS:
                _, depth = oracle.bfs(ref, int(c.root))
                assert c.reached == int((depth >= 0).sum()) and c.sum == float(depth[depth >= 0].sum())
            else:
                d, _ = oracle.dijkstra_f32(ref, int(c.root))
                fin = d[d >= 0]
                assert c.reached == fin.size
                assert c.sum == float(np.floor(fin.astype(np.float64) * 16777216.0).sum()) / 16777216.0
    finally:
        sess.lib.g500_server_destroy(h)
        sess.release(m)
