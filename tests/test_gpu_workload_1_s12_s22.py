"""Parity at the plain-method workloads, through the C-ABI:

- SCALE 12 edgefactor 16, all 64 Graph500 roots: BFS depths equal the
  oracle's BFS, parent trees pass rules 1-7 (library and oracle validators),
  TEPS edge counts equal; SSSP (delta 0.1, the reference's value) bit-exact
  against the oracle's FP32 Dijkstra with the certificate holding, and the
  largest |f32 - f64| against the oracle's FP64 Dijkstra below 1e-5.
- SCALE 22 edgefactor 16: the whole tuple list and CSR bit-exact against the
  oracle's, the 64 roots equal, all 64 roots' BFS depths equal the oracle's
  and pass rules 1-7, SSSP for all 64 roots at delta 0.05, 0.1 and 0.25:
  equal bit for bit across deltas and certified; bit-exact against the
  oracle's FP32 Dijkstra on four roots.
"""
import numpy as np
import pytest

from bigcheck import nthreads, say
from conftest import oracle_graph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sess():
    import g500
    s = g500.Session(arena_bytes=24 << 30)
    yield s
    s.close()


@pytest.fixture(scope="module")
def s12(sess, oracle):
    e, w, ref = oracle_graph(12)
    tg = sess.generate(12, 16)
    ge, gw = sess.tuples_to_host(tg)
    assert np.array_equal(ge, e) and np.array_equal(gw.view(np.uint32), w.view(np.uint32))
    g = sess.csr(tg, 1 << 12)
    rs, col, cw = sess.csr_to_host(g)
    assert np.array_equal(rs, ref.rowstarts) and np.array_equal(col, ref.column)
    assert np.array_equal(cw.view(np.uint32), ref.weights.view(np.uint32))
    roots = sess.roots(g)
    assert np.array_equal(roots, oracle.roots(ref)) and len(roots) == 64
    return g, ref, roots


def test_s12_bfs_all_roots(sess, oracle, s12):
    g, ref, roots = s12
    b = sess.bfs_create(g)
    for r in roots:
        r = int(r)
        sess.bfs(g, b, r)
        pred = sess.bfs_pred(b)
        ok, err = sess.bfs_validate(g, b, r)
        assert ok and err == 0, r
        assert oracle.bfs_validate(ref, pred, r) == 0, r
        _, depth = oracle.bfs(ref, r)
        assert np.array_equal(sess.bfs_depth(g, b, r), depth), r
        assert sess.edge_count(g, b.pred) == oracle.edge_count(ref, pred), r


def test_s12_sssp_all_roots(sess, oracle, s12):
    g, ref, roots = s12
    s = sess.sssp_create(g, 0.1)
    worst = 0.0
    for r in roots:
        r = int(r)
        sess.sssp(g, s, r)
        d, p = sess.sssp_result(s)
        d_ref, _ = oracle.dijkstra_f32(ref, r)
        assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32)), r
        assert oracle.sssp_check(ref, r, d, p) == 0, r
        d64 = oracle.dijkstra_f64(ref, r)
        fin = d64 >= 0
        assert np.array_equal(d >= 0, fin), r
        worst = max(worst, float(np.max(np.abs(d[fin].astype(np.float64) - d64[fin]))))
    say("S12, 64 roots: max |f32 - f64| =", worst)
    assert worst < 1e-5


def test_s12_batched_all_roots(sess, oracle, s12):
    g, ref, roots = s12
    m = sess.mark()
    bb = sess.batch_create(g)
    sess.batch(g, bb, roots)
    for i, r in enumerate(roots):
        _, depth = oracle.bfs(ref, int(r))
        assert np.array_equal(sess.batch_depth(g, bb, i), depth), (i, int(r))
    width = int(sess.get_param("sssp_batch_max"))
    sb = sess.sssp_batch_create(g, 0.1)
    hub = sess.sssp_hub_distances(g, roots)
    for offsets in (None, hub):  # shared buckets, then per-root lags from the hub distances
        for k in range(0, 64, width):
            sess.sssp_batch(g, sb, roots[k:k + width], None if offsets is None else offsets[k:k + width])
            for i, r in enumerate(roots[k:k + width]):
                d, p = sess.sssp_batch_result(g, sb, i)
                d_ref, _ = oracle.dijkstra_f32(ref, int(r))
                assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32)), (k + i, int(r))
                assert oracle.sssp_check(ref, int(r), d, p) == 0, (k + i, int(r))
    sess.release(m)


# --------------------------------------------------------------------- SCALE 22
@pytest.fixture(scope="module")
def s22(sess, oracle):
    say("S22: oracle generator and CSR")
    e, w, ref = oracle_graph(22)
    tg = sess.generate(22, 16)
    ge, gw = sess.tuples_to_host(tg)
    assert np.array_equal(ge, e)
    assert np.array_equal(gw.view(np.uint32), w.view(np.uint32))
    del ge, gw
    g = sess.csr(tg, 1 << 22)
    rs, col, cw = sess.csr_to_host(g)
    assert np.array_equal(rs, ref.rowstarts) and np.array_equal(col, ref.column)
    assert np.array_equal(cw.view(np.uint32), ref.weights.view(np.uint32))
    del rs, col, cw
    roots = sess.roots(g)
    assert np.array_equal(roots, oracle.roots(ref)) and len(roots) == 64
    say("S22: generator, CSR and roots equal the oracle's")
    return g, ref, roots


def test_s22_bfs_all_roots(sess, oracle, s22):
    g, ref, roots = s22
    m = sess.mark()
    b = sess.bfs_create(g)
    for i, r in enumerate(roots):
        r = int(r)
        sess.bfs(g, b, r)
        ok, err = sess.bfs_validate(g, b, r)
        assert ok and err == 0, r
        assert np.array_equal(sess.bfs_depth(g, b, r), oracle.bfs_depth_omp(ref, r, nthreads())), r
        if i < 2:  # the oracle's own validator and edge count on two roots
            pred = sess.bfs_pred(b)
            assert oracle.bfs_validate(ref, pred, r) == 0
            assert sess.edge_count(g, b.pred) == oracle.edge_count(ref, pred)
    sess.release(m)


@pytest.mark.parametrize("half", [0, 1])
def test_s22_sssp_delta_sweep(sess, oracle, s22, half):
    """Each half of the 64 roots: distances equal across delta 0.05, 0.1 and
    0.25 bit for bit and certified at each delta."""
    g, ref, roots = s22
    m = sess.mark()
    s = sess.sssp_create(g)
    for r in roots[32 * half:32 * half + 32]:
        r = int(r)
        first = None
        for delta in (0.05, 0.1, 0.25):
            s.delta = delta
            sess.sssp(g, s, r)
            d, p = sess.sssp_result(s)
            assert oracle.sssp_check(ref, r, d, p) == 0, (r, delta)
            if first is None:
                first = d
            else:
                assert np.array_equal(d.view(np.uint32), first.view(np.uint32)), (r, delta)
    sess.release(m)


def test_s22_sssp_matches_dijkstra(sess, oracle, s22):
    g, ref, roots = s22
    m = sess.mark()
    s = sess.sssp_create(g, 0.1)
    for r in roots[:4]:
        r = int(r)
        sess.sssp(g, s, r)
        d, _ = sess.sssp_result(s)
        d_ref, _ = oracle.dijkstra_f32(ref, r)
        assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32)), r
    d64 = oracle.dijkstra_f64(ref, r)
    fin = d64 >= 0
    err = float(np.max(np.abs(d[fin].astype(np.float64) - d64[fin])))
    say("S22: max |f32 - f64| =", err)
    assert err < 1e-5
    sess.release(m)
