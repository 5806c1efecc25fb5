"""The no-allocation validation entry point (oned_graph_bfs_validate_noalloc)
against oned_graph_bfs_validate through the C-ABI: the same verdict, error
count and depth array on valid and corrupted parent trees, on two calls in a
row with the second graph smaller (stale scratch bits must not leak), and a
clear error above its vertex maximum."""
import ctypes as C

import numpy as np
import pytest

from conftest import oracle_graph

pytestmark = pytest.mark.gpu


def corrupted(pred, root, n, seed):
    """The valid tree and seven kinds of planted mistakes (on the host)."""
    rng = np.random.default_rng(seed)
    reached = np.flatnonzero(pred >= 0)
    out = [pred.copy()]
    for kind in range(7):
        p = pred.copy()
        v = int(rng.choice(reached[reached != root]))
        if kind == 0: p[v] = n + 9
        if kind == 1: p[root] = -1
        if kind == 2: p[v] = -1
        if kind == 3: p[v] = v
        if kind == 4: p[v] = int(rng.choice(reached))
        if kind == 5: p[v] = -7
        if kind == 6 and (pred < 0).any(): p[int(rng.choice(np.flatnonzero(pred < 0)))] = root
        out.append(p)
    return out


def noalloc(sess, g, pred_dev, root, depth_dev):
    import g500
    err = g500.u64()
    before = sess.mark()
    ok = sess.lib.oned_graph_bfs_validate_noalloc(C.byref(g), pred_dev, int(root), depth_dev, C.byref(err))
    sess.check()
    assert sess.mark() == before  # nothing taken from the arena
    return bool(ok), err.value


def test_noalloc_validation_matches(oracle):
    import g500
    s = g500.Session(arena_bytes=8 << 30)
    try:
        for scale in (16, 12):  # the second call on a smaller graph
            e, w, ref = oracle_graph(scale)
            g = s.csr(s.upload_tuples(e, w), 1 << scale)
            n = int(g.nlocalverts)
            b = s.bfs_create(g)
            depth_dev = s.lib._worker_arena_push_array(C.byref(s.ws), n, 8, 8, C.byref(g500.u64()))
            assert depth_dev
            root = int(oracle.roots(ref)[0])
            s.bfs(g, b, root)
            pred = s.bfs_pred(b)
            for i, p in enumerate(corrupted(pred, root, n, scale)):
                s.to_device(b.pred, p)
                ok_a, err_a = s.bfs_validate(g, b, root)
                depth_a = s.bfs_depth(g, b, root)
                s.to_device(depth_dev, np.full(n, 12345, np.int64))
                ok_b, err_b = noalloc(s, g, b.pred, root, depth_dev)
                depth_b = s.to_host(depth_dev, n, np.int64)
                assert (ok_b, err_b) == (ok_a, err_a), (scale, i)
                assert np.array_equal(depth_b, depth_a), (scale, i)
                assert err_b == oracle.bfs_validate(ref, p, root), (scale, i)
                if i == 0:
                    assert ok_b and err_b == 0
                if i in (1, 2, 3, 4, 6):  # out of range, root cleared, reached cut off, self parent, negative
                    assert not ok_b and err_b > 0, (scale, i)
        # above the maximum: a clear error, nothing launched
        big = g500.OnedGraph()
        big.nlocalverts = big.nglobalverts = (1 << 27) + 1
        with pytest.raises(g500.LibraryError, match="no-allocation maximum"):
            noalloc(s, big, b.pred, 0, depth_dev)
    finally:
        s.close()
