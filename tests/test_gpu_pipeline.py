"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Bars: generator edges/weights and CSR arrays bit-exact; BFS depths bit-exact
with parent trees passing the Graph500 rules (GPU validator and oracle
validator); validation error counts and depth arrays equal on corrupted
parent arrays; SSSP FP32 distances bit-exact against the oracle's FP32
Dijkstra with the parent certificate holding; roots and TEPS edge counts equal.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import oracle_graph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sess():
    import g500
    s = g500.Session(arena_bytes=48 << 30)
    yield s
    s.close()


@pytest.fixture
def arena(sess):
    m = sess.mark()
    yield sess
    sess.release(m)


def say(*a):
    print(*a, flush=True)


# ------------------------------------------------------------------ generator
@pytest.mark.parametrize("scale,ef,s1,s2", [(1, 16, 2, 3), (5, 16, 2, 3), (10, 16, 2, 3), (12, 16, 2, 3),
                                            (12, 4, 2, 3), (11, 32, 7, 11)])
def test_generator_bitexact(arena, oracle, scale, ef, s1, s2):
    say("generator", scale, ef, s1, s2)
    tg = arena.generate(scale, ef, s1, s2)
    e, w = arena.tuples_to_host(tg)
    eo, wo = oracle.kronecker(scale, ef, s1, s2)
    assert np.array_equal(e, eo)
    assert np.array_equal(w.view(np.uint32), wo.view(np.uint32))


@pytest.mark.parametrize("scale,start,end", [(16, 1000, 70001), (20, 2**20 - 300, 2**20 + 5000),
                                             (28, 2**33 + 17, 2**33 + 3000), (26, 2**30 - 999, 2**30)])
def test_generator_ranges(arena, oracle, scale, start, end):
    say("generator range", scale, start, end)
    tg = arena.generate(scale, 16, start=start, end=end)
    e, w = arena.tuples_to_host(tg)
    eo, wo = oracle.kronecker(scale, 16, start=start, end=end)
    assert np.array_equal(e, eo)
    assert np.array_equal(w.view(np.uint32), wo.view(np.uint32))


# ------------------------------------------------------------------------ CSR
def check_csr(sess, oracle, n, edges, weights):
    tg = sess.upload_tuples(edges, weights)
    g = sess.csr(tg, n)
    rs, col, w = sess.csr_to_host(g)
    ref = oracle.csr(n, edges, weights)
    assert np.array_equal(rs, ref.rowstarts)
    assert np.array_equal(col, ref.column)
    assert np.array_equal(w.view(np.uint32), ref.weights.view(np.uint32))
    deg = np.diff(rs.astype(np.int64))
    assert g.notisolated == int((deg > 0).sum())
    return g


@pytest.mark.parametrize("scale,ef", [(1, 1), (6, 4), (10, 16), (12, 16), (16, 16)])
def test_csr_matches_oracle(arena, oracle, scale, ef):
    say("csr", scale, ef)
    e, w, _ = oracle_graph(scale, ef)
    check_csr(arena, oracle, 1 << scale, e, w)


def test_csr_from_generated_tuples(arena, oracle):
    tg = arena.generate(14, 16)
    g = arena.csr(tg, 1 << 14)
    rs, col, w = arena.csr_to_host(g)
    _, _, ref = oracle_graph(14, 16)
    assert np.array_equal(rs, ref.rowstarts) and np.array_equal(col, ref.column)
    assert np.array_equal(w.view(np.uint32), ref.weights.view(np.uint32))


def test_csr_edge_cases(arena, oracle):
    rng = np.random.default_rng(5)
    cases = []
    cases.append((10, np.zeros((0, 2), np.uint64), np.zeros(0, np.float32)))                      # no edges
    cases.append((10, np.array([[3, 3], [7, 7]], np.uint64), np.array([0.5, 0.25], np.float32)))   # loops only
    # a star of 20000 leaves (radix tier), rows of 70 (block tier) and 40 (wave tier)
    n = 50000
    star = np.stack([np.zeros(20000, np.uint64), np.arange(1, 20001, dtype=np.uint64)], 1)
    mid = np.stack([np.full(70 * 30, 0, np.uint64) + np.repeat(np.arange(30000, 30030, dtype=np.uint64), 70),
                    rng.integers(0, n, 70 * 30).astype(np.uint64)], 1)
    small = np.stack([np.repeat(np.arange(40000, 40100, dtype=np.uint64), 40),
                      rng.integers(0, n, 4000).astype(np.uint64)], 1)
    e = np.concatenate([star, mid, small, star[:500]])                                        # duplicates
    w = rng.random(len(e)).astype(np.float32)
    w[:7] = 0.0
    w[7:20] = 1.0
    w[-500:] = w[:500]                                                                          # tied duplicates
    cases.append((n, e, w))
    # 3 vertices, n not a multiple of 64
    cases.append((3, np.array([[0, 1], [1, 2], [2, 0], [0, 1]], np.uint64), np.array([.1, .2, .3, .1], np.float32)))
    for n, e, w in cases:
        say("csr case", n, len(e))
        g = check_csr(arena, oracle, n, e, w)
        check_sssp_copy(arena, g)


def test_traversal_copy_is_a_relabeling(arena, oracle):
    """The degree-ordered traversal copy holds the same rows as the CSR under
    a permutation, in decreasing degree order, entry order unchanged."""
    e, w, ref = oracle_graph(14)
    n = 1 << 14
    g = check_csr(arena, oracle, n, e, w)
    perm = arena.to_host(g.perm, n, np.uint32).astype(np.int64)
    iperm = arena.to_host(g.iperm, n, np.uint32).astype(np.int64)
    rs_t = arena.to_host(g.rowstarts_t, n + 1, np.uint64).astype(np.int64)
    col_t = arena.to_host(g.column_t, g.nlocaledges, np.uint32).astype(np.int64)
    w_t = arena.to_host(g.weights_t, g.nlocaledges, np.float32)
    assert np.array_equal(np.sort(perm), np.arange(n)) and np.array_equal(iperm[perm], np.arange(n))
    rs = ref.rowstarts.astype(np.int64)
    deg = np.diff(rs)
    deg_t = np.diff(rs_t)
    assert np.array_equal(deg_t, deg[iperm])
    assert np.all(np.diff(np.minimum(deg_t, 4095)) <= 0)
    src = np.repeat(rs[iperm] - rs_t[:-1], deg_t) + np.arange(g.nlocaledges)
    assert np.array_equal(col_t, perm[ref.column.astype(np.int64)[src]])
    assert np.array_equal(w_t.view(np.uint32), ref.weights[src].view(np.uint32))


def check_sssp_copy(sess, g):
    """The SSSP copy is the traversal copy without exactly the entries that
    the two-hop rule drops (hubs: traversal ids 0..31), rows and entry order
    otherwise unchanged. Returns (kept, total) entries."""
    from conftest import drop_entries, prune_rule
    n, m = g.nlocalverts, g.nlocaledges
    assert g.rowstarts_s and g.column_s and g.weights_s
    rs_t = sess.to_host(g.rowstarts_t, n + 1, np.uint64)
    col_t = sess.to_host(g.column_t, m, np.uint32)
    w_t = sess.to_host(g.weights_t, m, np.float32)
    # hubs: the 32 largest degrees of the traversal copy, ties to the smaller id
    hubs = np.argsort(-np.diff(rs_t.astype(np.int64)), kind="stable")[:32]
    drop = prune_rule(rs_t, col_t, w_t, hubs)
    rs_k, col_k, w_k = drop_entries(rs_t, col_t, w_t, drop)
    rs_s = sess.to_host(g.rowstarts_s, n + 1, np.uint64)
    assert np.array_equal(rs_s, rs_k)
    m_s = int(rs_s[-1])
    assert np.array_equal(sess.to_host(g.column_s, m_s, np.uint32), col_k)
    assert np.array_equal(sess.to_host(g.weights_s, m_s, np.float32).view(np.uint32), w_k.view(np.uint32))
    return m_s, m


@pytest.mark.parametrize("scale", [12, 14])
def test_sssp_copy_drops_two_hop_dominated_entries(arena, oracle, scale):
    e, w, ref = oracle_graph(scale)
    g = check_csr(arena, oracle, 1 << scale, e, w)
    kept, total = check_sssp_copy(arena, g)
    say("sssp copy", scale, "kept", kept, "of", total)
    assert kept < 0.8 * total


def test_sssp_copy_hubs_above_exact_degree_range(arena, oracle):
    """40 vertices of degree above 4095 share one degree class of the
    traversal copy in no particular order: the SSSP copy still takes the 32
    largest degrees as hubs, and SSSP distances match the oracle."""
    rng = np.random.default_rng(11)
    n = 60000
    hubs = rng.choice(n, 40, replace=False)
    parts = []
    for i, h in enumerate(hubs):
        d = 4100 + 97 * i
        parts.append(np.stack([np.full(d, h, np.uint64), rng.integers(0, n, d).astype(np.uint64)], 1))
    parts.append(rng.integers(0, n, (120000, 2)).astype(np.uint64))
    e = np.concatenate(parts)
    w = rng.random(len(e)).astype(np.float32)
    g = check_csr(arena, oracle, n, e, w)
    kept, total = check_sssp_copy(arena, g)
    say("sssp copy, 40 hubs", kept, "of", total)
    assert kept < total
    ref = oracle.csr(n, e, w)
    s = arena.sssp_create(g)
    for r in (int(hubs[0]), int(e[-1, 0])):
        arena.sssp(g, s, r)
        d, p = arena.sssp_result(s)
        d_ref, _ = oracle.dijkstra_f32(ref, r)
        assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32))
        assert oracle.sssp_check(ref, r, d, p) == 0


def test_sssp_copy_distance_bound(arena, oracle):
    """Distances at or above 2048 void the SSSP copy's guarantee: on a path of
    2199 edges of weight 0.99 both SSSP kernels finish and report error 32;
    built without the copy, the same graph runs exact."""
    import g500
    n = 2200
    e = np.stack([np.arange(n - 1, dtype=np.uint64), np.arange(1, n, dtype=np.uint64)], 1)
    w = np.full(n - 1, 0.99, np.float32)
    ref = oracle.csr(n, e, w)
    g = check_csr(arena, oracle, n, e, w)
    assert g.rowstarts_s
    s = arena.sssp_create(g)
    s.delta = 1.0
    with pytest.raises(g500.LibraryError, match="error 32 "):
        arena.sssp(g, s, 0)
    sb = arena.sssp_batch_create(g)
    sb.delta = 1.0
    with pytest.raises(g500.LibraryError, match="error 32 "):
        arena.sssp_batch(g, sb, np.array([0, 7], np.uint64))
    arena.set_param("sssp_prune", 0)
    try:
        g = check_csr(arena, oracle, n, e, w)
        assert not g.rowstarts_s
        s = arena.sssp_create(g)
        s.delta = 1.0
        arena.sssp(g, s, 0)
        d, p = arena.sssp_result(s)
        d_ref, _ = oracle.dijkstra_f32(ref, 0)
        assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32)) and d.max() > 2048
        assert oracle.sssp_check(ref, 0, d, p) == 0
        sb = arena.sssp_batch_create(g)
        sb.delta = 1.0
        sssp_batch_check(arena, oracle, g, ref, sb, np.array([0, 7], np.uint64))
    finally:
        arena.set_param("sssp_prune", 1)


def test_sssp_copy_skipped_without_arena_room(arena, oracle):
    """The SSSP copy is an optimization: with arena room for the CSR and the
    traversal copy but not for the SSSP copy's hub table (128 B per vertex),
    the graph keeps no SSSP copy and SSSP relaxes the traversal copy."""
    e, w, ref = oracle_graph(12)
    a = arena.ws.gpu_arena
    saved = a.reserved
    a.reserved = a.used + (6 << 20)  # the SSSP copy needs about 9.6 MB here
    try:
        g = check_csr(arena, oracle, 1 << 12, e, w)
    finally:
        a.reserved = saved
    assert g.perm and not g.rowstarts_s
    s = arena.sssp_create(g)
    for r in oracle.roots(ref)[:2]:
        arena.sssp(g, s, int(r))
        d, p = arena.sssp_result(s)
        d_ref, _ = oracle.dijkstra_f32(ref, int(r))
        assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32))


def test_without_sssp_copy(arena, oracle):
    """With pruning switched off both SSSP kernels relax the traversal copy."""
    e, w, ref = oracle_graph(12)
    arena.set_param("sssp_prune", 0)
    try:
        g = check_csr(arena, oracle, 1 << 12, e, w)
        assert g.perm and not g.rowstarts_s and not g.column_s
        s = arena.sssp_create(g)
        sb = arena.sssp_batch_create(g)
        roots = oracle.roots(ref)
        for r in roots[:3]:
            arena.sssp(g, s, int(r))
            d, p = arena.sssp_result(s)
            d_ref, _ = oracle.dijkstra_f32(ref, int(r))
            assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32))
            assert oracle.sssp_check(ref, int(r), d, p) == 0
        sssp_batch_check(arena, oracle, g, ref, sb, roots[:16])
    finally:
        arena.set_param("sssp_prune", 1)


def test_without_traversal_copy(arena, oracle):
    """With the copy switched off the kernels run on the CSR itself."""
    e, w, ref = oracle_graph(12)
    arena.set_param("relabel", 0)
    try:
        g = check_csr(arena, oracle, 1 << 12, e, w)
        assert not g.perm and not g.column_t and not g.rowstarts_s
        arena.bfs_state = arena.bfs_create(g)
        s = arena.sssp_create(g)
        sb = arena.sssp_batch_create(g)
        bb = arena.batch_create(g)
        roots = oracle.roots(ref)
        for r in roots[:3]:
            bfs_check(arena, oracle, g, ref, int(r))
            arena.sssp(g, s, int(r))
            d, p = arena.sssp_result(s)
            d_ref, _ = oracle.dijkstra_f32(ref, int(r))
            assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32))
            assert oracle.sssp_check(ref, int(r), d, p) == 0
        batch_check(arena, oracle, g, ref, bb, roots)
        sssp_batch_check(arena, oracle, g, ref, sb, roots[:16])
    finally:
        arena.set_param("relabel", 1)


# ------------------------------------------------------------------------ BFS
def bfs_check(sess, oracle, g, ref, root):
    sess.bfs(g, sess.bfs_state, root)
    pred = sess.bfs_pred(sess.bfs_state)
    _, depth_ref = oracle.bfs(ref, root)
    assert oracle.bfs_validate(ref, pred, root) == 0
    ok, nerr = sess.bfs_validate(g, sess.bfs_state, root)
    assert ok and nerr == 0
    depth = sess.bfs_depth(g, sess.bfs_state, root)
    assert np.array_equal(depth, depth_ref)
    assert sess.edge_count(g, sess.bfs_state.pred) == oracle.edge_count(ref, pred)
    return pred


@pytest.mark.parametrize("scale", [10, 12, 16])
def test_bfs_depths_and_rules(arena, oracle, scale):
    e, w, ref = oracle_graph(scale)
    g = check_csr(arena, oracle, 1 << scale, e, w)
    arena.bfs_state = arena.bfs_create(g)
    roots = oracle.roots(ref)
    assert np.array_equal(arena.roots(g), roots)
    for r in roots[:8]:
        say("bfs", scale, int(r))
        bfs_check(arena, oracle, g, ref, int(r))


# Beamer's rule: top-down -> bottom-up when m_f > m_u / alpha (a large alpha
# switches early); bottom-up -> top-down when n_f < n / beta (a large beta stays).
@pytest.mark.parametrize("alpha,beta,hub", [(1e-9, 18, 2048), (1e9, 1e9, 2048), (15, 18, 4), (2, 4, 64)])
def test_bfs_direction_modes(arena, oracle, alpha, beta, hub):
    e, w, ref = oracle_graph(14)
    g = check_csr(arena, oracle, 1 << 14, e, w)
    arena.bfs_state = arena.bfs_create(g)
    old = {k: arena.get_param(k) for k in ("bfs_alpha", "bfs_beta", "bfs_hub_degree")}
    try:
        arena.set_param("bfs_alpha", alpha)
        arena.set_param("bfs_beta", beta)
        arena.set_param("bfs_hub_degree", hub)
        for r in oracle.roots(ref)[:4]:
            say("bfs modes", alpha, beta, hub, int(r))
            bfs_check(arena, oracle, g, ref, int(r))
            modes, sizes = arena.level_log(arena.bfs_state)
            if alpha < 1e-3:
                assert not any(modes)      # top-down only
            if alpha > 1e8 and beta > 1e8:
                assert all(modes)          # bottom-up only
    finally:
        for k, v in old.items():
            arena.set_param(k, v)


# ----------------------------------------------------------- batched BFS (64 roots)
def batch_check(sess, oracle, g, ref, b, roots):
    sess.batch(g, b, roots)
    for i, r in enumerate(roots):
        _, depth_ref = oracle.bfs(ref, int(r))
        depth = sess.batch_depth(g, b, i)
        assert np.array_equal(depth, depth_ref), (i, int(r))


@pytest.mark.parametrize("scale", [10, 12, 16])
def test_bfs_batch_depths(arena, oracle, scale):
    e, w, ref = oracle_graph(scale)
    g = check_csr(arena, oracle, 1 << scale, e, w)
    b = arena.batch_create(g)
    roots = oracle.roots(ref)
    say("batch", scale, len(roots))
    batch_check(arena, oracle, g, ref, b, roots)
    # fewer roots, repeats, an isolated vertex as a root
    isolated = int(np.flatnonzero(np.diff(ref.rowstarts.astype(np.int64)) == 0)[0])
    some = np.array([roots[3], roots[3], isolated, roots[0], roots[5]], np.uint64)
    say("batch repeats", scale, some.tolist())
    batch_check(arena, oracle, g, ref, b, some)


# a huge alpha switches to bottom-up at once; a tiny one stays top-down; a
# small hub degree sends most top-down records through the grid-wide hub
# lists; a small heavy degree sends most bottom-up vertices through the
# grid-wide chunk pass, a huge one none
@pytest.mark.parametrize("alpha,beta,hub,heavy", [(1e-9, 18, 2048, 1024), (1e9, 1e9, 2048, 1024),
                                                  (1e9, 1e9, 2048, 1), (1e9, 1e9, 2048, 1 << 30),
                                                  (15, 18, 4, 16), (2, 4, 64, 64)])
def test_bfs_batch_direction_modes(arena, oracle, alpha, beta, hub, heavy):
    e, w, ref = oracle_graph(14)
    g = check_csr(arena, oracle, 1 << 14, e, w)
    b = arena.batch_create(g)
    old = {k: arena.get_param(k) for k in ("batch_alpha", "batch_beta", "bfs_hub_degree", "batch_heavy_degree")}
    try:
        arena.set_param("batch_alpha", alpha)
        arena.set_param("batch_beta", beta)
        arena.set_param("bfs_hub_degree", hub)
        arena.set_param("batch_heavy_degree", heavy)
        say("batch modes", alpha, beta, hub, heavy)
        batch_check(arena, oracle, g, ref, b, oracle.roots(ref))
        modes, recs, ms = arena.batch_level_log(b)
        if alpha < 1e-3:
            assert not any(modes)
        if alpha > 1e8 and beta > 1e8:
            assert all(modes[1:])
    finally:
        for k, v in old.items():
            arena.set_param(k, v)


def test_bfs_batch_errors_return(arena):
    import g500
    e, w, ref = oracle_graph(10)
    g = arena.csr(arena.upload_tuples(e, w), 1 << 10)
    b = arena.batch_create(g)
    for roots in ([], list(range(65)), [1 << 10]):
        with pytest.raises(g500.LibraryError):
            arena.batch(g, b, np.array(roots, np.uint64))
    arena.batch(g, b, np.array([1], np.uint64))
    with pytest.raises(g500.LibraryError):
        arena.batch_depth(g, b, 1)


def test_validation_parity_on_corrupted_trees(arena, oracle):
    e, w, ref = oracle_graph(12)
    g = check_csr(arena, oracle, 1 << 12, e, w)
    b = arena.bfs_create(g)
    root = int(oracle.roots(ref)[0])
    arena.bfs(g, b, root)
    pred = arena.bfs_pred(b)
    rng = np.random.default_rng(1)
    reached = np.flatnonzero(pred >= 0)
    cases = []
    for kind in range(7):
        p = pred.copy()
        v = int(rng.choice(reached[reached != root]))
        if kind == 0: p[v] = g.nlocalverts + 9
        if kind == 1: p[root] = -1
        if kind == 2: p[v] = -1
        if kind == 3: p[v] = v
        if kind == 4: p[v] = int(rng.choice(reached))
        if kind == 5: p[v] = -7
        if kind == 6: p[int(rng.choice(np.flatnonzero(pred < 0)))] = root if (pred < 0).any() else root
        cases.append(p)
    for p in cases:
        arena.to_device(b.pred, p)
        ok, nerr = arena.bfs_validate(g, b, root)
        ref_err = oracle.bfs_validate(ref, p, root)
        assert nerr == ref_err and ok == (ref_err == 0)
        depth = arena.bfs_depth(g, b, root)
        assert np.array_equal(depth, oracle.dist_from_pred(ref, p, root))


def test_bfs_errors_return(arena):
    import g500
    e, w, ref = oracle_graph(10)
    tg = arena.upload_tuples(e, w)
    g = arena.csr(tg, 1 << 10)
    b = arena.bfs_create(g)
    with pytest.raises(g500.LibraryError):
        arena.bfs(g, b, 1 << 10)
    with pytest.raises(g500.LibraryError):
        s = arena.sssp_create(g)
        arena.sssp(g, s, (1 << 10) + 5)


# ----------------------------------------------------------------------- SSSP
@pytest.mark.parametrize("scale", [10, 12, 16])
def test_sssp_bitexact(arena, oracle, scale):
    e, w, ref = oracle_graph(scale)
    g = check_csr(arena, oracle, 1 << scale, e, w)
    s = arena.sssp_create(g)
    for r in oracle.roots(ref)[:4]:
        r = int(r)
        d_ref, _ = oracle.dijkstra_f32(ref, r)
        for delta in (0.1, 0.05, 0.25, 0.003, 4.0):
            say("sssp", scale, r, delta)
            s.delta = delta
            arena.sssp(g, s, r)
            d, p = arena.sssp_result(s)
            assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32))
            assert oracle.sssp_check(ref, r, d, p) == 0
            assert arena.edge_count(g, s.pred) == oracle.edge_count(ref, p)


def test_sssp_max_error_vs_f64(arena, oracle):
    e, w, ref = oracle_graph(14)
    g = check_csr(arena, oracle, 1 << 14, e, w)
    s = arena.sssp_create(g)
    r = int(oracle.roots(ref)[0])
    arena.sssp(g, s, r)
    d, _ = arena.sssp_result(s)
    d64 = oracle.dijkstra_f64(ref, r)
    err = float(np.max(np.abs(d.astype(np.float64) - d64)))
    say("sssp max |f32 - f64| =", err)
    assert err < 1e-5


# ------------------------------------------------------- batched SSSP (16 roots)
def sssp_batch_check(sess, oracle, g, ref, s, roots):
    sess.sssp_batch(g, s, roots)
    for i, r in enumerate(roots):
        r = int(r)
        d_ref, _ = oracle.dijkstra_f32(ref, r)
        d, p = sess.sssp_batch_result(g, s, i)
        assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32)), (i, r)
        assert oracle.sssp_check(ref, r, d, p) == 0, (i, r)


@pytest.mark.parametrize("scale", [10, 12, 16])
def test_sssp_batch_bitexact(arena, oracle, scale):
    e, w, ref = oracle_graph(scale)
    g = check_csr(arena, oracle, 1 << scale, e, w)
    s = arena.sssp_batch_create(g)
    roots = oracle.roots(ref)
    width = int(arena.get_param("sssp_batch_max"))
    # (delta, mid, long, far thresholds): all three deferral classes, mid below
    # delta (no mid class), all below delta (no window), narrow and wide
    # windows, no deferral at all, thresholds out of order (clamped: classes
    # empty), far only
    for delta, wm, wl, wf in ((0.02, 0.05, 0.2, 0.5), (0.1, 0.05, 0.2, 0.4), (0.25, 0.05, 0.2, 0.2),
                              (0.02, 0.0, 0.05, 0.3), (0.02, 0.0, 0.0, 0.0), (0.02, 0.1, 0.7, 0.9),
                              (0.02, 0.5, 0.2, 0.1), (0.02, 0.0, 0.0, 0.3)):
        s.delta, s.mid_weight, s.long_weight, s.far_weight = delta, wm, wl, wf
        say("sssp batch", scale, delta, wm, wl, wf)
        sssp_batch_check(arena, oracle, g, ref, s, roots[:width])
    # fewer roots, repeats, an isolated vertex as a root
    s.delta, s.mid_weight, s.long_weight, s.far_weight = 0.1, 0.05, 0.2, 0.5
    isolated = int(np.flatnonzero(np.diff(ref.rowstarts.astype(np.int64)) == 0)[0])
    some = np.array([roots[20], roots[20], isolated, roots[17]], np.uint64)
    say("sssp batch repeats", scale, some.tolist())
    sssp_batch_check(arena, oracle, g, ref, s, some)


@pytest.mark.parametrize("scale", [10, 12, 16])
def test_sssp_batch_lags(arena, oracle, scale):
    """Per-root bucket lags order the work only: with the roots' hub
    distances, random offsets (lags of up to 100 buckets), offsets without a
    value (negative, NaN, inf) and one huge offset (lags at their cap, crossed
    by min-step jumps), every root's distances equal the oracle's FP32 Dijkstra
    bit for bit and the parents pass the certificate."""
    e, w, ref = oracle_graph(scale)
    g = arena.csr(arena.upload_tuples(e, w), 1 << scale)
    roots = oracle.roots(ref)
    width = int(arena.get_param("sssp_batch_max"))
    rr = roots[:width]
    hub = arena.sssp_hub_distances(g, rr)
    deg = np.diff(ref.rowstarts.astype(np.int64))
    d_hub, _ = oracle.dijkstra_f32(ref, int(np.argmax(deg)))
    assert np.array_equal(hub.view(np.uint32), d_hub[rr.astype(np.int64)].view(np.uint32))
    odd = np.random.default_rng(scale).uniform(0.0, 2.0, width).astype(np.float32)
    odd[::5] = -1.0
    odd[1::7] = np.nan
    odd[2::9] = np.inf
    huge = hub.copy()
    huge[3] = 1.0e6
    s = arena.sssp_batch_create(g)
    for delta, wm, wl, wf in ((0.02, 0.05, 0.2, 0.5), (0.1, 0.05, 0.2, 0.4), (0.02, 0.0, 0.0, 0.0)):
        s.delta, s.mid_weight, s.long_weight, s.far_weight = delta, wm, wl, wf
        for name, offsets in (("hub", hub), ("random", odd), ("huge", huge)):
            say("sssp batch lags", scale, delta, name)
            arena.sssp_batch(g, s, rr, offsets)
            for i, root in enumerate(rr):
                d_ref, _ = oracle.dijkstra_f32(ref, int(root))
                d, p = arena.sssp_batch_result(g, s, i)
                assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32)), (delta, name, i)
                assert oracle.sssp_check(ref, int(root), d, p) == 0, (delta, name, i)
    # fewer roots, repeats and an isolated vertex, with lags
    isolated = int(np.flatnonzero(deg == 0)[0])
    some = np.array([roots[20], roots[20], isolated, roots[17]], np.uint64)
    s.delta, s.mid_weight, s.long_weight, s.far_weight = 0.1, 0.05, 0.2, 0.5
    arena.sssp_batch(g, s, some, np.array([0.9, 0.1, -1.0, 0.4], np.float32))
    for i, root in enumerate(some):
        d_ref, _ = oracle.dijkstra_f32(ref, int(root))
        d, p = arena.sssp_batch_result(g, s, i)
        assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32)), ("repeats", i)
    with pytest.raises(ValueError):
        arena.sssp_batch(g, s, some, np.zeros(3, np.float32))


@pytest.mark.parametrize("scale", [10, 12, 16])
def test_sssp_batch_pull(arena, oracle, scale):
    """Pulled far-class flushes (every far flush pulled, or those of large
    windows) give the oracle's distances bit for bit and parents that pass
    the certificate; launches without offsets push every flush."""
    e, w, ref = oracle_graph(scale)
    g = arena.csr(arena.upload_tuples(e, w), 1 << scale)
    roots = oracle.roots(ref)
    width = int(arena.get_param("sssp_batch_max"))
    rr = roots[:width]
    hub = arena.sssp_hub_distances(g, rr)
    d_ref = [oracle.dijkstra_f32(ref, int(r))[0] for r in rr]
    s = arena.sssp_batch_create(g)
    old = arena.get_param("batch_pull")
    try:
        for share in (1e-9, 0.05):
            arena.set_param("batch_pull", share)
            for delta, wm, wl, wf in ((0.02, 0.05, 0.2, 0.2), (0.1, 0.05, 0.2, 0.4), (0.02, 0.0, 0.0, 0.0)):
                s.delta, s.mid_weight, s.long_weight, s.far_weight = delta, wm, wl, wf
                for offsets in (None, hub):
                    arena.sssp_batch(g, s, rr, offsets)
                    pw = arena.sssp_batch_pull_work(s)
                    say("sssp batch pull", scale, share, delta, "lags", offsets is not None, pw)
                    if offsets is None:  # without offsets no root is known to share the hub's component
                        assert pw["pulls"] == 0
                    elif share < 1e-6 and wf > delta:
                        assert pw["pulls"] > 0
                    for i, root in enumerate(rr):
                        d, p = arena.sssp_batch_result(g, s, i)
                        assert np.array_equal(d.view(np.uint32), d_ref[i].view(np.uint32)), (share, delta, i)
                        assert oracle.sssp_check(ref, int(root), d, p) == 0, (share, delta, i)
    finally:
        arena.set_param("batch_pull", old)


def test_sssp_batch_matches_plain(arena, oracle):
    """Batched and single-root kernels give the same distances on a larger graph."""
    e, w, ref = oracle_graph(18)
    g = arena.csr(arena.upload_tuples(e, w), 1 << 18)
    sb = arena.sssp_batch_create(g)
    sp = arena.sssp_create(g)
    roots = arena.roots(g)[:16]
    plain = []
    for r in roots:
        arena.sssp(g, sp, int(r))
        plain.append(arena.sssp_result(sp)[0])
    hub = arena.sssp_hub_distances(g, roots)
    for delta, wm, wl, wf in ((0.1, 0.05, 0.2, 0.5), (0.0125, 0.05, 0.25, 0.5)):
        sb.delta, sb.mid_weight, sb.long_weight, sb.far_weight = delta, wm, wl, wf
        for offsets in (None, hub):
            arena.sssp_batch(g, sb, roots, offsets)
            flushes = arena.sssp_batch_phase_ms(sb)[5]
            say("sssp batch S18", delta, wm, wl, wf, "lags", offsets is not None, "flushes", flushes)
            assert flushes > 0
            for i, r in enumerate(roots):
                d, p = arena.sssp_batch_result(g, sb, i)
                assert np.array_equal(d.view(np.uint32), plain[i].view(np.uint32)), (delta, wm, wl, wf, i)
                assert oracle.sssp_check(ref, int(r), d, p) == 0


def test_sssp_batch_diagnostics(arena, oracle):
    """Work counters and the phase log of a batched run are consistent."""
    e, w, ref = oracle_graph(14)
    g = arena.csr(arena.upload_tuples(e, w), 1 << 14)
    s = arena.sssp_batch_create(g, 0.02)
    s.mid_weight, s.long_weight, s.far_weight = 0.05, 0.2, 0.4
    roots = arena.roots(g)[:int(arena.get_param("sssp_batch_max"))]
    arena.sssp_batch(g, s, roots)
    wk = arena.sssp_batch_work(s)
    for seg in ("light", "medium", "mid", "long", "far"):
        items, entries, pairs = wk[seg]
        assert entries >= items and pairs >= entries, (seg, wk[seg])
    assert wk["medium"][1] > 0 and wk["mid"][1] > 0 and wk["long"][1] > 0 and wk["far"][1] > 0
    kind, items, ms = arena.sssp_batch_step_log(s)
    ph = arena.sssp_batch_phase_ms(s)
    b, li, imp = arena.sssp_batch_stats(s)
    assert len(kind) > 0 and set(kind) <= {0, 1, 2, 3, 4}
    assert kind.count(1) == b and kind.count(2) == b  # a dense scan and a medium phase per bucket
    assert kind.count(3) == int(ph[5])               # one log entry per flush phase
    assert kind.count(0) == li and imp > 0
    assert abs(sum(ms) - (ph[0] + ph[1] + ph[2])) < 0.05 * ph[3] + 0.1


def test_sssp_batch_errors_return(arena):
    import g500
    e, w, ref = oracle_graph(10)
    g = arena.csr(arena.upload_tuples(e, w), 1 << 10)
    s = arena.sssp_batch_create(g)
    width = int(arena.get_param("sssp_batch_max"))
    for roots in ([], list(range(width + 1)), [1 << 10]):
        with pytest.raises(g500.LibraryError):
            arena.sssp_batch(g, s, np.array(roots, np.uint64))
    for wl in (-1.0, float("nan"), float("inf")):
        s.long_weight = wl
        with pytest.raises(g500.LibraryError):
            arena.sssp_batch(g, s, np.array([1], np.uint64))
    s.long_weight = 0.2
    for wm in (-1.0, float("nan"), float("inf")):
        s.mid_weight = wm
        with pytest.raises(g500.LibraryError):
            arena.sssp_batch(g, s, np.array([1], np.uint64))
    s.mid_weight = 0.05
    for wf in (-1.0, float("nan"), float("inf")):
        s.far_weight = wf
        with pytest.raises(g500.LibraryError):
            arena.sssp_batch(g, s, np.array([1], np.uint64))


# ------------------------------------------------------------- knobs / tuner
@pytest.mark.parametrize("block", [128, 256])
def test_block_shapes(arena, oracle, block):
    """Persistent kernels with smaller blocks (more blocks in the grid barrier)."""
    e, w, ref = oracle_graph(13)
    g = check_csr(arena, oracle, 1 << 13, e, w)
    knobs = ("bfs_block", "sssp_block", "batch_bfs_block", "batch_sssp_block")
    old = {k: arena.get_param(k) for k in knobs}
    try:
        for k in knobs:
            arena.set_param(k, block)
        arena.bfs_state = arena.bfs_create(g)
        s = arena.sssp_create(g)
        roots = oracle.roots(ref)
        for r in roots[:2]:
            bfs_check(arena, oracle, g, ref, int(r))
            arena.sssp(g, s, int(r))
            d, p = arena.sssp_result(s)
            d_ref, _ = oracle.dijkstra_f32(ref, int(r))
            assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32))
        batch_check(arena, oracle, g, ref, arena.batch_create(g), roots)
        sssp_batch_check(arena, oracle, g, ref, arena.sssp_batch_create(g), roots[:16])
    finally:
        for k, v in old.items():
            arena.set_param(k, v)
    import g500
    with pytest.raises(KeyError):
        arena.set_param("bfs_block", 300)


def test_autotuner_applies_candidates(arena, oracle):
    from g500 import autotune
    e, w, ref = oracle_graph(14)
    g = check_csr(arena, oracle, 1 << 14, e, w)
    keys = ("bfs_alpha", "bfs_beta", "bfs_block", "sssp_block", "sssp_pull", "batch_alpha", "batch_heavy_degree",
            "batch_bfs_block", "batch_sssp_block")
    old = {k: arena.get_param(k) for k in keys}
    try:
        out = autotune.tune(arena, g, batch=True)
        p = out["params"]
        assert (p["bfs_alpha"], p["bfs_beta"]) in autotune.BFS_SWITCH and p["bfs_block"] in autotune.BLOCKS
        assert p["sssp_delta"] in autotune.SSSP_DELTA and p["sssp_pull"] in autotune.SSSP_PULL
        assert p["batch_bfs_block"] in autotune.BLOCKS and p["batch_sssp_block"] in autotune.BLOCKS
        assert (p["batch_sssp_delta"], p["batch_sssp_mid_weight"], p["batch_sssp_long_weight"],
                p["batch_sssp_far_weight"]) in autotune.BATCH_SSSP
        assert p["batch_alpha"] in autotune.BATCH_ALPHA and p["batch_heavy_degree"] in autotune.BATCH_HEAVY
        for k in keys:
            assert arena.get_param(k) == p[k]
        roots = oracle.roots(ref)
        arena.bfs_state = arena.bfs_create(g)
        bfs_check(arena, oracle, g, ref, int(roots[0]))
        sb = arena.sssp_batch_create(g, p["batch_sssp_delta"])
        sb.mid_weight, sb.long_weight = p["batch_sssp_mid_weight"], p["batch_sssp_long_weight"]
        sb.far_weight = p["batch_sssp_far_weight"]
        sssp_batch_check(arena, oracle, g, ref, sb, roots[:16])
    finally:
        for k, v in old.items():
            arena.set_param(k, v)


# --------------------------------------------------------------------- runner
def test_runner_small():
    import g500
    env = dict(os.environ, G500_ARENA_BYTES=str(8 << 30))
    r = subprocess.run([g500.RUNNER_PATH, "10", "16"], capture_output=True, timeout=300, env=env)
    out = r.stdout.decode()
    say(out[-2000:], r.stderr.decode()[-2000:])
    assert r.returncode == 0
    kv = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip()] = v.replace("!", "").strip()
    assert kv["SCALE"] == "10" and kv["NBFS"] == "64"
    assert float(kv["bfs  harmonic_mean_TEPS"]) > 0
    assert float(kv["sssp median_TEPS"]) > 0


def test_plain_launch_path(arena, oracle):
    """Persistent kernels launched without the cooperative API (the fallback
    used when another process holds the device's cooperative queue)."""
    e, w, ref = oracle_graph(12)
    g = check_csr(arena, oracle, 1 << 12, e, w)
    b = arena.bfs_create(g)
    s = arena.sssp_create(g)
    arena.set_param("coop_launch", 0)
    try:
        for r in oracle.roots(ref)[:3]:
            r = int(r)
            arena.bfs(g, b, r)
            pred = arena.bfs_pred(b)
            assert oracle.bfs_validate(ref, pred, r) == 0
            arena.sssp(g, s, r)
            d, p = arena.sssp_result(s)
            d_ref, _ = oracle.dijkstra_f32(ref, r)
            assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32))
    finally:
        arena.set_param("coop_launch", 1)


# ------------------------------------------------------------ tuple dump / load
def _dump(sess, tg, start, path):
    os.environ["G500_DUMP_PATH"] = path
    try:
        sess.lib.tuple_graph_dump(C.byref(tg), int(start))
        sess.check()
    finally:
        del os.environ["G500_DUMP_PATH"]


def _load(sess, start, count, path):
    import g500
    tg = g500.TupleGraph()
    os.environ["G500_DUMP_PATH"] = path
    try:
        sess.lib.tuple_graph_load(C.byref(sess.ws), C.byref(tg), int(start), int(count))
        sess.check()
    finally:
        del os.environ["G500_DUMP_PATH"]
    return tg


def _file_tuples(path, a, b):
    e = np.fromfile(path, dtype="<u8", count=2 * (b - a), offset=16 * a).reshape(-1, 2)
    w = np.fromfile(path + ".weights", dtype="<f4", count=b - a, offset=4 * a)
    return e, w


def test_tuple_dump_load_roundtrip(arena, oracle, tmp_path):
    """graph_generation.cu:314-396: raw little-endian {u64 v0, u64 v1} per edge
    in <path>, f32 per edge in <path>.weights, each file nglobaledges long, a
    range written at its start_edge offset. Two ranges dumped into the same
    files (as two ranks would) give the whole list; loads return the bytes."""
    scale, ef = 12, 16
    E = ef << scale
    eo, wo = oracle.kronecker(scale, ef)
    path = str(tmp_path / "s12.bin")
    half = E // 2 + 123
    for a, b in ((0, half), (half, E)):
        tg = arena.generate(scale, ef, start=a, end=b)
        _dump(arena, tg, a, path)
    assert os.path.getsize(path) == 16 * E and os.path.getsize(path + ".weights") == 4 * E
    e, w = _file_tuples(path, 0, E)
    assert np.array_equal(e, eo) and np.array_equal(w.view(np.uint32), wo.view(np.uint32))
    for a, b in ((0, E), (777, 3001), (E - 5, E)):
        tg = _load(arena, a, b - a, path)
        assert tg.edges_size == b - a and tg.max_edges_size == b - a
        e, w = arena.tuples_to_host(tg)
        assert np.array_equal(e, eo[a:b]) and np.array_equal(w.view(np.uint32), wo[a:b].view(np.uint32))
    # a range dumped alone: the rest of the files reads as zeros
    path2 = str(tmp_path / "range.bin")
    tg = arena.generate(scale, ef, start=1000, end=5000)
    _dump(arena, tg, 1000, path2)
    e, w = _file_tuples(path2, 0, E)
    assert np.array_equal(e[1000:5000], eo[1000:5000]) and not e[:1000].any() and not e[5000:].any()
    assert np.array_equal(w[1000:5000].view(np.uint32), wo[1000:5000].view(np.uint32)) and not w[5000:].any()
    # a graph loaded from the files builds the oracle's CSR
    tg = _load(arena, 0, E, path)
    tg.nglobaledges = E
    g = arena.csr(tg, 1 << scale)
    rs, col, _ = arena.csr_to_host(g)
    ref = oracle_graph(scale)[2]
    assert np.array_equal(rs, ref.rowstarts) and np.array_equal(col, ref.column)
    with pytest.raises(Exception):
        _load(arena, 0, 10, str(tmp_path / "missing.bin"))


def test_tuple_dump_load_past_2gib(arena, oracle, tmp_path):
    """A range whose byte count passes 2^31 (the reference casts the count
    to int, graph_generation.cu:340): written and read whole."""
    import shutil
    scale, ef = 24, 16
    cnt = (1 << 27) + 4096                     # 2 GiB + 64 KiB of edges
    if shutil.disk_usage(str(tmp_path)).free < 8 << 30:
        pytest.skip("needs 8 GiB of free space in the temporary directory")
    path = str(tmp_path / "big.bin")
    tg = arena.generate(scale, ef, start=0, end=cnt)
    _dump(arena, tg, 0, path)
    assert os.path.getsize(path) == 16 * (ef << scale)
    for a, b in ((0, 4096), (cnt - 8192, cnt)):
        e, w = _file_tuples(path, a, b)
        eo, wo = oracle.kronecker(scale, ef, start=a, end=b)
        assert np.array_equal(e, eo) and np.array_equal(w.view(np.uint32), wo.view(np.uint32))
    m = arena.mark()
    back = _load(arena, 0, cnt, path)
    for a, b in ((0, 4096), ((1 << 27) - 100, cnt)):
        e = arena.to_host(back.edges + 16 * a, 2 * (b - a), np.uint64).reshape(-1, 2)
        w = arena.to_host(back.weights + 4 * a, b - a, np.float32)
        eo, wo = oracle.kronecker(scale, ef, start=a, end=b)
        assert np.array_equal(e, eo) and np.array_equal(w.view(np.uint32), wo.view(np.uint32))
    arena.release(m)
    os.remove(path)
    os.remove(path + ".weights")


# ------------------------------------------------------------- certificates
def _dev(sess, arr):
    import g500
    arr = np.ascontiguousarray(arr)
    p = sess.lib._worker_arena_push_array(C.byref(sess.ws), arr.size, arr.itemsize, 256, C.byref(g500.u64()))
    assert p
    sess.to_device(p, arr)
    return p


def test_depth_check_matches_oracle(arena, oracle):
    """The device depth certificate counts what the oracle's counts: zero on
    the BFS depths (plain and batched), the same count on planted mistakes."""
    e, w, ref = oracle_graph(12)
    g = check_csr(arena, oracle, 1 << 12, e, w)
    roots = oracle.roots(ref)[:4]
    bb = arena.batch_create(g)
    arena.batch(g, bb, roots)
    rng = np.random.default_rng(2)
    for i, r in enumerate(roots):
        r = int(r)
        _, truth = oracle.bfs(ref, r)
        cases = [truth]
        for _ in range(12):
            d = truth.copy()
            vs = rng.choice(ref.n, int(rng.integers(1, 4)), replace=False)
            d[vs] = rng.integers(-3, 8, len(vs))
            cases.append(d)
        for d in cases:
            m = arena.mark()
            assert arena.depth_check(g, _dev(arena, d), r) == oracle.depth_check(ref, d, r)
            arena.release(m)
        assert arena.depth_check(g, _dev(arena, arena.batch_depth(g, bb, i)), r) == 0


def test_sssp_check_matches_oracle(arena, oracle):
    """The device SSSP certificate counts what the oracle's counts: zero on
    FP32 Dijkstra results, the same count on planted mistakes, parent cycles
    through zero-gain entries included."""
    e, w, ref = oracle_graph(12)
    g = check_csr(arena, oracle, 1 << 12, e, w)
    rng = np.random.default_rng(4)
    for r in oracle.roots(ref)[:3]:
        r = int(r)
        d0, p0 = oracle.dijkstra_f32(ref, r)
        reached = np.flatnonzero((d0 > 0))
        cases = [(d0, p0)]
        for kind in range(7):
            d, p = d0.copy(), p0.copy()
            v = int(rng.choice(reached))
            if kind == 0: d[v] = np.nextafter(d[v], np.float32(0))
            if kind == 1: d[v] = np.nextafter(d[v], np.float32(9))
            if kind == 2: p[v] = v
            if kind == 3: p[v] = ref.n + 3
            if kind == 4: d[v], p[v] = -1.0, -1
            if kind == 5: p[r] = v
            if kind == 6: p[int(rng.choice(reached))] = v; p[v] = int(rng.choice(reached))
            cases.append((d, p))
        for d, p in cases:
            m = arena.mark()
            got = arena.sssp_check(g, _dev(arena, d), _dev(arena, p), r)
            arena.release(m)
            assert got == oracle.sssp_check(ref, r, d, p), r
    # zero-weight entries: tight both ways at equal distance
    e2 = np.array([[0, 1], [1, 2], [2, 3], [0, 3]], np.uint64)
    w2 = np.array([0.5, 0.0, 0.25, 2.0], np.float32)
    g2 = check_csr(arena, oracle, 4, e2, w2)
    ref2 = oracle.csr(4, e2, w2)
    d = np.array([0.0, 0.5, 0.5, 0.75], np.float32)
    for p in ([0, 0, 1, 2], [0, 2, 1, 2], [0, 0, 1, 0]):
        p = np.array(p, np.int64)
        m = arena.mark()
        got = arena.sssp_check(g2, _dev(arena, d), _dev(arena, p), 0)
        arena.release(m)
        assert got == oracle.sssp_check(ref2, 0, d, p), p
    assert oracle.sssp_check(ref2, 0, d, np.array([0, 2, 1, 2], np.int64)) == 3  # the 1 <-> 2 cycle and 3 below it


def test_sssp_root_order(arena, oracle):
    """oned_graph_sssp_root_order sorts roots by their FP32 distance from the
    highest-degree vertex (smallest id among ties), unreached roots first."""
    e, w, ref = oracle_graph(12)
    g = check_csr(arena, oracle, 1 << 12, e, w)
    deg = np.diff(ref.rowstarts.astype(np.int64))
    hub = int(np.flatnonzero(deg == deg.max())[0])
    d, _ = oracle.dijkstra_f32(ref, hub)
    roots = np.concatenate([oracle.roots(ref), np.flatnonzero(d < 0)[:2].astype(np.uint64)])
    order = arena.sssp_root_order(g, roots)
    assert sorted(order.tolist()) == list(range(len(roots)))
    key = d[roots.astype(np.int64)]
    assert np.array_equal(order, np.argsort(key, kind="stable"))
    assert len(arena.sssp_root_order(g, roots[:0])) == 0
    with pytest.raises(Exception):
        arena.sssp_root_order(g, np.array([1 << 12], np.uint64))
