"""Pins the oracle's generator, CSR, BFS, validation, SSSP, roots and
statistics by invariants and independent library routines (numpy, scipy)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

from conftest import oracle_graph

P = 2**31 - 1


# ---------------------------------------------------------------- generator
def test_vertices_in_range_and_count(oracle):
    for scale in (1, 5, 10):
        e, w = oracle.kronecker(scale, 16)
        assert e.shape == (16 << scale, 2)
        assert int(e.max()) < (1 << scale)


@pytest.mark.parametrize("scale", [1, 7, 16])
def test_scramble_is_bijection(oracle, scale):
    v0, v1 = oracle.scramble_values(oracle.mrg_seed(2, 3))
    img = [oracle.lib().o_scramble(v, scale, v0, v1) for v in range(1 << scale)]
    assert sorted(img) == list(range(1 << scale))


def test_unscrambled_endpoints_ordered_and_quadrants(oracle):
    scale = 12
    v0, v1 = oracle.scramble_values(oracle.mrg_seed(2, 3))
    inv = np.zeros(1 << scale, np.uint64)
    for v in range(1 << scale):
        inv[oracle.lib().o_scramble(v, scale, v0, v1)] = v
    e, _ = oracle.kronecker(scale, 16)
    src, tgt = inv[e[:, 0]], inv[e[:, 1]]
    assert np.all(src <= tgt)  # the diagonal-block swap (src offset <= tgt offset)
    top = scale - 1
    a = (src >> np.uint64(top)) & np.uint64(1)
    b = (tgt >> np.uint64(top)) & np.uint64(1)
    frac00 = np.mean((a == 0) & (b == 0))
    frac01 = np.mean((a == 0) & (b == 1))
    frac11 = np.mean((a == 1) & (b == 1))
    assert np.all(~((a == 1) & (b == 0)))
    assert abs(frac00 - 0.57) < 0.01 and abs(frac01 - 0.38) < 0.01 and abs(frac11 - 0.05) < 0.005


def test_weights_are_scaled_f32_of_mrg_outputs(oracle):
    _, w = oracle.kronecker(10, 16)
    u = w.astype(np.float64) * 2.0**31
    assert np.all(u == np.floor(u)) and np.all(w >= 0) and np.all(w <= 1.0)
    assert abs(float(w.mean()) - 0.5) < 0.01


def test_generator_range_split_is_consistent(oracle):
    e, w = oracle.kronecker(10, 16)
    e2, w2 = oracle.kronecker(10, 16, start=1000, end=9000)
    assert np.array_equal(e[1000:9000], e2) and np.array_equal(w[1000:9000].view(np.uint32), w2.view(np.uint32))


def test_single_edge_second_restatement(oracle):
    """A Python restatement of one edge (matrix-power skip, big-integer MRG,
    Python scramble values and scramble)."""
    scale = 10
    seed = [3, 1, 4, 1, 1]
    from test_oracle_mrg import apply, mat_pow, py_steps
    from test_oracle_pins import py_scramble, py_scramble_values
    v0, v1 = py_scramble_values(seed)
    e, w = oracle.kronecker(scale, 16, start=0, end=300)
    for ei in (0, 1, 255, 256, 299):
        z = apply(mat_pow(ei * 2**64), seed)
        half, s, t = 1 << scale, 0, 0
        while half > 1:
            while True:
                z, (u,) = py_steps(z, 1)
                if u >= P % 10000:
                    break
            u %= 10000
            sq = 1 if u < 1900 else 2 if u < 3800 else 0 if u < 9500 else 3
            so, to = sq // 2, sq % 2
            if s == t and so > to:
                so, to = to, so
            half //= 2
            s += half * so
            t += half * to
        assert py_scramble(s, scale, v0, v1) == int(e[ei, 0])
        assert py_scramble(t, scale, v0, v1) == int(e[ei, 1])
        z, (u,) = py_steps(z, 1)
        assert np.float32(w[ei]) == np.float32(np.float32(u) * np.float32(2.0**-31))


# ---------------------------------------------------------------------- CSR
def numpy_csr(n, e, w):
    keep = e[:, 0] != e[:, 1]
    a, b, ww = e[keep, 0], e[keep, 1], w[keep]
    src = np.concatenate([a, b]).astype(np.int64)
    dst = np.concatenate([b, a]).astype(np.uint32)
    wb = np.concatenate([ww, ww]).view(np.uint32)
    order = np.lexsort((dst, wb, src))
    rs = np.zeros(n + 1, np.uint64)
    rs[1:] = np.cumsum(np.bincount(src, minlength=n))
    return rs, dst[order], wb[order].view(np.float32)


@pytest.mark.parametrize("scale,ef", [(1, 16), (6, 4), (10, 16), (12, 16)])
def test_csr_matches_numpy(oracle, scale, ef):
    e, w, g = oracle_graph(scale, ef)
    rs, col, cw = numpy_csr(1 << scale, e, w)
    assert np.array_equal(g.rowstarts, rs)
    assert np.array_equal(g.column, col)
    assert np.array_equal(g.weights.view(np.uint32), cw.view(np.uint32))


# ---------------------------------------------------------------------- BFS
def scipy_graph(g):
    n = g.n
    rows = np.repeat(np.arange(n), np.diff(g.rowstarts.astype(np.int64)))
    return sp.csr_matrix((np.ones(g.m), (rows, g.column.astype(np.int64))), shape=(n, n))


def test_bfs_depths_match_scipy(oracle, g12):
    _, _, g = g12
    A = scipy_graph(g)
    for root in oracle.roots(g)[:8]:
        pred, depth = oracle.bfs(g, int(root))
        ref = cg.shortest_path(A, unweighted=True, indices=int(root))
        ref = np.where(np.isinf(ref), -1, ref).astype(np.int64)
        assert np.array_equal(depth, ref)
        assert np.array_equal(oracle.dist_from_pred(g, pred, int(root)), ref)
        assert oracle.bfs_validate(g, pred, int(root)) == 0
        pomp = oracle.bfs_omp(g, int(root), 4)
        assert oracle.bfs_validate(g, pomp, int(root)) == 0
        assert np.array_equal(oracle.dist_from_pred(g, pomp, int(root)), ref)


def test_validation_catches_planted_mistakes(oracle, g12):
    _, _, g = g12
    root = int(oracle.roots(g)[0])
    pred, depth = oracle.bfs(g, root)
    reached = np.flatnonzero(pred >= 0)
    v = int(reached[reached != root][5])
    unreached = np.flatnonzero(pred == -1)
    cases = []
    p = pred.copy(); p[v] = g.n + 5; cases.append(p)                # rule 1
    p = pred.copy(); p[root] = -1; cases.append(p)                  # rule 2
    p = pred.copy(); p[v] = -1; cases.append(p)                     # rule 4
    if len(unreached):
        p = pred.copy(); p[int(unreached[0])] = root; cases.append(p)  # rule 6/7
    nonnb = [u for u in range(g.n) if u != v and u not in set(g.column[g.rowstarts[v]:g.rowstarts[v + 1]].tolist()) and pred[u] >= 0]
    p = pred.copy(); p[v] = nonnb[0]; cases.append(p)               # rule 6/7 (non-edge)
    p = pred.copy(); p[v] = v; cases.append(p)                      # self parent
    for p in cases:
        assert oracle.bfs_validate(g, p, root) > 0


def test_dist_from_pred_unresolved_marker(oracle, g10):
    _, _, g = g10
    root = int(oracle.roots(g)[0])
    pred, _ = oracle.bfs(g, root)
    v = int(np.flatnonzero((pred >= 0) & (np.arange(g.n) != root))[3])
    p = pred.copy()
    p[v] = v
    d = oracle.dist_from_pred(g, p, root)
    assert d[v] == np.iinfo(np.int64).max


def test_edge_count(oracle, g12):
    e, _, g = g12
    root = int(oracle.roots(g)[0])
    pred, _ = oracle.bfs(g, root)
    r = pred >= 0
    keep = (e[:, 0] != e[:, 1]) & r[e[:, 0].astype(np.int64)]
    assert oracle.edge_count(g, pred) == int(keep.sum())


# --------------------------------------------------------------------- SSSP
def numpy_certificate(g, root, dist, pred):
    """Violations of: d[v] <= fl32(d[u]+w) on every edge; a tight parent edge
    for every reached non-root vertex; parent chains reach the root."""
    n = g.n
    src = np.repeat(np.arange(n), np.diff(g.rowstarts.astype(np.int64)))
    dst = g.column.astype(np.int64)
    reached = dist >= 0
    du = dist[src]
    cand = (du + g.weights).astype(np.float32)
    ru = reached[src]
    bad = np.sum(ru & (~reached[dst] | (dist[dst] > cand)))
    p = pred.copy()
    nr = reached & (np.arange(n) != root)
    bad += np.sum(nr & ((p < 0) | (p >= n)))
    # tight: entry (v -> pred[v]) with d[pred]+w == d[v]
    pv = np.where(p >= 0, p, 0)
    tight_entry = (dst == pv[src]) & nr[src] & ((dist[dst] + g.weights).astype(np.float32) == dist[src])
    has = np.zeros(n, bool)
    has[src[tight_entry]] = True
    bad += np.sum(nr & ~has)
    anc = np.where(reached, p, -1)
    anc[root] = root
    anc = np.where(anc < 0, root, anc)
    for _ in range(int(np.ceil(np.log2(max(n, 2)))) + 2):
        anc = anc[anc]
    bad += np.sum(reached & (anc != root))
    bad += np.sum(~reached & ((dist != -1.0) | (pred != -1)))
    return int(bad)


def test_dijkstra_f64_matches_scipy(oracle, g12):
    _, _, g = g12
    n = g.n
    rows = np.repeat(np.arange(n), np.diff(g.rowstarts.astype(np.int64)))
    # scipy sums the f64 values of the f32 weights; duplicate entries keep their
    # minimum through the (row, col) -> min reduction below.
    import pandas as pd
    df = pd.DataFrame({"r": rows, "c": g.column.astype(np.int64), "w": g.weights.astype(np.float64)})
    df = df.groupby(["r", "c"], as_index=False)["w"].min()
    # zero weights would be dropped by scipy's sparse format; nudge check
    assert (df.w > 0).all()
    A = sp.csr_matrix((df.w.values, (df.r.values, df.c.values)), shape=(n, n))
    for root in oracle.roots(g)[:6]:
        ref = cg.dijkstra(A, indices=int(root))
        ref = np.where(np.isinf(ref), -1.0, ref)
        assert np.array_equal(oracle.dijkstra_f64(g, int(root)), ref)


def test_dijkstra_f32_certificate_and_delta_stepping(oracle, g12):
    _, _, g = g12
    for root in oracle.roots(g)[:6]:
        root = int(root)
        d, p = oracle.dijkstra_f32(g, root)
        assert numpy_certificate(g, root, d, p) == 0
        assert oracle.sssp_check(g, root, d, p) == 0
        for delta in (0.05, 0.1, 0.25, 2.0):
            d2, p2 = oracle.delta_stepping_f32(g, root, delta)
            assert np.array_equal(d.view(np.uint32), d2.view(np.uint32))
            assert oracle.sssp_check(g, root, d2, p2) == 0
        d64 = oracle.dijkstra_f64(g, root)
        assert np.array_equal(d >= 0, d64 >= 0)
        assert float(np.max(np.abs(d.astype(np.float64) - d64))) < 1e-5


def test_sssp_certificate_catches_mistakes(oracle, g10):
    _, _, g = g10
    root = int(oracle.roots(g)[0])
    d, p = oracle.dijkstra_f32(g, root)
    v = int(np.flatnonzero((d > 0))[7])
    bad = d.copy(); bad[v] = np.nextafter(bad[v], np.float32(0))
    assert oracle.sssp_check(g, root, bad, p) > 0 and numpy_certificate(g, root, bad, p) > 0
    bad = d.copy(); bad[v] = np.nextafter(bad[v], np.float32(9))
    assert oracle.sssp_check(g, root, bad, p) > 0 and numpy_certificate(g, root, bad, p) > 0
    pp = p.copy(); pp[v] = v
    assert oracle.sssp_check(g, root, d, pp) > 0 and numpy_certificate(g, root, d, pp) > 0


# ------------------------------------------------------------------- roots
def test_roots(oracle, g12):
    _, _, g = g12
    r = oracle.roots(g)
    assert len(r) == 64 and len(set(r.tolist())) == 64
    deg = np.diff(g.rowstarts.astype(np.int64))
    assert np.all(deg[r.astype(np.int64)] > 0)
    d = oracle.random_numbers(2, 2, 3, 0)
    first = int((d[0] + d[1]) * g.n) % g.n
    if deg[first] > 0:
        assert int(r[0]) == first


def test_roots_tiny_graph_exhausts_counter(oracle):
    e, w, g = oracle_graph(1, 1)
    r = oracle.roots(g)
    assert len(r) <= 2


# -------------------------------------------------------------- statistics
@pytest.mark.parametrize("n", [5, 64, 7])
def test_statistics(oracle, n):
    x = np.random.default_rng(n).random(n)
    s = oracle.statistics(x)
    assert s[0] == x.min() and s[4] == x.max()
    assert np.isclose(s[2], np.median(x))
    assert np.isclose(s[5], x.mean()) and np.isclose(s[6], x.std(ddof=1))
    if n in (5, 64):
        assert np.isclose(s[1], np.percentile(x, 25, method="midpoint"))
        assert np.isclose(s[3], np.percentile(x, 75, method="midpoint"))


# ------------------------------------------------------- two-hop pruned SSSP
@pytest.mark.parametrize("scale,hubset", [(10, "top"), (12, "top"), (12, "random")])
def test_two_hop_pruning_keeps_f32_distances(oracle, scale, hubset):
    """Dropping the entries that a two-hop path through a hub beats by more
    than 2^-10 leaves every f32 Dijkstra distance bit-identical, for any hub
    set; the results still pass the certificate on the full graph."""
    from conftest import drop_entries, prune_rule
    e, w, g = oracle_graph(scale)
    deg = np.diff(g.rowstarts.astype(np.int64))
    if hubset == "top":
        hubs = np.argsort(-deg, kind="stable")[:32]
    else:
        hubs = np.random.default_rng(5).choice(g.n, 32, replace=False)
    drop = prune_rule(g.rowstarts, g.column, g.weights, hubs)
    assert drop.sum() > (0.2 * g.m if hubset == "top" else 0)
    rs_k, col_k, w_k = drop_entries(g.rowstarts, g.column, g.weights, drop)
    pruned = oracle.Csr(g.n, rs_k, col_k.astype(np.uint32), w_k.astype(np.float32))
    for r in oracle.roots(g)[:6]:
        d_full, _ = oracle.dijkstra_f32(g, int(r))
        d_pr, p_pr = oracle.dijkstra_f32(pruned, int(r))
        assert np.array_equal(d_full.view(np.uint32), d_pr.view(np.uint32))
        assert oracle.sssp_check(g, int(r), d_pr, p_pr) == 0


def test_two_hop_rule_pins(oracle):
    """Hand cases of the rule (dyadic weights, exact in f32 and f64): the path
    1-0-2 through hub 0 (0.125 + 0.25) beats a direct entry 1-2 only by more
    than the margin 2^-10, and never an entry of weight >= 2."""
    from conftest import prune_rule
    m = 2.0 ** -10
    edges = [(1, 0, 0.125), (2, 0, 0.25),
             (1, 2, 0.375 + m),            # beaten by exactly the margin: kept
             (1, 2, 0.375 + m + 2 ** -20),  # beaten by more: dropped
             (1, 2, 0.5),                  # dropped
             (1, 2, 2.5)]                  # w >= 2: kept
    n = 3
    rows = [[] for _ in range(n)]
    for u, v, x in edges:
        rows[u].append((x, v))
        rows[v].append((x, u))
    rs = np.zeros(n + 1, np.int64)
    col, wt = [], []
    for u in range(n):
        for x, v in sorted(rows[u]):
            col.append(v)
            wt.append(x)
        rs[u + 1] = len(col)
    wt = np.array(wt, np.float32)
    assert np.array_equal(wt.astype(np.float64), np.array([x for u in range(n) for x, _ in sorted(rows[u])]))
    drop = prune_rule(rs, np.array(col), wt, [0])
    dropped = sorted(float(wt[k]) for u in range(n) for k in range(rs[u], rs[u + 1]) if drop[k] and u == 1)
    assert dropped == [0.375 + m + 2 ** -20, 0.5]
    # the reverse entries (row 2) get the same decisions
    assert sorted(float(wt[k]) for k in range(rs[2], rs[3]) if drop[k]) == dropped


# ------------------------------------------- checkers for the large graphs
@pytest.mark.parametrize("scale", [10, 12, 14])
def test_bfs_depth_omp_equals_serial(oracle, scale):
    _, _, g = oracle_graph(scale)
    for r in oracle.roots(g)[:6]:
        _, depth = oracle.bfs(g, int(r))
        for nt in (1, 3, 8):
            assert np.array_equal(oracle.bfs_depth_omp(g, int(r), nt), depth)


@pytest.mark.parametrize("scale", [10, 13])
def test_edge_digests_agree_and_catch_mistakes(oracle, scale):
    e, w, g = oracle_graph(scale)
    d, cnt = oracle.tuple_digest(e, w)
    assert cnt == g.m and d == oracle.csr_digest(g)
    assert np.array_equal(oracle.rowstarts_of(g.n, e), g.rowstarts)
    rng = np.random.default_rng(scale)
    k = int(rng.integers(0, g.m))
    col = g.column.copy()
    col[k] = (col[k] + 1) % g.n                                        # one wrong column
    assert oracle.csr_digest(oracle.Csr(g.n, g.rowstarts, col, g.weights)) != d
    wt = g.weights.copy()
    wt.view(np.uint32)[k] ^= 1                                         # one weight ulp off
    assert oracle.csr_digest(oracle.Csr(g.n, g.rowstarts, g.column, wt)) != d
    u = int(np.searchsorted(g.rowstarts.astype(np.int64), k, side="right") - 1)
    rs = g.rowstarts.copy()
    rs[u + 1:] -= np.uint64(1)                                         # entry k dropped
    keep = np.ones(g.m, bool)
    keep[k] = False
    assert oracle.csr_digest(oracle.Csr(g.n, rs, g.column[keep], g.weights[keep])) != d
    e2 = e.copy()
    e2[0] = e2[0][::-1]                                                # endpoints of one edge swapped: same multiset
    assert oracle.tuple_digest(e2, w)[0] == d


def test_depth_certificate_is_exact(oracle, g10):
    """o_depth_check is zero on the BFS depths and only there: every depth
    array over {-1..4} of a 6-vertex graph (with an isolated vertex) is
    enumerated; on SCALE 10 the scipy depths pass and single-vertex
    perturbations are all caught."""
    import itertools
    e = np.array([[0, 1], [1, 2], [2, 3], [0, 3], [3, 4], [1, 1]], np.uint64)
    g = oracle.csr(6, e, np.full(len(e), 0.5, np.float32))
    adj = sp.csr_matrix((np.ones(g.m), g.column.astype(np.int64), g.rowstarts.astype(np.int64)), shape=(6, 6))
    for root in (0, 3, 5):
        sd = cg.shortest_path(adj, unweighted=True, indices=root)
        truth = np.where(np.isinf(sd), -1, sd).astype(np.int64)
        passing = [d for d in itertools.product(range(-1, 5), repeat=6)
                   if oracle.depth_check(g, np.array(d, np.int64), root) == 0]
        assert passing == [tuple(truth)], root
    _, _, g = g10
    adj = sp.csr_matrix((np.ones(g.m), g.column.astype(np.int64), g.rowstarts.astype(np.int64)), shape=(g.n, g.n))
    rng = np.random.default_rng(3)
    for root in oracle.roots(g)[:3]:
        root = int(root)
        sd = cg.shortest_path(adj, unweighted=True, indices=root)
        truth = np.where(np.isinf(sd), -1, sd).astype(np.int64)
        assert oracle.depth_check(g, truth, root) == 0
        for v in rng.choice(g.n, 60, replace=False):
            for dv in (-2, -1, 0, 1, truth[v] - 1, truth[v] + 1, truth[v] + 2):
                if dv == truth[v]:
                    continue
                d = truth.copy()
                d[v] = dv
                assert oracle.depth_check(g, d, root) > 0, (root, v, dv)


@pytest.mark.parametrize("scale", [10, 13])
def test_chunked_degrees_and_row_order(oracle, scale):
    """Degrees added chunk by chunk equal numpy's bincount of the non-loop
    endpoints (and so the CSR's rowstarts); the oracle CSR has no order
    violations, and each planted one is found."""
    e, w, g = oracle_graph(scale)
    n = g.n
    deg = np.zeros(n, np.uint64)
    for a in range(0, len(e), 3001):                                   # ragged chunks
        assert oracle.degree_add(n, e[a:a + 3001], deg) == 0
    keep = e[:, 0] != e[:, 1]
    ref = np.bincount(e[keep].reshape(-1).astype(np.int64), minlength=n)
    assert np.array_equal(deg, ref.astype(np.uint64))
    assert np.array_equal(np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64), g.rowstarts)
    bad = e[:5].copy()
    bad[2, 1] = n                                                      # endpoint out of range
    assert oracle.degree_add(n, bad, np.zeros(n, np.uint64)) == 1
    assert oracle.csr_order_errors(g) == 0
    rs = g.rowstarts.astype(np.int64)
    u = int(np.argmax(np.diff(rs)))                                    # the longest row
    b = rs[u]
    col = g.column.copy()
    col[[b, b + 1]] = col[[b + 1, b]]                                  # two entries swapped
    wt = g.weights.copy()
    wt[[b, b + 1]] = wt[[b + 1, b]]
    same = g.weights[b] == g.weights[b + 1] and g.column[b] == g.column[b + 1]
    assert oracle.csr_order_errors(oracle.Csr(n, g.rowstarts, col, wt)) == (0 if same else 1)
    wt2 = g.weights.copy()
    wt2[b] = np.float32(2.0)                                           # a heavier first entry
    assert oracle.csr_order_errors(oracle.Csr(n, g.rowstarts, g.column, wt2)) == 1
    rs2 = g.rowstarts.copy()
    rs2[u + 1] = rs2[u] - np.uint64(1) if rs2[u] > 0 else rs2[u + 1]   # a backwards step
    assert oracle.csr_order_errors(oracle.Csr(n, rs2, g.column, g.weights)) >= (1 if rs2[u] > 0 else 0)


@pytest.mark.parametrize("scale", [12, 14])
def test_delta_stepping_omp_equals_dijkstra(oracle, scale):
    _, _, g = oracle_graph(scale)
    for r in oracle.roots(g)[:3]:
        r = int(r)
        d_ref, _ = oracle.dijkstra_f32(g, r)
        for delta in (0.003, 0.05, 0.1, 0.25, 4.0):
            for nt in (1, 8):
                d, p = oracle.delta_stepping_omp(g, r, delta, nt)
                assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32)), (r, delta, nt)
                assert oracle.sssp_check(g, r, d, p) == 0


def test_delta_stepping_omp_parents_acyclic_on_tiny_weights(oracle):
    """Vertices at equal distance joined by a weight below half an ulp are
    tight both ways; parents still form a tree (a per-vertex choice of any
    tight neighbour would close 1 <-> 2 here)."""
    e = np.array([[0, 1], [0, 2], [1, 2], [2, 3], [1, 3]], np.uint64)
    w = np.array([0.5, 0.5, 1e-9, 0.25, 0.25], np.float32)
    g = oracle.csr(4, e, w)
    for nt in (1, 4):
        d, p = oracle.delta_stepping_omp(g, 0, 0.1, nt)
        assert list(d) == [0.0, 0.5, 0.5, 0.75]
        assert oracle.sssp_check(g, 0, d, p) == 0
    bad = np.array([0, 2, 1, 1], np.int64)  # 1 <-> 2 cycle
    assert oracle.sssp_check(g, 0, d, bad) > 0
