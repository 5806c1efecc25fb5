"""Checks of device results at the workload sizes (SCALE 22-26) against the
oracle, with host memory bounded: the tuple list is streamed in chunks and
only the CSR is held on the host.

Bars (the same as at small sizes, where the oracle's own arrays are compared):
- generator: sampled edge ranges bit-exact against the oracle generator;
- CSR: rowstarts equal the prefix sum of the tuple list's degrees, every row
  in the canonical (weight bits, column) order, and the multiset of (u, v, w)
  entries equal to the one the tuple list stands for (order-free digest), so
  the arrays are the oracle's arrays;
- BFS: depths bit-exact against the oracle's BFS; parent trees pass rules 1-7;
- SSSP: the parent certificate, which fixes the FP32 distances uniquely
  (every label-correcting schedule ends at the same fixed point), plus bit
  equality between schedules (deltas, batched and plain kernels).
"""
import os

import numpy as np

U64 = (1 << 64) - 1


def say(*a):
    print(*a, flush=True)


def nthreads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def stream_tuples(sess, tg, chunk=1 << 26):
    for a in range(0, int(tg.edges_size), chunk):
        k = min(chunk, int(tg.edges_size) - a)
        e = sess.to_host(tg.edges + 16 * a, 2 * k, np.uint64).reshape(-1, 2)
        w = sess.to_host(tg.weights + 4 * a, k, np.float32)
        yield a, e, w


def generator_samples(sess, oracle, tg, scale, ef, ranges):
    """Edge ranges [a, b) of the device tuple list equal the oracle's."""
    for a, b in ranges:
        e = sess.to_host(tg.edges + 16 * a, 2 * (b - a), np.uint64).reshape(-1, 2)
        w = sess.to_host(tg.weights + 4 * a, b - a, np.float32)
        eo, wo = oracle.kronecker(scale, ef, start=a, end=b)
        assert np.array_equal(e, eo), (scale, ef, a, b)
        assert np.array_equal(w.view(np.uint32), wo.view(np.uint32)), (scale, ef, a, b)


def csr_against_tuples(sess, oracle, tg, g):
    """Checks g's CSR against the device tuple list it was built from and
    returns its host copy (an oracle.Csr)."""
    n = int(g.nlocalverts)
    deg = np.zeros(n, np.uint64)
    dig, cnt = 0, 0
    for _, e, w in stream_tuples(sess, tg):
        assert oracle.degree_add(n, e, deg) == 0
        d, c = oracle.tuple_digest(e, w)
        dig = (dig + d) & U64
        cnt += c
    say("tuples streamed:", int(tg.edges_size), "edges,", cnt, "entries")
    rs, col, cw = sess.csr_to_host(g)
    ref_rs = np.zeros(n + 1, np.uint64)
    np.cumsum(deg, out=ref_rs[1:])
    assert np.array_equal(rs, ref_rs)
    assert int(rs[-1]) == cnt == int(g.nlocaledges)
    assert int(g.notisolated) == int(np.count_nonzero(deg))
    h = oracle.Csr(n, rs, col, cw)
    assert oracle.csr_order_errors(h) == 0
    assert oracle.csr_digest(h) == dig
    return h


def bfs_plain(sess, oracle, g, h, roots, n_oracle, depths=None):
    """Single-root BFS from every root: the library's rules 1-7 validation
    passes; the first n_oracle roots' depths equal the oracle's. Depths of
    the roots in `depths` (a dict root -> None) are filled in."""
    b = sess.bfs_create(g)
    for i, r in enumerate(roots):
        r = int(r)
        sess.bfs(g, b, r)
        ok, err = sess.bfs_validate(g, b, r)
        assert ok and err == 0, (r, err)
        if i < n_oracle or (depths is not None and r in depths):
            d = sess.bfs_depth(g, b, r)
            if i < n_oracle:
                assert np.array_equal(d, oracle.bfs_depth_omp(h, r, nthreads())), r
            if depths is not None and r in depths:
                depths[r] = d
    return b


def bfs_batched(sess, g, roots, depths):
    """One batched launch over `roots`: depths of the roots in `depths` equal
    the single-root kernel's."""
    bb = sess.batch_create(g)
    ms = sess.batch(g, bb, roots)
    say(f"batched BFS, {len(roots)} roots: {ms:.1f} ms")
    for i, r in enumerate(roots):
        r = int(r)
        if r in depths:
            assert np.array_equal(sess.batch_depth(g, bb, i), depths[r]), (i, r)
    return bb


def sssp_plain(sess, oracle, g, h, roots, deltas):
    """Single-root SSSP from each root for each delta: the certificate holds
    and the distances are equal bit for bit across deltas. Returns root ->
    distances."""
    s = sess.sssp_create(g)
    out = {}
    for r in roots:
        r = int(r)
        for j, delta in enumerate(deltas):
            s.delta = delta
            ms = sess.sssp(g, s, r)
            d, p = sess.sssp_result(s)
            say(f"sssp root {r} delta {delta}: {ms:.1f} ms")
            assert oracle.sssp_check(h, r, d, p) == 0, (r, delta)
            if j == 0:
                out[r] = d
            else:
                assert np.array_equal(d.view(np.uint32), out[r].view(np.uint32)), (r, delta)
    return out


def sssp_batched(sess, oracle, g, h, roots, plain, certify, delta=0.05, offsets=None):
    """One batched launch over `roots` (with per-root bucket lags from
    `offsets`; "hub": the roots' hub distances): distances of roots in `plain`
    equal the single-root kernel's bit for bit; the roots in `certify` pass
    the certificate. Returns the launch's distances of the certified roots."""
    sb = sess.sssp_batch_create(g, delta)
    if isinstance(offsets, str) and offsets == "hub":
        offsets = sess.sssp_hub_distances(g, roots)
    ms = sess.sssp_batch(g, sb, roots, offsets)
    say(f"batched SSSP, {len(roots)} roots, delta {delta}, lags {'on' if offsets is not None else 'off'}: "
        f"{ms:.1f} ms")
    got = {}
    for i, r in enumerate(roots):
        r = int(r)
        if r not in plain and r not in certify:
            continue
        d, p = sess.sssp_batch_result(g, sb, i)
        if r in plain:
            assert np.array_equal(d.view(np.uint32), plain[r].view(np.uint32)), (i, r)
        if r in certify:
            assert oracle.sssp_check(h, r, d, p) == 0, (i, r)
            got[r] = d
    return sb, got
