"""Parity at the large workloads, through the C-ABI, one graph resident at a
time: Kronecker SCALE 24 and 26, edgefactor 16 and 32 (SCALE 26 edgefactor 32
has about 2^32 CSR entries, past the signed 32-bit range).

For each graph:
- generator: edge ranges at the start, middle and end of the tuple list (and
  around 2^31 where the list is that long) bit-exact against the oracle;
- CSR: checked against the whole device tuple list (degree prefix sum,
  canonical row order, entry multiset), so it is the oracle's CSR; the roots
  equal the oracle's root rule on it;
- plain BFS from every root passes the library's rules 1-7 validation, depths
  of the first roots equal the oracle's BFS; one 64-root batched BFS launch
  gives the plain kernel's depths;
- plain SSSP certified at delta 0.05 and 0.1 and equal across them; one
  32-root batched launch (and at SCALE 24, also a 16-root one) gives the plain
  kernel's distances bit for bit, and further roots of it are certified; the
  same launch with per-root bucket lags from the roots' hub distances, and
  with pulled far-class flushes, gives the same distances.
The certificate fixes the FP32 distances exactly (any label-correcting
schedule ends at the same fixed point), so no full Dijkstra is run here.
"""
import numpy as np
import pytest

from bigcheck import (bfs_batched, bfs_plain, csr_against_tuples, generator_samples, say, sssp_batched,
                      sssp_plain)

pytestmark = pytest.mark.gpu

# (scale, edgefactor, arena GiB, plain BFS roots validated, oracle-depth roots, batched-vs-plain BFS roots,
#  plain SSSP roots)
CONFIGS = {
    (24, 16): (64, 64, 2, 8, 4),
    (24, 32): (96, 64, 2, 8, 4),
    (26, 16): (200, 64, 2, 8, 4),
    (26, 32): (228, 16, 2, 4, 2),
}


class Graph:
    pass


@pytest.fixture(scope="module", params=list(CONFIGS), ids=lambda p: f"S{p[0]}ef{p[1]}")
def big(request, oracle):
    import g500
    scale, ef = request.param
    arena_gib = CONFIGS[request.param][0]
    say(f"S{scale} ef{ef}: session with a {arena_gib} GiB arena")
    s = g500.Session(arena_bytes=arena_gib << 30)
    x = Graph()
    x.sess, x.scale, x.ef = s, scale, ef
    x.nbfs, x.noracle, x.nbatch_cmp, x.nsssp = CONFIGS[request.param][1:]
    x.tg = s.generate(scale, ef)
    say(f"S{scale} ef{ef}: generated {x.tg.edges_size} edges in {s.lib.g500_last_run_ms():.1f} ms")
    x.g = s.csr(x.tg, 1 << scale)
    say(f"S{scale} ef{ef}: CSR with {x.g.nlocaledges} entries in {s.lib.g500_last_run_ms():.1f} ms")
    x.h = None
    yield x
    x.h = None
    s.close()


def test_generator_csr_roots(big, oracle):
    s, scale, ef = big.sess, big.scale, big.ef
    E = ef << scale
    k = 1 << 16
    ranges = [(0, k), (E // 2 - k // 2, E // 2 + k // 2), (E - k, E)]
    if E > (1 << 31):
        ranges.append(((1 << 31) - k // 2, (1 << 31) + k // 2))
    generator_samples(s, oracle, big.tg, scale, ef, ranges)
    big.h = csr_against_tuples(s, oracle, big.tg, big.g)
    if (scale, ef) == (26, 32):
        assert big.g.nlocaledges > (1 << 31)
    big.roots = s.roots(big.g)
    assert np.array_equal(big.roots, oracle.roots(big.h)) and len(big.roots) == 64
    say(f"S{scale} ef{ef}: generator samples, CSR and roots equal the oracle's")


def test_bfs_plain_and_batched(big, oracle):
    assert big.h is not None, "needs test_generator_csr_roots"
    s, g, roots = big.sess, big.g, big.roots
    m = s.mark()
    keep = {int(r): None for r in roots[:big.nbatch_cmp]}
    bfs_plain(s, oracle, g, big.h, roots[:big.nbfs], big.noracle, depths=keep)
    s.release(m)
    bfs_batched(s, g, roots, keep)
    s.release(m)


def test_sssp_plain_and_batched(big, oracle):
    assert big.h is not None, "needs test_generator_csr_roots"
    s, g, roots = big.sess, big.g, big.roots
    width = int(s.get_param("sssp_batch_max"))
    m = s.mark()
    plain = sssp_plain(s, oracle, g, big.h, roots[:big.nsssp], (0.05, 0.1))
    s.release(m)
    certify = {int(r) for r in roots[big.nsssp:big.nsssp + 2]}
    _, got = sssp_batched(s, oracle, g, big.h, roots[:width], plain, certify, delta=0.025)
    s.release(m)
    # the same launch with per-root bucket lags from the hub distances: the same distances
    same = dict(plain)
    same.update(got)
    sssp_batched(s, oracle, g, big.h, roots[:width], same, certify, delta=0.025, offsets="hub")
    s.release(m)
    # and with the far-class flushes of large windows pulled by their targets
    old = s.get_param("batch_pull")
    s.set_param("batch_pull", 0.05)
    try:
        sb, _ = sssp_batched(s, oracle, g, big.h, roots[:width], same, certify, delta=0.025, offsets="hub")
        say(f"S{big.scale} ef{big.ef}: pulled flushes {s.sssp_batch_pull_work(sb)}")
    finally:
        s.set_param("batch_pull", old)
    s.release(m)
    if big.scale == 24:  # the 16-root batches of the batched workload
        plain.update(got)
        sssp_batched(s, oracle, g, big.h, roots[:16], plain, set(), delta=0.05)
        s.release(m)
