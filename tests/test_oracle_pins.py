"""Pins the oracle's scramble values, vertex scramble, whole edges and the
64-root sequence with Python big-integer restatements written from the
reference text, by routes that share no arithmetic with oracle/oracle.c:

- val0/val1 (graph_generation.cu:290-300): the seed skipped by
  50 * 2^128 + 7 * 2^64 through a square-and-multiply power of the 5x5
  companion matrix, four big-integer MRG steps, products mod 2^64;
- the scramble (graph_generation.cu:171-179, bit reversal :145-169): bit
  reversal by reversing the 64-character binary string;
- make_one_edge (graph_generation.cu:114-143, :186-207) on top of both;
- the root loop (examples/graph500_runner.cu:43-74): the MRG stream from the
  seed skipped by 2 * 2^128, stepped one value at a time (no skip matrices),
  degrees from numpy.

Each check is shown to fail for the mistakes a restatement could plausibly
make (swapped or dropped scramble values, an altered constant, a missing bit
reversal, another double literal, an off-by-one in the root counter).
"""
import numpy as np
import pytest

from conftest import oracle_graph
from test_oracle_mrg import P, apply, mat_pow, py_steps

M64 = (1 << 64) - 1
K0 = 0x4519840211493211  # graph_generation.cu:174
K1 = 0x3050852102C843A5  # graph_generation.cu:176
INV_P = float(".0000000004656612875245796924106")  # graph_generation.cu:93


def py_brev(x):
    return int(format(x & M64, "064b")[::-1], 2)


def py_scramble_values(seed, skip=(50, 7)):
    hi, mid = skip
    z = apply(mat_pow(hi * 2**128 + mid * 2**64), [int(v) for v in seed])
    _, (a, b, c, d) = py_steps(z, 4)
    return (a * 0xFFFFFFFF + b) & M64, (c * 0xFFFFFFFF + d) & M64


def py_scramble(v, lgn, val0, val1, k0=K0, k1=K1, brev=py_brev, add=None):
    v = (v + (val0 + val1 if add is None else add)) & M64
    v = (v * (val0 | k0)) & M64
    v = brev(v) >> (64 - lgn)
    v = (v * (val1 | k1)) & M64
    return brev(v) >> (64 - lgn)


def py_edge(seed, lgn, ei, scr):
    """(src, tgt, weight bits) of edge ei: the seed skipped by ei * 2^64, one
    quadrant per level (reject u < 2^31-1 mod 10000, then B/C/A/D by u mod
    10000), the diagonal swap, both ends through scr, then the weight."""
    z = apply(mat_pow(ei * 2**64), [int(v) for v in seed])
    half, s, t = 1 << lgn, 0, 0
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
    z, (u,) = py_steps(z, 1)
    w = np.float32(np.float32(u) * np.float32(2.0**-31))
    return scr(s), scr(t), int(w.view(np.uint32))


def py_seed(s1, s2):
    # graph_generation.cu:248-254
    return [(s1 & 0x3FFFFFFF) + 1, ((s1 >> 30) & 0x3FFFFFFF) + 1, (s2 & 0x3FFFFFFF) + 1,
            ((s2 >> 30) & 0x3FFFFFFF) + 1, ((s2 >> 60) << 4) + (s1 >> 60) + 1]


def py_roots(n, degree, s1=2, s2=3, max_roots=64, literal=INV_P, counter_step=2):
    """examples/graph500_runner.cu:43-74 over the stepped stream: the two
    values of make_random_numbers(2, s1, s2, counter) are the stream's outputs
    2*counter + 1 and 2*counter + 2 after the skip by 2 * 2^128."""
    z = apply(mat_pow(2 * 2**128), py_seed(s1, s2))
    made = 0  # outputs drawn so far from z
    roots, counter = [], 0
    while len(roots) < max_roots:
        found = None
        while True:
            target = 2 * counter  # skip to output index 2*counter, then draw two
            if target > made:
                z, _ = py_steps(z, target - made)
                made = target
            z, (u0, u1) = py_steps(z, 2)
            made += 2
            d0, d1 = u0 * literal, u1 * literal
            root = int((d0 + d1) * float(n)) % n
            counter += counter_step
            if counter > 2 * n:
                break
            if root in roots or degree[root] == 0:
                continue
            found = root
            break
        if found is None:
            break
        roots.append(found)
    return roots


def degrees(n, edges):
    keep = edges[:, 0] != edges[:, 1]
    ends = np.concatenate([edges[keep, 0], edges[keep, 1]]).astype(np.int64)
    return np.bincount(ends, minlength=n)


# ------------------------------------------------------------ scramble values
@pytest.mark.parametrize("s1,s2", [(2, 3), (7, 11), (0, 0), ((1 << 63) + 12345, (5 << 60) + (77 << 30) + 9)])
def test_scramble_values(oracle, s1, s2):
    seed = py_seed(s1, s2)
    assert list(oracle.mrg_seed(s1, s2)) == seed
    assert oracle.scramble_values(seed) == py_scramble_values(seed)


def test_scramble_values_mistakes_are_caught(oracle):
    seed = py_seed(2, 3)
    got = oracle.scramble_values(seed)
    v0, v1 = py_scramble_values(seed)
    assert got != (v1, v0)                                  # swapped
    assert got != py_scramble_values(seed, skip=(7, 50))    # skip exponents swapped
    assert got != py_scramble_values(seed, skip=(50, 0))    # mid exponent dropped
    z = apply(mat_pow(50 * 2**128 + 7 * 2**64), seed)
    _, (a, b, c, d) = py_steps(z, 4)
    assert got != ((a << 32) + b, (c << 32) + d)            # 2^32 instead of 2^32 - 1


# ------------------------------------------------------------------ scramble
@pytest.mark.parametrize("lgn", [1, 2, 5, 12, 16, 26, 31, 32])
def test_scramble(oracle, lgn):
    v0, v1 = py_scramble_values(py_seed(2, 3))
    rng = np.random.default_rng(lgn)
    vs = sorted({0, 1, (1 << lgn) - 1, *(int(x) for x in rng.integers(0, 1 << lgn, 200, dtype=np.uint64))})
    for v in vs:
        assert oracle.lib().o_scramble(v, lgn, v0, v1) == py_scramble(v, lgn, v0, v1), v


def test_scramble_mistakes_are_caught(oracle):
    lgn = 12
    v0, v1 = py_scramble_values(py_seed(2, 3))
    got = [oracle.lib().o_scramble(v, lgn, v0, v1) for v in range(1 << lgn)]
    # only the low lgn bits of each multiplier reach the result: flip the
    # lowest such bit that val0 (val1) leaves to the constant
    b0 = next(i for i in range(lgn) if not (v0 >> i) & 1)
    b1 = next(i for i in range(lgn) if not (v1 >> i) & 1)
    variants = {
        "val0/val1 swapped": lambda v: py_scramble(v, lgn, v1, v0),
        "val1 dropped from the offset": lambda v: py_scramble(v, lgn, v0, v1, add=v0),
        "K0 one bit off": lambda v: py_scramble(v, lgn, v0, v1, k0=K0 ^ (1 << b0)),
        "K1 one bit off": lambda v: py_scramble(v, lgn, v0, v1, k1=K1 ^ (1 << b1)),
        "no bit reversal": lambda v: py_scramble(v, lgn, v0, v1, brev=lambda x: x),
        "32-bit halves not swapped": lambda v: py_scramble(
            v, lgn, v0, v1, brev=lambda x: ((py_brev(x) >> 32) | ((py_brev(x) & 0xFFFFFFFF) << 32))),
    }
    for name, f in variants.items():
        assert [f(v) for v in range(1 << lgn)] != got, name


# ------------------------------------------------------------- whole edges
@pytest.mark.parametrize("scale,eis", [(10, (0, 1, 255, 256, 16383)), (12, (7, 65535)),
                                       (26, (0, 2**20 + 3, 2**30 - 1)), (32, (2**33 + 17,))])
def test_edges_from_python_restatement(oracle, scale, eis):
    seed = py_seed(2, 3)
    v0, v1 = py_scramble_values(seed)
    scr = lambda x: py_scramble(x, scale, v0, v1)
    for ei in eis:
        e, w = oracle.kronecker(scale, 16, start=ei, end=ei + 1)
        s, t, wb = py_edge(seed, scale, ei, scr)
        assert (int(e[0, 0]), int(e[0, 1]), int(w.view(np.uint32)[0])) == (s, t, wb), ei


def test_edge_mistakes_are_caught(oracle):
    scale = 10
    seed = py_seed(2, 3)
    v0, v1 = py_scramble_values(seed)
    e, _ = oracle.kronecker(scale, 16, start=0, end=64)
    got = [(int(a), int(b)) for a, b in e]
    swapped = [py_edge(seed, scale, ei, lambda x: py_scramble(x, scale, v1, v0))[:2] for ei in range(64)]
    assert swapped != got
    right = [py_edge(seed, scale, ei, lambda x: py_scramble(x, scale, v0, v1))[:2] for ei in range(64)]
    assert right == got


# --------------------------------------------------------------------- roots
def test_double_literal_is_inverse_of_p():
    assert INV_P == float.fromhex("0x1.00000002p-31") and INV_P == 1.0 / P


@pytest.mark.parametrize("scale,ef,s1,s2", [(10, 16, 2, 3), (12, 16, 2, 3), (16, 16, 2, 3), (12, 4, 7, 11)])
def test_roots_full_sequence(oracle, scale, ef, s1, s2):
    n = 1 << scale
    e, w, g = oracle_graph(scale, ef)
    deg = degrees(n, e)
    assert np.array_equal(deg, np.diff(g.rowstarts.astype(np.int64)))
    ref = py_roots(n, deg, s1, s2)
    got = [int(r) for r in oracle.roots(g, s1, s2)]
    assert got == ref
    assert len(ref) == 64 and len(set(ref)) == 64 and all(deg[r] > 0 for r in ref)


def test_roots_exhaust_the_counter(oracle):
    """A graph with two non-isolated vertices: the loop stops when the
    counter passes 2n and returns fewer than 64 roots."""
    n = 1 << 6
    rs = np.zeros(n + 1, np.uint64)
    rs[6:] = 1
    rs[40:] = 2
    deg = np.diff(rs.astype(np.int64))
    ref = py_roots(n, deg)
    got = [int(r) for r in oracle.roots(oracle.Csr(n, rs, np.zeros(2, np.uint32), np.zeros(2, np.float32)))]
    assert got == ref and len(ref) <= 2 and set(ref) <= {5, 39}


def test_root_mistakes_are_caught(oracle):
    scale = 12
    n = 1 << scale
    e, w, g = oracle_graph(scale)
    deg = degrees(n, e)
    got = [int(r) for r in oracle.roots(g)]
    assert py_roots(n, deg, counter_step=1) != got            # counter += 1
    assert py_roots(n, deg, s1=3, s2=2) != got                # user seeds swapped
    assert py_roots(n, np.ones(n, np.int64)) != got           # isolated vertices not skipped


def test_roots_at_scale_26():
    """The root arithmetic at n = 2^26 (every vertex of degree 1, so no
    isolation skips): equal to the oracle, and sensitive to the double
    literal, which a small n does not resolve (2^-31 instead of 1/(2^31-1)
    moves (d0 + d1) * n by up to 0.06 here)."""
    import pyoracle
    n = 1 << 26
    rs = np.arange(n + 1, dtype=np.uint64)
    got = [int(r) for r in pyoracle.roots(pyoracle.Csr(n, rs, np.zeros(1, np.uint32), np.zeros(1, np.float32)))]
    ones = np.ones(n, np.int8)
    assert got == py_roots(n, ones)
    assert got != py_roots(n, ones, literal=2.0**-31)
