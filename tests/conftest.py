import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "cuda-graph-traversal_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle


_GRAPHS = {}


def oracle_graph(scale, edgefactor=16, s1=2, s2=3):
    """(edges, weights, csr) from the CPU oracle, cached per session."""
    key = (scale, edgefactor, s1, s2)
    if key not in _GRAPHS:
        import pyoracle
        e, w = pyoracle.kronecker(scale, edgefactor, s1, s2)
        g = pyoracle.csr(1 << scale, e, w)
        _GRAPHS[key] = (e, w, g)
    return _GRAPHS[key]


@pytest.fixture(scope="session")
def g10():
    return oracle_graph(10)


@pytest.fixture(scope="session")
def g12():
    return oracle_graph(12)


def rng(seed=0):
    return np.random.default_rng(seed)


def prune_rule(rowstarts, column, weights, hubs, margin=2.0 ** -10, wmax=2.0):
    """Entries (u, v, w) that a two-hop path u-x-v through a vertex x of
    `hubs` beats by more than `margin`: hw[u][x] + hw[v][x] + margin < w in
    f64, w < wmax, with hw[y][x] the lightest weight of an entry (y, x) (+inf
    without one). Returns the boolean drop mask over the entries."""
    rs = np.asarray(rowstarts, np.int64)
    col = np.asarray(column, np.int64)
    w = np.asarray(weights, np.float32).astype(np.float64)
    n = len(rs) - 1
    hubs = np.asarray(hubs, np.int64)
    slot = np.full(n, -1, np.int64)
    slot[hubs] = np.arange(len(hubs))
    src = np.repeat(np.arange(n), np.diff(rs))
    hw = np.full((n, len(hubs)), np.inf)
    to_hub = slot[col] >= 0
    np.minimum.at(hw, (src[to_hub], slot[col[to_hub]]), w[to_hub])
    best = np.full(len(col), np.inf)
    for x in range(len(hubs)):
        best = np.minimum(best, hw[src, x] + hw[col, x])
    return (w < wmax) & (best + margin < w)


def drop_entries(rowstarts, column, weights, drop):
    """(rowstarts, column, weights) of the CSR without the entries in `drop`,
    rows and entry order otherwise unchanged."""
    rs = np.asarray(rowstarts, np.int64)
    n = len(rs) - 1
    src = np.repeat(np.arange(n), np.diff(rs))
    keep = ~drop
    rs_k = np.zeros(n + 1, np.uint64)
    rs_k[1:] = np.cumsum(np.bincount(src[keep], minlength=n))
    return rs_k, np.asarray(column)[keep], np.asarray(weights)[keep]
