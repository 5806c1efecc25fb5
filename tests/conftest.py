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
