"""The autotuner's candidate lists are well formed (host logic, no GPU).

A batched SSSP candidate (delta, W_mid, W_long, W_far) must keep the class
floors ordered, delta <= W_mid <= W_long <= W_far: the kernel clamps an
out-of-order floor onto its neighbour, so a mistyped candidate would silently
time a different setting than the one the tuner reports.
"""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "cuda-graph-traversal_amd"))

from g500 import autotune  # noqa: E402


def test_batch_sssp_candidates_ordered_and_distinct():
    seen = set()
    for cand in autotune.BATCH_SSSP:
        d, wm, wl, wf = cand
        assert 0.0 < d <= wm <= wl <= wf < 1.0, cand
        assert cand not in seen, cand
        seen.add(cand)


def test_scalar_candidate_lists():
    assert all(d > 0.0 for d in autotune.SSSP_DELTA)
    assert all(0.0 <= p < 1.0 for p in autotune.SSSP_PULL)
    assert all(0.0 <= p < 1.0 for p in autotune.BATCH_PULL)
    assert all(a > 0.0 and b > 0.0 for a, b in autotune.BFS_SWITCH)
    assert all(a > 0.0 for a in autotune.BATCH_ALPHA)
    assert all(b in (128, 256, 512) for b in autotune.BLOCKS)
    for lst in (autotune.SSSP_DELTA, autotune.SSSP_PULL, autotune.BATCH_PULL, autotune.BFS_SWITCH,
                autotune.BATCH_ALPHA, autotune.BATCH_HEAVY, autotune.BLOCKS):
        assert len(set(lst)) == len(lst), lst
