"""Autotuner for the traversal knobs of one resident graph.

It times each candidate setting through the C-ABI on tuning roots and keeps
the fastest one. The roots are non-isolated vertices of the giant component,
drawn from a fixed numpy seed, so they are distinct from the Graph500 roots
that a benchmark measures. Candidates:
  plain BFS:     (bfs_alpha, bfs_beta) switch thresholds, bfs_block
  batched BFS:   batch_alpha, batch_heavy_degree, then batch_bfs_block
  plain SSSP:    delta, sssp_block, sssp_pull (the bucket share above which
                 heavy relaxations are pulled)
  batched SSSP:  (delta, mid_weight, long_weight, far_weight) candidates, then
                 batch_sssp_block and batch_pull (the window share above which
                 far-class flushes are pulled) for the best of them (with
                 per-root bucket lags from the tuning roots' hub distances
                 unless lags=False)
Every candidate gives the same depths and distances (the tuner changes only
schedules), so the choice is by time alone. Plain BFS candidates are timed on
16 roots and plain SSSP ones on 4 (the mean of their times; a few roots whose
level 2 frontier lies just below the switch threshold take several times the
median, so few roots misjudge the thresholds), batched ones on one launch each.
"""
import numpy as np

BFS_SWITCH = [(15.0, 60.0), (30.0, 60.0), (60.0, 60.0), (120.0, 60.0), (60.0, 18.0)]
BLOCKS = [256, 512]
BATCH_ALPHA = [4.0, 15.0, 60.0]
BATCH_HEAVY = [4096, 1 << 20]
SSSP_DELTA = [0.025, 0.05, 0.1]
SSSP_PULL = [0.02, 0.05, 0.2]
# (delta, mid, long, far); far == long leaves the far class empty
BATCH_SSSP = [(0.025, 0.075, 0.25, 0.5), (0.025, 0.05, 0.15, 0.3), (0.025, 0.075, 0.25, 0.25),
              (0.0125, 0.05, 0.2, 0.5), (0.05, 0.1, 0.3, 0.6), (0.05, 0.075, 0.2, 0.4),
              (0.025, 0.1, 0.35, 0.35), (0.0375, 0.075, 0.25, 0.25), (0.025, 0.05, 0.2, 0.2),
              (0.025, 0.05, 0.15, 0.15), (0.025, 0.05, 0.1, 0.1), (0.02, 0.05, 0.15, 0.15),
              (0.02, 0.04, 0.15, 0.15), (0.02, 0.05, 0.2, 0.2), (0.0175, 0.05, 0.15, 0.15)]
BATCH_PULL = [0.0, 0.1, 0.2, 0.3]


def tuning_roots(sess, g, count, seed=2024):
    """`count` giant-component roots from a fixed seed."""
    n = int(g.nlocalverts)
    deg = np.diff(sess.to_host(g.rowstarts, n + 1, np.uint64).astype(np.int64))
    cand = np.flatnonzero(deg > 0)
    rng = np.random.default_rng(seed)
    b = sess.bfs_create(g)
    sess.bfs(g, b, int(np.argmax(deg)))
    n_giant = sum(sess.level_log(b)[1])
    out = []
    for v in rng.permutation(cand)[: 8 * count]:
        sess.bfs(g, b, int(v))
        if sum(sess.level_log(b)[1]) == n_giant:
            out.append(int(v))
            if len(out) == count:
                break
    return np.array(out, np.uint64)


def _best(results):
    return min(results, key=lambda r: r["ms"])


def tune(sess, g, batch=True, roots_plain=4, roots_bfs=16, seed=2024, log=None, lags=True):
    """Tunes the knobs on graph g, applies the best settings, and returns
    them with the measurements: {"params": ..., "trials": [...]}."""
    say = log or (lambda *a: None)
    mark = sess.mark()  # the tuner's traversal states are released at the end
    try:
        return _tune(sess, g, batch, roots_plain, roots_bfs, seed, say, lags)
    finally:
        sess.release(mark)


def _tune(sess, g, batch, roots_plain, roots_bfs, seed, say, lags):
    roots = tuning_roots(sess, g, max(64, roots_plain, roots_bfs), seed)
    plain = roots[:roots_plain]
    trials = []
    old = {k: sess.get_param(k) for k in ("bfs_alpha", "bfs_beta", "bfs_block", "sssp_block", "batch_alpha",
                                          "batch_heavy_degree", "batch_bfs_block", "batch_sssp_block", "batch_pull")}

    b = sess.bfs_create(g)
    res = []
    for (al, be) in BFS_SWITCH:
        for blk in BLOCKS:
            sess.set_param("bfs_alpha", al)
            sess.set_param("bfs_beta", be)
            sess.set_param("bfs_block", blk)
            ms = float(np.mean([sess.bfs(g, b, r) for r in roots[:roots_bfs]]))
            res.append({"kind": "bfs", "bfs_alpha": al, "bfs_beta": be, "bfs_block": blk, "ms": ms})
    best_bfs = _best(res)
    trials += res
    for k in ("bfs_alpha", "bfs_beta", "bfs_block"):
        sess.set_param(k, best_bfs[k])
    say(f"[tune] bfs {best_bfs}")

    s = sess.sssp_create(g)
    res = []
    for d in SSSP_DELTA:
        for blk in BLOCKS:
            for pull in SSSP_PULL:
                sess.set_param("sssp_block", blk)
                sess.set_param("sssp_pull", pull)
                s.delta = d
                ms = float(np.mean([sess.sssp(g, s, r) for r in plain]))
                res.append({"kind": "sssp", "delta": d, "sssp_block": blk, "sssp_pull": pull, "ms": ms})
    best_sssp = _best(res)
    trials += res
    sess.set_param("sssp_block", best_sssp["sssp_block"])
    sess.set_param("sssp_pull", best_sssp["sssp_pull"])
    say(f"[tune] sssp {best_sssp}")

    params = {"bfs_alpha": best_bfs["bfs_alpha"], "bfs_beta": best_bfs["bfs_beta"],
              "bfs_block": best_bfs["bfs_block"], "sssp_block": best_sssp["sssp_block"],
              "sssp_delta": best_sssp["delta"], "sssp_pull": best_sssp["sssp_pull"]}
    if batch:
        bb = sess.batch_create(g)
        res = []
        for al in BATCH_ALPHA:
            for hv in BATCH_HEAVY:
                sess.set_param("batch_alpha", al)
                sess.set_param("batch_heavy_degree", hv)
                ms = sess.batch(g, bb, roots[:64])
                res.append({"kind": "batch_bfs", "batch_alpha": al, "batch_heavy_degree": hv, "ms": ms})
        best_bb = _best(res)
        sess.set_param("batch_alpha", best_bb["batch_alpha"])
        sess.set_param("batch_heavy_degree", best_bb["batch_heavy_degree"])
        for blk in BLOCKS:  # block shape for the best thresholds
            sess.set_param("batch_bfs_block", blk)
            res.append(dict(best_bb, batch_bfs_block=blk, ms=sess.batch(g, bb, roots[:64])))
        best_bb = _best(res)
        best_bb.setdefault("batch_bfs_block", int(old["batch_bfs_block"]))
        sess.set_param("batch_bfs_block", best_bb["batch_bfs_block"])
        trials += res
        say(f"[tune] batched bfs {best_bb}")
        sb = sess.sssp_batch_create(g)
        width = int(sess.get_param("sssp_batch_max"))
        offs = sess.sssp_hub_distances(g, roots[:width]) if lags else None
        res = []
        for d, wm, wl, wf in BATCH_SSSP:
            sb.delta, sb.mid_weight, sb.long_weight, sb.far_weight = d, wm, wl, wf
            ms = sess.sssp_batch(g, sb, roots[:width], offs)
            res.append({"kind": "batch_sssp", "delta": d, "mid_weight": wm, "long_weight": wl, "far_weight": wf,
                        "ms": ms})
        best_bs = _best(res)
        sb.delta, sb.mid_weight = best_bs["delta"], best_bs["mid_weight"]
        sb.long_weight, sb.far_weight = best_bs["long_weight"], best_bs["far_weight"]
        for blk in BLOCKS:  # block shape for the best schedule
            sess.set_param("batch_sssp_block", blk)
            res.append(dict(best_bs, batch_sssp_block=blk, ms=sess.sssp_batch(g, sb, roots[:width], offs)))
        best_bs = _best(res)
        best_bs.setdefault("batch_sssp_block", int(old["batch_sssp_block"]))
        sess.set_param("batch_sssp_block", best_bs["batch_sssp_block"])
        if lags:  # pulled flushes serve only roots with an offset
            for share in BATCH_PULL:
                sess.set_param("batch_pull", share)
                res.append(dict(best_bs, batch_pull=share, ms=sess.sssp_batch(g, sb, roots[:width], offs)))
            best_bs = _best(res)
        best_bs.setdefault("batch_pull", float(old["batch_pull"]))
        sess.set_param("batch_pull", best_bs["batch_pull"])
        trials += res
        say(f"[tune] batched sssp {best_bs}")
        params.update({"batch_alpha": best_bb["batch_alpha"], "batch_heavy_degree": best_bb["batch_heavy_degree"],
                       "batch_bfs_block": best_bb["batch_bfs_block"], "batch_sssp_block": best_bs["batch_sssp_block"],
                       "batch_sssp_delta": best_bs["delta"], "batch_sssp_mid_weight": best_bs["mid_weight"],
                       "batch_sssp_long_weight": best_bs["long_weight"], "batch_sssp_far_weight": best_bs["far_weight"],
                       "batch_pull": best_bs["batch_pull"]})
    else:
        for k in ("batch_alpha", "batch_heavy_degree", "batch_bfs_block", "batch_sssp_block", "batch_pull"):
            sess.set_param(k, old[k])
    return {"params": params, "trials": trials, "tuning_roots": len(roots)}
