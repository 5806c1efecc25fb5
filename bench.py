#!/usr/bin/env python3
"""Graph500 BFS + SSSP benchmark on one MI355X.

Workload (default): Kronecker graph SCALE 26, edgefactor 16 (seed 2, 3),
generated and CSR-built on the GPU, 64 roots from the Graph500 root rule. All
inputs are resident in HBM before any timed region. Times are the kernels'
HIP-event durations on the library stream.

Batched mode (default): one step = the 64 roots, traversed by one batched BFS
launch (64-bit per-vertex root masks) and batched delta-stepping SSSP launches
of the library's width (32 roots each, shared buckets).
Plain mode: one step = one root, single-root BFS then single-root SSSP.

`value` is whole-job throughput in GTEPS: the TEPS edges of every traversal
(m_c = undirected edges of the root's component, as the Graph500 runner counts
them; once for the BFS and once for the SSSP of each root) over the summed
kernel time. A plain reference pass over the same roots is reported beside the
batched value (`plain`), with the runner's harmonic means; `batched_vs_plain`
is the ratio of the two whole-job throughputs.

Prints one JSON line. --gpus above 1 is rejected (one GPU).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cuda-graph-traversal_amd"))

METRIC = "harmonic-mean GTEPS BFS+SSSP @SCALE26 ef16, batched vs plain; p99 req latency"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
BFS_GROUP = 64     # roots per batched BFS launch (the SSSP launch width is the library's sssp_batch_max)


def hm(times_s, edges):
    return len(times_s) / float(np.sum(np.asarray(times_s) / np.asarray(edges, dtype=np.float64)))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sess, g, roots, m_c, budget_s, nthreads, full=False):
    """The CPU oracle on the host cores, over complete traversals of the same
    graph and giant-component roots (in root order): OpenMP level-synchronous
    BFS and OpenMP delta-stepping (delta 0.1, the reference's value), each from
    as many roots as fit half the time budget (at least 2 and 1). With `full`,
    the planned baseline: OpenMP BFS from all giant roots, delta-stepping from
    8, and serial queue BFS and binary-heap Dijkstra from 2 roots each.
    Timing is the traversal call alone, as the runner times its kernels."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    rs, col, w = sess.csr_to_host(g)
    ref = O.Csr(int(g.nlocalverts), rs, col, w)
    giant = [i for i in range(len(roots)) if m_c[i] >= 0.5 * m_c.max()]

    def run(fn, limit, at_least, at_most):
        done, spent = [], 0.0
        for i in giant[:at_most]:
            if len(done) >= at_least and spent >= limit:
                break
            t0 = time.perf_counter()
            fn(int(roots[i]))
            dt = time.perf_counter() - t0
            done.append((i, dt))
            spent += dt
        return done

    half = budget_s / 2
    bfs = run(lambda r: O.bfs_omp(ref, r, nthreads), half, 2, len(giant) if full else 64)
    ds = run(lambda r: O.delta_stepping_omp(ref, r, 0.1, nthreads), half, 8 if full else 1, 8 if full else 64)
    gteps = lambda runs: sum(m_c[i] for i, _ in runs) / sum(t for _, t in runs) / 1e9
    b, d = gteps(bfs), gteps(ds)
    out = {
        "value": 2.0 / (1.0 / b + 1.0 / d), "unit": "GTEPS", "cores": nthreads, "kind": "port",
        "cpu": cpu_model(), "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
        "bfs_omp": {"roots": [int(roots[i]) for i, _ in bfs], "s": [round(t, 3) for _, t in bfs],
                    "gteps": round(b, 4)},
        "sssp_omp": {"roots": [int(roots[i]) for i, _ in ds], "s": [round(t, 3) for _, t in ds],
                     "gteps": round(d, 4), "delta": 0.1},
    }
    if full:
        ser = run(lambda r: O.bfs(ref, r, False), 0.0, 2, 2)
        dij = run(lambda r: O.dijkstra_f32(ref, r), 0.0, 2, 2)
        out["bfs_serial"] = {"roots": [int(roots[i]) for i, _ in ser], "s": [round(t, 3) for _, t in ser],
                             "gteps": round(gteps(ser), 4)}
        out["dijkstra_serial"] = {"roots": [int(roots[i]) for i, _ in dij], "s": [round(t, 3) for _, t in dij],
                                  "gteps": round(gteps(dij), 4)}
    scale = int(np.log2(g.nlocalverts))
    out["sample"] = (f"oracle on the same SCALE {scale} CSR ({out['cpu']}, nproc {out['nproc']}, {nthreads} threads "
                     f"used): complete OpenMP level-synchronous BFS from {len(bfs)} giant-component roots "
                     f"({sum(t for _, t in bfs):.1f} s, {b:.3f} GTEPS) and complete OpenMP delta-stepping "
                     f"(delta 0.1) from {len(ds)} ({sum(t for _, t in ds):.1f} s, {d:.4f} GTEPS); "
                     f"value = 2/(1/BFS + 1/SSSP), GTEPS = sum of m_c / sum of times")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--mode", choices=("batched", "plain"), default="batched")
    ap.add_argument("--steps", type=int, default=1,
                    help="timed steps (batched: all 64 roots; plain: one root)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-roots", type=int, default=64,
                    help="roots of the plain reference pass in batched mode (0: skip)")
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--delta", type=float, default=0.1, help="SSSP delta when not tuned")
    ap.add_argument("--plain-delta", type=float, default=None,
                    help="single-root SSSP delta when not tuned (default: --delta)")
    ap.add_argument("--sssp-pull", type=float, default=None,
                    help="single-root SSSP pull ratio when not tuned (default: the library's)")
    ap.add_argument("--long-weight", type=float, default=0.15,
                    help="batched SSSP long-edge threshold (deferred over windows this wide) when not tuned")
    ap.add_argument("--mid-weight", type=float, default=0.05,
                    help="batched SSSP mid-edge threshold (deferred over windows this wide) when not tuned")
    ap.add_argument("--far-weight", type=float, default=0.3,
                    help="batched SSSP far-edge threshold (deferred over windows this wide) when not tuned")
    ap.add_argument("--no-root-order", dest="root_order", action="store_false",
                    help="SSSP launches take the roots in root-rule order instead of by hub distance")
    ap.add_argument("--no-root-lag", dest="root_lag", action="store_false",
                    help="batched SSSP roots share one bucket sequence (no per-root lags from hub distances)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (knobs at their defaults, delta from --delta)")
    ap.add_argument("--arena-gib", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=30.0,
                    help="time budget of the CPU baseline's complete traversals")
    ap.add_argument("--cpu-full", action="store_true",
                    help="the full CPU baseline: OpenMP BFS from every giant root, delta-stepping from 8, "
                         "serial BFS and Dijkstra from 2 (minutes at SCALE 26)")
    ap.add_argument("--serve-requests", type=int, default=2048,
                    help="requests of the serving leg (0: skip); resident graphs S16-S24")
    ap.add_argument("--serve-rate", type=float, default=1000.0, help="Poisson arrivals per second")
    ap.add_argument("--serve-sweep", default="300,5000,20000,50000",
                    help="further offered rates (comma-separated, empty for none) reported in serving_sweep")
    ap.add_argument("--calibrate-mib", type=int, default=0,
                    help="also run the device copy kernel on this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import g500
    sess = g500.Session(arena_bytes=args.arena_gib << 30)
    lib = sess.lib
    global SSSP_GROUP
    SSSP_GROUP = int(sess.get_param("sssp_batch_max"))
    log = lambda *a: print(*a, file=sys.stderr, flush=True)

    log(f"[bench] generating SCALE {args.scale} edgefactor {args.edgefactor}")
    tg = sess.generate(args.scale, args.edgefactor)
    gen_ms = lib.g500_last_run_ms()
    log(f"[bench] generator {gen_ms:.1f} ms; building CSR")
    g = sess.csr(tg, 1 << args.scale)
    csr_ms = lib.g500_last_run_ms()
    log(f"[bench] CSR {csr_ms:.1f} ms, {g.nlocaledges} entries")
    if args.calibrate_mib:
        nb = args.calibrate_mib << 20
        m = sess.mark()
        src = lib._arena_push(g500.C.byref(sess.ws.gpu_arena), nb, 256)
        dst = lib._arena_push(g500.C.byref(sess.ws.gpu_arena), nb, 256)
        lib.g500_copy_device(dst, src, nb)
        log(f"[bench] calibration copy of {args.calibrate_mib} MiB: {lib.g500_last_run_ms():.3f} ms")
        sess.release(m)
    tuned = None
    delta_plain = delta_batch = args.delta
    if args.plain_delta is not None:
        delta_plain = args.plain_delta
    if args.sssp_pull is not None:
        sess.set_param("sssp_pull", args.sssp_pull)
    long_weight, mid_weight, far_weight = args.long_weight, args.mid_weight, args.far_weight
    if not args.no_tune:
        from g500 import autotune
        t0 = time.perf_counter()
        tuned = autotune.tune(sess, g, batch=args.mode == "batched", log=log, lags=args.root_lag)["params"]
        delta_plain = tuned["sssp_delta"]
        delta_batch = tuned.get("batch_sssp_delta", delta_plain)
        long_weight = tuned.get("batch_sssp_long_weight", long_weight)
        mid_weight = tuned.get("batch_sssp_mid_weight", mid_weight)
        far_weight = tuned.get("batch_sssp_far_weight", far_weight)
        log(f"[bench] autotuner ({time.perf_counter() - t0:.1f} s, untimed): {tuned}")
    roots = sess.roots(g)
    nroots = len(roots)
    states = sess.mark()  # traversal states below are released before the serving leg
    bfs = sess.bfs_create(g)
    sssp = sess.sssp_create(g, delta_plain)

    # per-root TEPS numerators and reached counts (untimed, as in the runner)
    m_c = np.zeros(nroots, np.float64)
    n_c = np.zeros(nroots, np.float64)
    for i in range(nroots):
        sess.bfs(g, bfs, roots[i])
        m_c[i] = sess.edge_count(g, bfs.pred)
        n_c[i] = sum(sess.level_log(bfs)[1])
    log(f"[bench] {nroots} roots; edges per root ~{np.median(m_c):.4g}, reached ~{np.median(n_c):.4g}; "
        f"smallest components (edges): {sorted(m_c)[:4]}")

    checks = {"plain_bfs": 0, "plain_bfs_failures": 0, "plain_sssp": 0, "plain_sssp_failures": 0,
              "batched_bfs": 0, "batched_bfs_failures": 0, "batched_sssp": 0, "batched_sssp_failures": 0}

    def plain_pass(idx, label, validate=False):
        """Times single-root BFS and SSSP from roots[idx]; with `validate`, each
        result is checked after its timed run (rules 1-7 for the BFS tree, the
        certificate for the SSSP distances and parents)."""
        tb = np.zeros(len(idx))
        ts = np.zeros(len(idx))
        for i, ri in enumerate(idx):
            r = int(roots[ri])
            tb[i] = sess.bfs(g, bfs, r) * 1e-3
            if validate:
                ok, _ = sess.bfs_validate(g, bfs, r)
                checks["plain_bfs"] += 1
                checks["plain_bfs_failures"] += int(not ok)
            ts[i] = sess.sssp(g, sssp, r) * 1e-3
            if validate:
                checks["plain_sssp"] += 1
                checks["plain_sssp_failures"] += int(sess.sssp_check(g, sssp.dist, sssp.pred, r) != 0)
        log(f"[bench] {label}: bfs ms median {np.median(tb)*1e3:.3f} max {tb.max()*1e3:.3f}; "
            f"sssp ms median {np.median(ts)*1e3:.3f} max {ts.max()*1e3:.3f}")
        return tb, ts

    kern = {}
    order_ms = None
    if args.mode == "plain":
        for i in range(args.warmup):
            r = roots[(args.steps + i) % nroots]
            sess.bfs(g, bfs, r)
            sess.sssp(g, sssp, r)
        idx = [i % nroots for i in range(args.steps)]
        t0 = time.perf_counter()
        tb, ts = plain_pass(idx, "plain")
        t1 = time.perf_counter()
        plain_pass(idx[:64], "plain, validated (untimed)", validate=True)
        me, nv = m_c[idx], n_c[idx]
        edges_total = 2.0 * me.sum()
        kern_time = tb.sum() + ts.sum()
        kern["bfs_kernel"] = {"bytes": float((16.0 * me + 16.0 * nv).mean()), "time": float(tb.mean()),
                              "total": float(tb.sum()), "units": "one root"}
        kern["sssp_kernel"] = {"bytes": float((24.0 * me + 16.0 * nv).mean()), "time": float(ts.mean()),
                               "total": float(ts.sum()), "units": "one root"}
        plain = None
        sssp16 = None
        workload = (f"plain single-root BFS + delta-stepping SSSP per step (one root), Kronecker SCALE "
                    f"{args.scale} edgefactor {args.edgefactor}, roots from the Graph500 root rule")
    else:
        bb = sess.batch_create(g)
        sb = sess.sssp_batch_create(g, delta_batch)
        sb.long_weight, sb.mid_weight, sb.far_weight = long_weight, mid_weight, far_weight
        rsets = [list(range(k, min(k + BFS_GROUP, nroots))) for k in range(0, nroots, BFS_GROUP)]
        # Root grouping of the SSSP launches (once per graph, untimed, like the CSR build): the roots
        # of a step are sorted by their distance from the highest-degree vertex. Shortest paths of a
        # Kronecker graph run through its hubs, so roots at similar hub distances reach each vertex at
        # similar distances, and the roots of one launch settle it in the same buckets and windows.
        # Per-root bucket lags (once per graph, untimed, like the CSR build): the roots' distances
        # from the highest-degree vertex. A root's distance to most vertices of a Kronecker graph is
        # about its hub distance plus the vertex's, so a root that lags the others by its hub-distance
        # difference settles each vertex in the same shared buckets as they do.
        hubd = None
        if args.root_order or args.root_lag:
            t_o = time.perf_counter()
            hubd = sess.sssp_hub_distances(g, roots)
            order_ms = (time.perf_counter() - t_o) * 1e3
            log(f"[bench] hub distances of the roots (oned_graph_sssp_hub_distances, {order_ms:.1f} ms, untimed)")
        if args.root_order:  # launch groups of roots at similar hub distances
            rsets = [[grp[i] for i in np.argsort(hubd[grp], kind="stable")] for grp in rsets]

        def offsets_of(idx):
            return hubd[idx] if args.root_lag else None

        def batched_step(grp):
            rr = roots[grp]
            b = 0.0
            for k in range(0, len(rr), BFS_GROUP):
                b += sess.batch(g, bb, rr[k:k + BFS_GROUP]) * 1e-3
            s = 0.0
            for k in range(0, len(rr), SSSP_GROUP):
                s += sess.sssp_batch(g, sb, rr[k:k + SSSP_GROUP], offsets_of(grp[k:k + SSSP_GROUP])) * 1e-3
            return b, s

        for i in range(args.warmup):
            batched_step(rsets[(args.steps + i) % len(rsets)])
        tb = np.zeros(args.steps)
        ts = np.zeros(args.steps)
        me = np.zeros(args.steps)
        nv = np.zeros(args.steps)
        t0 = time.perf_counter()
        for i in range(args.steps):
            grp = rsets[i % len(rsets)]
            tb[i], ts[i] = batched_step(grp)
            me[i] = m_c[grp].sum()
            nv[i] = n_c[grp].sum()
        t1 = time.perf_counter()
        # untimed: every result of one more step is checked on the device (the
        # BFS depths by the depth certificate, the SSSP distances and parents by
        # the SSSP certificate), launch by launch
        n = int(g.nlocalverts)
        mark_v = sess.mark()
        depth_dev = lib._worker_arena_push_array(g500.C.byref(sess.ws), n, 8, 256, g500.C.byref(g500.u64()))
        rr = roots[rsets[0]]
        for k in range(0, len(rr), BFS_GROUP):
            part = rr[k:k + BFS_GROUP]
            sess.batch(g, bb, part)
            for i, r in enumerate(part):
                lib.oned_graph_bfs_batch_depth(g500.C.byref(sess.ws), g500.C.byref(g), g500.C.byref(bb), i, depth_dev)
                checks["batched_bfs"] += 1
                checks["batched_bfs_failures"] += int(sess.depth_check(g, depth_dev, int(r)) != 0)
        sess.release(mark_v)
        for k in range(0, len(rr), SSSP_GROUP):
            part = rr[k:k + SSSP_GROUP]
            sess.sssp_batch(g, sb, part, offsets_of(rsets[0][k:k + SSSP_GROUP]))
            for i, r in enumerate(part):
                checks["batched_sssp"] += 1
                checks["batched_sssp_failures"] += int(
                    sess.sssp_check(g, sb.dist + 4 * n * i, sb.pred + 8 * n * i, int(r)) != 0)
        log(f"[bench] validation: {checks}")
        # the 16-root SSSP batches of the batched workload (reported, not in value)
        t16, m16 = [], 0.0
        for k in range(0, len(rr), 16):
            t16.append(sess.sssp_batch(g, sb, rr[k:k + 16], offsets_of(rsets[0][k:k + 16])) * 1e-3)
            m16 += float(m_c[rsets[0][k:k + 16]].sum())
        sssp16 = {"launches": len(t16), "roots_per_launch": 16, "ms_per_launch": round(float(np.mean(t16)) * 1e3, 3),
                  "gteps": round(m16 / sum(t16) / 1e9, 4)}
        log(f"[bench] 16-root SSSP launches: {sssp16}")
        nb = -(-len(rsets[0]) // BFS_GROUP)
        ns = -(-len(rsets[0]) // SSSP_GROUP)
        log(f"[bench] batched: bfs ms per step {np.median(tb)*1e3:.2f} ({nb} launch); "
            f"sssp ms per step {np.median(ts)*1e3:.1f} ({ns} launches)")
        edges_total = 2.0 * me.sum()
        kern_time = tb.sum() + ts.sum()
        # Shared-pass bytes of one launch, the least a batched launch moves when its roots share
        # every row pass: the union component's rows once (BFS: 8 B of columns per undirected edge,
        # SSSP: 16 B of columns and weights) plus the roots' own state, which does not shrink with
        # sharing (BFS: a 64-bit root mask probed per entry, 16 B per edge, and read, written and
        # queued per vertex, 24 B; SSSP: per root, the 4-byte target probe of both directed entries
        # of an edge, 8 B, and the 16-byte label per reached vertex).
        mu, nu = float(m_c.max()), float(n_c.max())
        kern["batch_bfs_kernel"] = {"bytes": float((16.0 * me + 16.0 * nv).mean()) / nb,
                                    "shared_pass": 24.0 * mu + 24.0 * nu,
                                    "time": float(tb.mean()) / nb, "total": float(tb.sum()),
                                    "units": f"{BFS_GROUP} roots"}
        kern["sssp_batch_kernel"] = {"bytes": float((24.0 * me + 16.0 * nv).mean()) / ns,
                                     "shared_pass": 16.0 * mu + float((8.0 * me + 16.0 * nv).mean()) / ns,
                                     "time": float(ts.mean()) / ns, "total": float(ts.sum()),
                                     "units": f"{SSSP_GROUP} roots"}
        plain = None
        if args.plain_roots:
            idx = list(range(min(args.plain_roots, nroots)))
            ptb, pts = plain_pass(idx, "plain reference", validate=True)
            pm = m_c[idx]
            giant = pm >= 0.5 * pm.max()
            pw = 2.0 * pm.sum() / (ptb.sum() + pts.sum()) / 1e9
            plain = {"roots": len(idx), "gteps": round(pw, 4),
                     "bfs_gteps": round(pm.sum() / ptb.sum() / 1e9, 4),
                     "sssp_gteps": round(pm.sum() / pts.sum() / 1e9, 4),
                     "bfs_hm_gteps": round(hm(ptb, pm) / 1e9, 6), "sssp_hm_gteps": round(hm(pts, pm) / 1e9, 6),
                     "giant_roots": int(giant.sum()),
                     "bfs_hm_gteps_giant": round(hm(ptb[giant], pm[giant]) / 1e9, 4),
                     "sssp_hm_gteps_giant": round(hm(pts[giant], pm[giant]) / 1e9, 4),
                     "bfs_ms_median": round(float(np.median(ptb)) * 1e3, 3),
                     "sssp_ms_median": round(float(np.median(pts)) * 1e3, 3)}
        workload = (f"batched: per step the {BFS_GROUP} Graph500 roots, one 64-bit-mask BFS launch and "
                    f"{BFS_GROUP // SSSP_GROUP} {SSSP_GROUP}-root delta-stepping SSSP launches, Kronecker SCALE "
                    f"{args.scale} edgefactor {args.edgefactor}")
    value = edges_total / kern_time / 1e9
    dom = max(kern, key=lambda k: kern[k]["total"])
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", f"traffic_s{args.scale}_ef{args.edgefactor}.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath))

    def roof(name):
        k = kern[name]
        ach = k["bytes"] / k["time"] / 1e9
        t = traffic.get(name)
        out = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
               "frac": round(ach * 1e9 / HBM_PEAK, 4), "traffic": t,
               "kernel": name, "bytes_per_launch": k["bytes"], "avg_launch_ms": k["time"] * 1e3,
               "launch_covers": k["units"]}
        if t:
            # HBM bytes actually moved (rocprofv3 counters of the committed profile) over this run's launch time
            out["traffic_gbs"] = round(t / k["time"] / 1e9, 2)
            out["traffic_frac"] = round(t / k["time"] / HBM_PEAK, 4)
            out["traffic_source"] = (f"profiles/{os.path.basename(tpath)}: FETCH_SIZE x2 + WRITE_SIZE per dispatch, "
                                     f"{traffic.get('commit', 'commit not recorded')}")
        if "shared_pass" in k:
            sp = k["shared_pass"] / k["time"]
            out["shared_pass_bytes"] = k["shared_pass"]
            out["shared_pass_gbs"] = round(sp / 1e9, 2)
            out["shared_pass_frac"] = round(sp / HBM_PEAK, 4)
            out["note"] = ("achieved/frac: per-root algorithmic bytes (BFS 16 m_c + 16 n_c, SSSP 24 m_c + 16 n_c) "
                           "summed over the roots of a launch; the roots share the row passes, so this can exceed "
                           "what the launch must move. shared_pass_frac: the least a launch moves with every row "
                           "pass shared (BFS 24 m + 24 n of the union component; SSSP 16 m of the union plus "
                           "8 m_c + 16 n_c per root), at most 1 at HBM speed; traffic_frac: bytes actually moved "
                           "(counters of the committed profile)")
        return out

    other = [k for k in kern if k != dom][0]
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round((t1 - t0) / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Graph500 Kronecker generator on the GPU, user seed (2, 3), U[0,1] f32 weights",
        "config": {"workload": workload, "scale": args.scale, "edgefactor": args.edgefactor,
                   "delta": delta_batch if args.mode == "batched" else delta_plain,
                   "plain_delta": delta_plain, "long_weight": long_weight if args.mode == "batched" else None,
                   "mid_weight": mid_weight if args.mode == "batched" else None,
                   "far_weight": far_weight if args.mode == "batched" else None,
                   "roots": nroots, "mode": args.mode, "tuned": tuned,
                   "sssp_root_order": ("hub distance" if args.root_order else "root rule")
                   if args.mode == "batched" else None,
                   "sssp_root_lags": ("hub distance" if args.root_lag else None) if args.mode == "batched" else None},
        "bfs_gteps": round(me.sum() / tb.sum() / 1e9, 4),
        "sssp_gteps": round(me.sum() / ts.sum() / 1e9, 4),
        "plain": plain,
        "batched_vs_plain": round(value / plain["gteps"], 3) if plain else None,
        "generator_ms": round(gen_ms, 3),
        "root_order_ms": round(order_ms, 3) if order_ms is not None else None,
        "csr_build_ms": round(csr_ms, 3),
        "roofline": roof(dom),
        "roofline_other": roof(other),
        "sssp16": sssp16,
        "validation": checks,
    }
    if args.mode == "plain":
        out["bfs_hm_gteps"] = round(hm(tb, me) / 1e9, 6)
        out["sssp_hm_gteps"] = round(hm(ts, me) / 1e9, 6)
    sess.release(states)
    if args.serve_requests:
        from g500 import serving
        log("[bench] serving leg: loading 8 resident graphs")
        res = serving.load_graphs(sess)
        m = sess.mark()
        out["serving"] = serving.poisson_run(sess, res, args.serve_rate, args.serve_requests)
        sess.release(m)
        log(f"[bench] serving: {out['serving']}")
        sweep = []
        for rate in [float(x) for x in args.serve_sweep.split(",") if x.strip()]:
            m = sess.mark()  # each run's server states are released before the next
            r = serving.poisson_run(sess, res, rate, args.serve_requests)
            sess.release(m)
            sweep.append({k: r[k] for k in ("rate_offered", "completed", "errors", "p50_ms", "p99_ms", "req_per_s",
                                            "gteps", "mean_batch", "arrival_s", "drain_s")})
            log(f"[bench] serving at {rate:g} req/s: {sweep[-1]}")
        out["serving_sweep"] = sweep
    else:
        out["serving"] = None
        out["serving_sweep"] = []
    if not args.no_cpu:
        # the host share of one GPU: OMP_NUM_THREADS where set (16 on the GPU hosts), else the affinity mask
        nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        log(f"[bench] CPU baseline ({nthreads} threads, {cpu_model()}, nproc {os.cpu_count()})")
        out["cpu_baseline"] = cpu_baseline(sess, g, roots, m_c, args.cpu_seconds, nthreads, args.cpu_full)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    sess.close()


if __name__ == "__main__":
    main()
