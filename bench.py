#!/usr/bin/env python3
"""Graph500 BFS + SSSP benchmark on one MI355X.

Workload (default): Kronecker graph SCALE 26, edgefactor 16 (seed 2, 3),
generated and CSR-built on the GPU, then one step per root: the plain
single-root BFS and delta-stepping SSSP (delta 0.1) through the C-ABI. All
inputs are resident in HBM before the timed region. Per-root times are the
kernels' HIP-event durations on the library stream; TEPS follow the Graph500
runner (m_c = undirected component edges, harmonic mean over roots);
`value` is the combined harmonic mean 2 / (1/HM_BFS + 1/HM_SSSP) in GTEPS.

Prints one JSON line (see the keys below). Options exist for smaller scales
and for the CPU baseline sample; --gpus above 1 is rejected (one GPU).
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


def hm(times_s, edges):
    return len(times_s) / float(np.sum(np.asarray(times_s) / np.asarray(edges, dtype=np.float64)))


def cpu_baseline(sess, g, roots, m_c, seconds, nthreads):
    """The CPU oracle on the host cores, on a bounded sample of the same graph:
    an OpenMP level-synchronous BFS from the first root (whole traversal) and
    a binary-heap FP32 Dijkstra from the same root stopped after `seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    rs, col, w = sess.csr_to_host(g)
    ref = O.Csr(int(g.nlocalverts), rs, col, w)
    root = int(roots[0])
    t0 = time.perf_counter()
    O.bfs_omp(ref, root, nthreads)
    tb = time.perf_counter() - t0
    scanned, settled, ts = O.dijkstra_f32_timed(ref, root, seconds)
    teps_b = m_c[0] / tb
    teps_s = (scanned / 2.0) / ts  # each undirected edge is scanned from both ends
    value = 2.0 / (1.0 / teps_b + 1.0 / teps_s) / 1e9
    return {
        "value": value, "unit": "GTEPS", "cores": nthreads, "kind": "port",
        "sample": (f"oracle on the same SCALE {int(np.log2(g.nlocalverts))} CSR, root {root}: OpenMP "
                   f"top-down BFS ({nthreads} threads, whole traversal, {tb:.2f} s, "
                   f"{teps_b / 1e9:.3f} GTEPS) and serial binary-heap FP32 Dijkstra (1 thread) stopped "
                   f"after {ts:.1f} s with {settled} vertices settled ({teps_s / 1e9:.4f} GTEPS as scanned "
                   f"entries/2 per second); value = 2/(1/BFS + 1/SSSP)"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64, help="roots timed (BFS + SSSP each)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--delta", type=float, default=0.1)
    ap.add_argument("--arena-gib", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
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
    roots = sess.roots(g)
    nroots = len(roots)
    bfs = sess.bfs_create(g)
    sssp = sess.sssp_create(g, args.delta)
    need = min(nroots, args.steps + args.warmup)

    # per-root TEPS numerators and reached counts (untimed, as in the runner)
    m_c = np.zeros(nroots, np.float64)
    n_c = np.zeros(nroots, np.float64)
    for i in range(need):
        sess.bfs(g, bfs, roots[i])
        m_c[i] = sess.edge_count(g, bfs.pred)
        n_c[i] = sum(sess.level_log(bfs)[1])
    log(f"[bench] {nroots} roots; edges per root ~{np.median(m_c[:need]):.4g}, reached ~{np.median(n_c[:need]):.4g}; "
        f"smallest components (edges): {sorted(m_c[:need])[:4]}")

    for i in range(args.warmup):
        r = roots[(args.steps + i) % nroots]
        sess.bfs(g, bfs, r)
        sess.sssp(g, sssp, r)

    tb = np.zeros(args.steps)
    ts = np.zeros(args.steps)
    idx = [i % nroots for i in range(args.steps)]
    # every entry point is blocking, so the device is idle when the clock starts and stops
    t0 = time.perf_counter()
    for i, ri in enumerate(idx):
        tb[i] = sess.bfs(g, bfs, roots[ri]) * 1e-3
        ts[i] = sess.sssp(g, sssp, roots[ri]) * 1e-3
    t1 = time.perf_counter()
    for i in np.argsort(tb)[:3]:
        log(f"[bench] root {int(roots[idx[i]])}: m_c {m_c[idx[i]]:.0f} bfs {tb[i]*1e3:.3f} ms sssp {ts[i]*1e3:.3f} ms")
    log(f"[bench] bfs ms median {np.median(tb)*1e3:.3f} max {tb.max()*1e3:.3f}; "
        f"sssp ms median {np.median(ts)*1e3:.3f} max {ts.max()*1e3:.3f}")
    me = m_c[idx]
    nv = n_c[idx]
    hm_b = hm(tb, me)
    hm_s = hm(ts, me)
    value = 2.0 / (1.0 / hm_b + 1.0 / hm_s) / 1e9

    bytes_b = 16.0 * me + 16.0 * nv
    bytes_s = 24.0 * me + 16.0 * nv
    kern = {
        "bfs_kernel": {"bytes": float(bytes_b.mean()), "time": float(tb.mean()), "total": float(tb.sum())},
        "sssp_kernel": {"bytes": float(bytes_s.mean()), "time": float(ts.mean()), "total": float(ts.sum())},
    }
    dom = max(kern, key=lambda k: kern[k]["total"])
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", f"traffic_s{args.scale}_ef{args.edgefactor}.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath))

    def roof(name):
        k = kern[name]
        ach = k["bytes"] / k["time"] / 1e9
        return {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(ach * 1e9 / HBM_PEAK, 4), "traffic": traffic.get(name),
                "kernel": name, "bytes_per_launch": k["bytes"], "avg_launch_ms": k["time"] * 1e3}

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
        "config": {"workload": (f"plain single-root BFS + delta-stepping SSSP per step, Kronecker SCALE "
                                f"{args.scale} edgefactor {args.edgefactor}, roots from the Graph500 root rule"),
                   "scale": args.scale, "edgefactor": args.edgefactor, "delta": args.delta,
                   "roots": nroots, "mode": "plain"},
        "bfs_hm_gteps": round(hm_b / 1e9, 4),
        "sssp_hm_gteps": round(hm_s / 1e9, 4),
        "generator_ms": round(gen_ms, 3),
        "csr_build_ms": round(csr_ms, 3),
        "roofline": roof(dom),
        "roofline_other": roof("sssp_kernel" if dom == "bfs_kernel" else "bfs_kernel"),
    }
    if not args.no_cpu:
        nthreads = min(16, len(os.sched_getaffinity(0)))
        log(f"[bench] CPU baseline sample ({nthreads} threads)")
        out["cpu_baseline"] = cpu_baseline(sess, g, roots, m_c, args.cpu_seconds, nthreads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    sess.close()


if __name__ == "__main__":
    main()
