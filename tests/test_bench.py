"""bench.py contract: one JSON line with the required keys; one GPU only."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_bench_rejects_more_than_one_gpu():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_bench_json_line_small(mode):
    cmd = [sys.executable, BENCH, "--scale", "14", "--steps", "2", "--warmup", "1", "--mode", mode,
           "--plain-roots", "8", "--arena-gib", "8", "--serve-requests", "0", "--cpu-seconds", "2"]
    if mode == "plain":
        cmd.append("--no-cpu")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    out = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in out, k
    assert out["n_gpus"] == 1 and out["steps"] == 2 and out["value"] > 0
    assert out["config"]["mode"] == mode
    roof = out["roofline"]
    # at SCALE 14 the kernels are latency-bound: frac can round to 0 at 4 decimals
    assert roof["bound"] == "hbm" and roof["peak"] == 8000.0 and roof["achieved"] > 0 and roof["frac"] < 1, roof
    v = out["validation"]
    assert v["plain_bfs"] > 0 and v["plain_sssp"] > 0
    assert v["plain_bfs_failures"] == 0 and v["plain_sssp_failures"] == 0
    if mode == "batched":
        assert out["plain"]["roots"] == 8 and out["batched_vs_plain"] > 0
        assert v["batched_bfs"] == 64 and v["batched_sssp"] == 64
        assert v["batched_bfs_failures"] == 0 and v["batched_sssp_failures"] == 0
        assert out["sssp16"]["launches"] == 4 and out["sssp16"]["gteps"] > 0
        cpu = out["cpu_baseline"]
        assert cpu["kind"] == "port" and cpu["value"] > 0 and cpu["cores"] >= 1
        assert len(cpu["bfs_omp"]["roots"]) >= 2 and len(cpu["sssp_omp"]["roots"]) >= 1
