"""bench.py end to end: the refusal of several GPUs (CPU), and the config 1
and reduced config 3 lines (GPU): one JSON line each, exact, with the
roofline and per-op fields filled."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args, timeout=150):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                       timeout=timeout, cwd=ROOT)
    return r


def json_line(r):
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1, r.stdout[-2000:]
    return json.loads(lines[0])


def test_bench_refuses_several_gpus():
    r = run_bench("--gpus", "2", timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.gpu
def test_bench_config1_line():
    print("bench c1", flush=True)
    d = json_line(run_bench("--config", "c1", "--steps", "2", "--warmup", "1", "--no-cpu-baseline"))
    assert d["check"]["bit_exact"] is True
    assert d["check"]["inserted"] == 500_000 and d["check"]["removed"] == 100_000
    assert d["value"] > 0 and d["n_gpus"] == 1
    assert set(d["ops"]) == {"insert", "get", "remove", "get2"}
    for op in d["ops"].values():
        assert op["device_ms"] > 0 and 0 < op["frac"] < 1
    assert d["roofline"]["peak"] == 8000.0 and d["stream_copy"]["best_gbs"] > 0


@pytest.mark.gpu
def test_bench_config3_reduced_line():
    print("bench c3 reduced", flush=True)
    d = json_line(run_bench("--config", "c3", "--scale", "0.02", "--steps", "1", "--warmup", "0",
                            "--no-cpu-baseline", "--no-autotune"))
    assert d["check"]["rows_exact"] is True
    assert d["check"]["probes_checked"] == 2_000_000
    assert d["value"] > 0 and 0 < d["roofline"]["frac"] < 1


@pytest.mark.gpu
def test_bench_default_line_has_every_section():
    """The default run (reduced): the account-state line with a section per
    other configuration and the host-buffer calls, every check passing."""
    print("bench default reduced", flush=True)
    d = json_line(run_bench("--scale", "0.02", "--steps", "2", "--warmup", "1", "--tables", "24", "--requests",
                            "60", "--rates", "2000,10000", "--cpu-seconds", "0.5", "--no-autotune", timeout=300))
    assert d["check"]["bit_exact"] is True
    for name in ("c1", "c3", "c4", "c5", "host_api"):
        assert d[name]["ok"] is True, (name, d[name])
    assert d["c1"]["check"]["bit_exact"] is True
    assert d["c3"]["check"]["rows_exact"] is True
    assert d["c4"]["check"]["fused_equals_plain"] and d["c4"]["check"]["oracle_exact"]
    assert d["c4"]["check"]["oracle_tables"] == 24
    assert all(l["mismatches"] == 0 and l["checked"] == 60 for l in d["c5"]["loads"])
    assert d["host_api"]["cap_2M"]["pageable_exact"] and d["host_api"]["cap_2M"]["pinned_exact"]
    assert d["targets"]["all_checks_pass"] is True
    assert d["cpu_baseline"]["cores"] == len(os.sched_getaffinity(0))
    assert len(json.dumps(d, separators=(",", ":"))) < 8000
