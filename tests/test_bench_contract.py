"""bench.py's command-line contract: one GPU only, and the one JSON line with
the fields its readers rely on (roofline, cpu_baseline, dense-input leg)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, timeout):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=timeout)


def test_bench_refuses_more_than_one_gpu():
    p = _run(["--gpus", "2"], 120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout.strip() == ""


def test_bench_rejects_unknown_leg():
    p = _run(["--leg", "nonsense"], 120)
    assert p.returncode == 2
    assert "--leg" in p.stderr


@pytest.mark.gpu
def test_bench_prints_one_json_line_with_the_contract_fields(cuda):
    print("[case] bench.py --steps 2 --warmup 1 (no cpu, batched, beijing legs)", flush=True)
    p = _run(["--steps", "2", "--warmup", "1", "--no-cpu-baseline", "--no-batched",
              "--no-beijing", "--quiet"], 600)
    assert p.returncode == 0, p.stderr[-2000:]
    lines = [ln for ln in p.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    line = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline", "dense_inputs"):
        assert k in line, k
    assert line["n_gpus"] == 1 and line["steps"] == 2 and line["warmup"] == 1
    assert line["unit"] == "Grays/s" and line["value"] > 0 and line["higher_is_better"]
    assert line["config"]["workload"] and "model" not in line["config"]
    roof = line["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in roof, k
    # bound = the binding resource against its physical peak, never above it
    assert roof["bound"] in ("valu", "hbm")
    assert 0 < roof["frac"] <= 1
    assert abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-9
    assert roof[{"valu": "valu_issue", "hbm": "hbm"}[roof["bound"]]]["frac"] == roof["frac"]
    assert roof["hbm"]["peak"] == 8000.0 and 0 < roof["hbm"]["frac"] <= 1
    assert roof["valu_issue"]["frac"] <= 1
    assert roof["frac"] >= max(roof["hbm"]["frac"], roof["valu_issue"]["frac"])
    # the algorithmic (LDS-served) tap figure is reported, not used as a bound
    assert roof["tap_gbs"] > 0 and roof["algorithmic_bytes"]["per_sample"] == 32
    # value = rays per pair / seconds per pair
    assert abs(line["value"] - 8847360 / (line["ms_per_step"] * 1e-3) / 1e9) \
        < 1e-6 * line["value"]
    dense = line["dense_inputs"]
    assert dense["adjoint_ms_median"] > 0 and 0 < dense["phantom_sino_nonzero_ray_frac"] < 1
