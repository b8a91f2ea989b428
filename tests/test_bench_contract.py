"""bench.py's output contract: one JSON line with the fields the headline
metric is read from (GPU), and a refusal of more than one GPU (CPU)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU only" in p.stderr
    assert p.stdout.strip() == ""


@pytest.mark.gpu
def test_headline_line_fields():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--steps", "20", "--warmup", "2",
                        "--no-cpu", "--no-fp32", "--no-extras"], cwd=ROOT, capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    lines = [l for l in p.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    out = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline"):
        assert k in out, k
    assert out["metric"] == "cell-steps/sec" and out["unit"] == "cell-steps/s"
    assert out["n_gpus"] == 1 and out["steps"] == 20 and out["warmup"] == 2
    assert out["dtype"] == "f64" and out["higher_is_better"] is True
    assert out["config"]["workload"] == "floodplain_207234" and out["config"]["cells"] == 207234
    # value is cells x steps / time, the roofline its algorithmic bytes
    assert abs(out["value"] - 207234 / (out["ms_per_step"] / 1e3)) <= 1e-6 * out["value"]
    r = out["roofline"]
    assert r["bound"] == "hbm" and r["peak"] == 8000.0 and r["unit"] == "GB/s"
    assert abs(r["achieved"] - out["value"] * 212 / 1e9) <= 1e-6 * r["achieved"]
    assert abs(r["frac"] - r["achieved"] / r["peak"]) <= 1e-12
    assert 0 < r["frac"] < 1
