"""bench.py's command-line contract on the host: one GPU only, and the pieces of
the JSON line that do not need a device."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2", "--steps", "1", "--warmup", "0"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_line_reports_bandwidth_from_the_kernel_time():
    import bench
    r = dict(mlups=1000.0, kernel_ms=0.5, steps_done=10, resets=0, cells=1_000_000)
    out = bench.line(r, 177)
    assert out["achieved_gbs"] == round(177 * 1_000_000 / 0.5e-3 / 1e9, 1)
    assert out["pct_hbm"] == round(1000.0 * 1e6 * 177 / 8e12, 4)
    json.dumps(out)


def test_counter_cases_name_the_measured_lattices():
    import bench
    assert bench.PMC_CASES["c2"] == dict(scale=1.0, precision="fp32")
    assert {c["precision"] for c in bench.PMC_CASES.values()} == {"fp32", "fp16", "bf16", "fp16_eq"}
    assert np.isclose(bench.HBM_PEAK_GBS, 8000.0)
