"""bench.py contract on the host: --gpus other than 1 exits with a message,
and every section the default run produces has its function."""

import importlib
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_multi_gpu_request_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_sections_exist():
    sys.path.insert(0, ROOT)
    bench = importlib.import_module("bench")
    for name in ("run_scan", "run_tuned", "run_large", "run_learned", "cpu_baseline", "flops_per_sample",
                 "calibration_copy", "main"):
        assert callable(getattr(bench, name)), name
