"""bench.py's command-line contract, without a GPU: one GPU only, sane step counts."""
from __future__ import annotations

import json
import subprocess
import sys

from conftest import ROOT


def run_bench(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run_bench("--gpus", "2")
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_step_counts_are_checked():
    for args in (("--steps", "0"), ("--warmup", "-1")):
        r = run_bench(*args)
        assert r.returncode == 2 and "--steps must be >= 1" in r.stderr, args
