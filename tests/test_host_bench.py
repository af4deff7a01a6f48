"""Host-side checks of bench.py's command line (no GPU: these exit before any device work)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    p = _bench("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert p.returncode == 2
    msg = json.loads(p.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"] and "--gpus 1" in msg["error"]


def test_unknown_workload_is_rejected():
    p = _bench("--workload", "no_such_workload")
    assert p.returncode == 2 and "invalid choice" in p.stderr
