"""bench.py's command line without a GPU: it measures one GPU only and says
so for any other count, before touching a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU only" in (r.stdout + r.stderr)
    assert r.stdout.strip() == "" or not r.stdout.strip().startswith("{")


def test_bench_rejects_unknown_workloads():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--workload", "preflop"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "invalid choice" in r.stderr
