"""bench.py's command line: one GPU only.  A request for more GPUs exits with
status 2 and says why, before anything touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60, cwd=ROOT)
    assert p.returncode == 2
    assert "one GPU only" in p.stderr and "--gpus 2" in p.stderr
    assert p.stdout == ""
