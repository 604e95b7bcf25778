"""bench.py's command line on the host: more than one GPU is refused with a
message (one grid's transform runs on one GPU) before anything touches a
device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 2
    assert "only one GPU" in out.stderr
    assert out.stdout == ""
