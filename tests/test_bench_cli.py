"""bench.py's command line on the CPU: more than one GPU is refused with a
message (this build measures one GPU), and bad step counts are rejected."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=300)


def test_more_than_one_gpu_is_refused():
    p = run("--gpus", "2")
    assert p.returncode == 2
    assert "one GPU" in p.stderr or "single MI355X" in p.stderr


def test_bad_step_counts():
    p = run("--steps", "0")
    assert p.returncode != 0
