"""bench.py's command line on the CPU: more than one GPU is refused with a
message (the counting path runs on one GPU), and the options parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_help_lists_the_workloads():
    r = _run("--help")
    assert r.returncode == 0
    for opt in ("--config", "--patterns", "--sample-frac", "--batched-scale", "--served-scales", "--autotune"):
        assert opt in r.stdout
    for cfg in ("rmat22", "rmat13", "chunglu3m"):
        assert cfg in r.stdout
