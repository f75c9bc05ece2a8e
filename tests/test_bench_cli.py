"""bench.py command line: runs that the single-GPU backend does not support end before
any device work, with a message and a nonzero status."""
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
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_unknown_config_is_refused():
    r = _run("--config", "no-such-config")
    assert r.returncode != 0
    assert "invalid choice" in r.stderr
