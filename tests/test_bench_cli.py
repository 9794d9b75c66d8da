"""bench.py's argument handling on the CPU: more than one GPU is refused
with a message (every run here has one GPU), and with no GPU visible the
measurement fails loudly instead of falling back to the host."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args],
                          capture_output=True, text=True, timeout=300)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "single MI355X" in r.stderr and r.stdout == ""


def test_no_gpu_fails_loudly():
    r = _run("--steps", "1", "--warmup", "0", "--extras", "")
    assert r.returncode == 1
    assert "no GPU visible" in r.stderr and r.stdout == ""


def test_help_lists_the_modes():
    r = _run("--help")
    assert r.returncode == 0
    for word in ("--mode", "--extras", "--tc-layout", "--tc-merge-max"):
        assert word in r.stdout
