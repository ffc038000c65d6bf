"""Command-line entry points (CPU): their options parse, and bench.py refuses
more than one GPU with a message and exit status 2."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, *args], cwd=ROOT, capture_output=True, text=True,
                          timeout=300)


def test_trace_bench_options():
    r = _run("-m", "ddhdg_amd.trace_bench", "--help")
    assert r.returncode == 0
    for opt in ("--dim", "--cells", "--degree", "--reps"):
        assert opt in r.stdout


def test_bench_options():
    r = _run("bench.py", "--help")
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--cells3d-k2", "--serve"):
        assert opt in r.stdout


def test_bench_refuses_several_gpus():
    r = _run("bench.py", "--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
