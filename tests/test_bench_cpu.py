"""bench.py's command line on the CPU: the one-GPU scope and the side-measurement
selection fail with a message before any GPU work (the measurements themselves
run in the GPU suite's bench runs)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr and r.stdout == ""


def test_unknown_side_measurement_is_refused():
    r = _bench("--only", "tv_256,nonesuch")
    assert r.returncode != 0
    assert "unknown side measurement" in r.stderr and "nonesuch" in r.stderr
    for name in ("batch", "cnn", "demo_384", "tgv_512", "tv_256"):
        assert name in r.stderr
    assert r.stdout == ""


def test_help_lists_the_options():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--only", "--no-graph", "--no-autotune", "--fixed-iters", "--calibrate-copy", "--no-warm"):
        assert opt in r.stdout
