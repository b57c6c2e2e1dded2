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


def test_cpu_baseline_threads_and_host_record(monkeypatch):
    """The CPU baseline uses OMP_NUM_THREADS where the host sets it (the
    CPU share of this GPU), else every physical core, and records the host."""
    import importlib
    sys.path.insert(0, ROOT)
    bench = importlib.import_module("bench")
    monkeypatch.setenv("OMP_NUM_THREADS", "5")
    assert bench.default_cpu_threads() == 5
    monkeypatch.delenv("OMP_NUM_THREADS")
    info = bench.host_info()
    assert bench.default_cpu_threads() == info.get("physical_cores", os.cpu_count())
    assert info["nproc"] == os.cpu_count() and "model" in info
    med, ts = bench.timed_cpu(iter([3.0, 1.0, 2.0, 5.0, 4.0, 9.0]).__next__)
    assert (med, ts) == (4.0, [1.0, 2.0, 5.0, 4.0, 9.0])
