"""The benchmark's command line without a GPU: more than one GPU is refused
before anything touches a device, and every workload is offered."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + list(args), cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    p = run("--gpus", "2")
    assert p.returncode == 2
    assert "exactly one GPU" in p.stderr and "--gpus 2" in p.stderr
    assert p.stdout == ""


def test_workloads_listed():
    p = run("--help")
    assert p.returncode == 0
    for w in ("spheres", "cornell", "graphs", "cache", "batch", "instances"):
        assert w in p.stdout
