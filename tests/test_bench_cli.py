"""bench.py's command-line contract, checked without a GPU: more than one GPU
is refused with a message, and the batched implicit workload refuses its
disabled learned methods before touching the device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout.strip() == ""


def test_c5_learned_methods_are_refused():
    r = run("--workload", "c5", "--c5-methods", "jacobi,learned-sym")
    assert r.returncode != 0
    assert "learned methods are disabled" in (r.stderr + r.stdout)
