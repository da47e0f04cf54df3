"""bench.py's command-line contract, on the CPU: more than one GPU is refused with a
message (the batched path keeps every source of a shot set on one GPU), and the options
that select what is measured are there."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_options():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--sweep", "--autotune", "--streams",
                "--no-2d", "--no-serve", "--no-cpu", "--copy-calibration", "--driver2d", "--driver3d"):
        assert opt in r.stdout, opt
