"""The bench command line on the host: more than one GPU is refused with a
message and exit code 2 before anything touches a device, and the options
the driver and the side workloads use parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "4", "--steps", "3"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout and "--gpus 4" in r.stdout


def test_help_lists_the_workloads():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--batched", "--serve", "--no-learned", "--no-small",
                "--backward-euler", "--viscous"):
        assert opt in r.stdout, opt
