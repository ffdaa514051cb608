"""bench.py's command line, without a GPU: more than one GPU is refused with
a message before anything touches a device, and the argument parser knows
every option the profile commands use."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_profile_command_options_parse():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--no-cpu-baseline", "--serve-jobs", "--traffic", "--one-group-frames",
                "--copy-probe", "--serve-rates", "--workload", "--autotune", "--tuning", "--serve-config"):
        assert opt in r.stdout, opt
