"""bench.py's command-line contract, checked without a GPU: more than one GPU
is refused with a message and exit status 2 before any device is touched,
and the flags the measurement commands use exist."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True,
                          text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""  # no JSON line


def test_flags_of_the_measurement_commands():
    r = _run("--help")
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--nfreq", "--fp64-sample", "--no-cpu", "--no-config0",
                 "--no-peaks", "--serve-requests", "--copy-calibration", "--assembly-nel", "--cpu-procs",
                 "--node-results"):
        assert flag in r.stdout, flag
