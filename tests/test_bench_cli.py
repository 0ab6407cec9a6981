"""bench.py's command line on the CPU: more than one GPU is refused with a
message before anything touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "single GPU" in r.stderr and r.stdout == ""


def test_unknown_section_is_an_argument_error():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--sections", "nope"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "unknown sections" in r.stderr
