"""bench.py's command line (CPU): a GPU count other than one is refused with
a message before anything touches the GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8, 0])
def test_gpu_counts_other_than_one_are_refused(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)], capture_output=True,
                       text=True, timeout=60, cwd=ROOT)
    assert r.returncode == 2
    assert f"--gpus {n} is not supported" in r.stderr
    assert r.stdout == ""  # no JSON line
