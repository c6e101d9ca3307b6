"""bench.py's command line on the CPU: a run asked for more than one GPU stops with a message and
exit status 2 before it touches a device (the kinetic step runs on one GPU; the velocity sharding
across GPUs is a design only)."""

from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    for argv, env in ((["--gpus", "2"], {}), ([], {"WORLD_SIZE": "2"})):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *argv], capture_output=True, text=True,
                           env={**os.environ, **env}, timeout=120)
        assert r.returncode == 2, (argv, r.returncode, r.stderr)
        assert "one GPU" in r.stderr and r.stdout == ""
