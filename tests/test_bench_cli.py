"""bench.py command line (CPU): the one-GPU contract and the workload list."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BENCH = os.path.join(os.path.dirname(HERE), "bench.py")


def run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 1" in r.stderr
    assert r.stdout == ""  # no JSON line


def test_world_size_above_one_is_refused():
    env = dict(os.environ, WORLD_SIZE="2")
    r = subprocess.run([sys.executable, BENCH], capture_output=True, text=True, timeout=120,
                       env=env)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_help_lists_every_workload():
    r = run("--help")
    assert r.returncode == 0
    for w in ("solvent10k", "vacuum1k", "helix60k", "batch512", "served", "direct", "frames"):
        assert w in r.stdout
