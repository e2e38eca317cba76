"""CPU model check of the tile work queue.

lib/dtk_queue_check runs the queue functions of csrc/dtk_queue.hpp (the
same __host__ __device__ code the persistent kernel runs) on many CPU
threads over a random wake graph, for the FIFO, KEY and SEED orders (with
deferral windows of 0 and 3 buckets), each with and without workers that keep
a tile woken during its own visit and visit it again at once.  It checks that tile ids and buckets read back
from the rings stay in range, that no tile is visited by two workers at
once, that no wake is lost and that the workers terminate.  It also checks,
without launching anything, that the grid of the adjoint's tile kernels is
accepted only within HIP's launch limits.  The queue's
replacement of the reference's task queue is dt/src/task_queue.cuh:20-150.
"""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "distance_transform_benchmarks_amd", "lib", "dtk_queue_check")


def test_queue_protocol_model_check():
    if not os.path.exists(CHECK):
        pytest.fail(f"{CHECK} missing: run __graft_entry__.build() first")
    r = subprocess.run([CHECK, "3"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "queue model check passed" in r.stdout
    # the adjoint's tile grid stays within HIP's launch limits
    assert "launch shape checks passed" in r.stdout
    # every order, both claim modes and the multi-thread cases ran
    for order in (0, 1, 2):
        for keep in (0, 1):
            assert f"order={order} window=0 keep={keep} tiles=40x25 threads=24" in r.stdout
    assert "order=1 window=3 keep=1 tiles=40x25 threads=24" in r.stdout
