"""The bench's compulsory-traffic model and its one-GPU contract (CPU only).

The byte counts are the published per-scheme totals of the compulsory HBM
traffic of one step (FP64 planes, ghosts ignored): SSP2(2,2,2) reads 55 and
writes 36 doubles per zone-update, SSP3(4,3,3) reads 178 and writes 70.
"""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_compulsory_bytes_per_zone_update():
    assert bench.step_bytes_per_zone(2) == 8 * (55 + 36) == 728
    assert bench.step_bytes_per_zone(4) == 8 * (178 + 70) == 1984
    # one stage: read U^n, P^n (19 planes), write U^{n+1}, P^{n+1}
    assert bench.step_bytes_per_zone(1) == 8 * (19 + 19)
    # each extra stage adds more than the one before (it re-reads every earlier right-hand side)
    d = [bench.step_bytes_per_zone(l + 1) - bench.step_bytes_per_zone(l) for l in range(1, 6)]
    assert all(b > a for a, b in zip(d, d[1:]))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
