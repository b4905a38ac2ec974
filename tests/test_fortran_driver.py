"""The Fortran host driver (mhd-kh-instability_amd/fortran/): the reference's
time loop calling reference-named shims over the iso_c_binding interface.

It must produce the same fields, bit for bit, as the fused cueva_step on the
same data, and its per-hglb global diagnostics must be cueva_reduce_globals'.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import cueva

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mhd-kh-instability_amd", "lib", "cueva_gpu_driver")


def test_driver_built_and_checks_arguments():
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (needs flang)")
    res = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert res.returncode == 2 and "usage" in res.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("fs,rec,order,kw", [(cueva.HLLC, 0, 2, {}), (cueva.HLLE, 1, 2, {"bz0": 0.1}),
                                             (cueva.HLL, 2, 4, {})], ids=["hllc_mp5", "hlle_mp7", "hll_mp9_ssp3"])
def test_driver_matches_fused_step(tmp_path, fs, rec, order, kw):
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (needs flang)")
    ni, nj, steps, sigma = 32, 64, 3, 1.0e6
    bz0 = kw.get("bz0", 0.0)
    prefix = str(tmp_path / "run")
    print("driver", fs, rec, order, flush=True)
    args = [DRIVER, str(ni), str(nj), str(steps), str(fs), str(rec), str(order), repr(sigma), "1", "0.5", "0.01",
            repr(bz0), "20261202", "1", prefix]
    res = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    p, prims = cueva.kh_problem(ni, nj, sigma=sigma, flux_solver=fs, rec_mp=rec, order=order, rotated=True, **kw)
    raw = open(prefix + ".bin", "rb").read()
    cnt = np.frombuffer(raw[:12], dtype="<i4").tolist()
    arr = np.frombuffer(raw[12:], dtype="<f8").reshape(19, p.jmax + 13, p.imax + 13)
    glob = np.loadtxt(prefix + "_globals.dat", ndmin=2)
    with cueva.Solver(p) as s:
        s.upload(prims)
        want_glob = []
        for h in range(steps):
            s.step(1)
            c, _ = s.step_counters()
            want_glob.append(list(s.globals().values()))
        pr, co = s.download()
    assert cnt == c
    ours = np.concatenate([pr, co])
    sl = (slice(None), slice(6, 6 + p.jmax + 1), slice(6, 6 + p.imax + 1))
    assert np.array_equal(arr[sl], ours[sl])
    assert glob.shape == (steps, 39)
    assert np.array_equal(glob[:, 2:], np.array(want_glob))
    assert np.array_equal(glob[:, 0], np.arange(1, steps + 1))
