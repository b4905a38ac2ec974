"""Global diagnostics and the signal-speed maximum on the GPU.

cueva_reduce_globals against the oracle's sequential restatement of
22_writer.f95:818-991 (same cell terms, reference summation order): the
device sums in a fixed tree order, so the sums agree to summation-order
rounding and the two maxima exactly.  cueva_max_speed against the same
per-cell formula evaluated by numpy on the downloaded fields.
"""
import sys

import numpy as np
import pytest

import crafted
import cueva
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("fs,kw", [(cueva.HLLC, {}), (cueva.HLLE, {"bz0": 0.1})], ids=["hllc", "hlle_bz"])
def test_globals_match_sequential_sums(fs, kw):
    say("globals 32x64", fs)
    p, prims = cueva.kh_problem(32, 64, sigma=1.0e6, flux_solver=fs, rotated=True, **kw)
    o = Oracle(p)
    o.upload(prims)
    o.step(3)
    ref = o.globals()
    with cueva.Solver(p) as s:
        s.upload(prims)
        s.step(3)
        s.step_counters()
        got = np.array(list(s.globals().values()))
        again = np.array(list(s.globals().values()))
    assert np.array_equal(got, again), "reduction is not deterministic"
    names = cueva.GLOBAL_NAMES
    for k, name in enumerate(names):
        if name.startswith("max_"):
            assert got[k] == ref[k], (name, got[k], ref[k])
        else:
            # area 2, fields O(1): summation-order rounding stays far below 1e-12
            assert abs(got[k] - ref[k]) <= 1e-12 * max(1.0, abs(ref[k])), (name, got[k], ref[k])


def test_globals_sum_d_is_the_area_weighted_d_plane():
    """sum_D is the sum of D*Delx*Dely over the interior of the evolved D plane."""
    p, prims = cueva.kh_problem(32, 64, sigma=1.0e3, rotated=True)
    with cueva.Solver(p) as s:
        s.upload(prims)
        s.step(2)
        s.step_counters()
        g = s.globals()
        pr, co = s.download()
    d = co[0, 6:6 + p.jmax + 1, 6:6 + p.imax + 1]
    dv = cueva.derived(p)
    dx, dy = dv["dx"], dv["dy"]
    assert g["sum_D"] == pytest.approx(float(np.sum(d * dx * dy)), rel=1e-13)


def acoustic(gam, rho, p, v2, vn):
    h = rho + (gam / (gam - 1.0)) * p
    cs2 = (gam * p) / h
    root = np.sqrt(cs2 * (1.0 - v2) * ((1.0 - v2 * cs2) - vn * vn * (1.0 - cs2)))
    den = 1.0 - v2 * cs2
    lp = (vn * (1.0 - cs2) + root) / den
    lm = (vn * (1.0 - cs2) - root) / den
    return np.maximum(np.abs(lp), np.abs(lm))


@pytest.mark.parametrize("v0", [0.5, 0.9])
def test_max_speed(v0):
    say("max_speed v0", v0)
    p, prims = cueva.kh_problem(64, 128, sigma=1.0e3, v0=v0, rotated=False)
    with cueva.Solver(p) as s:
        s.upload(prims)
        s.step(1)
        s.step_counters()
        got = s.max_speed()
        pr = s.download(cons=False)
    sl = (slice(6, 6 + p.jmax + 1), slice(6, 6 + p.imax + 1))
    vx, vy, vz = pr[5][sl], pr[6][sl], pr[7][sl]
    rho, pp = pr[12][sl], pr[13][sl]
    v2 = (vx * vx + vy * vy) + vz * vz
    want = max(acoustic(p.gamma, rho, pp, v2, vx).max(), acoustic(p.gamma, rho, pp, v2, vy).max())
    assert got == pytest.approx(want, rel=1e-15, abs=0)
    assert 0.0 < got < 1.0  # bounded by the light speed that sets the reference's dt


@pytest.mark.parametrize("name", ["kh_hllc_mp5_32x64", "kh_hlle_mp7_bz_32x64", "kh_cfg1_hll_mp5_128x256"])
def test_globals_match_reference_lector(name):
    """Against the reference's own lector output after the fixture's last step."""
    import json
    import os
    meta = json.load(open(os.path.join(os.path.dirname(__file__), "golden", name + ".json")))
    if "globals_final" not in meta:
        pytest.skip("fixture without the reference's global diagnostics")
    ni, nj = meta["native_grid"]
    say("globals vs reference", name)
    p, prims = crafted.problem(meta)
    with cueva.Solver(p) as s:
        s.upload(prims)
        s.step(meta["nsteps"])
        s.step_counters()
        got = s.globals()
    spread = max(meta["dumps"][str(meta["nsteps"])]["libm_spread"].values())
    for n, want in meta["globals_final"].items():
        if n.startswith("max_") and spread == 0.0:
            assert got[n] == want, (n, got[n], want)
        else:
            # summation order, and beyond the bitwise fixtures the reference's own libm spread
            tol = max(1e-12, 10 * spread)
            assert abs(got[n] - want) <= tol * max(1.0, abs(want)), (n, got[n], want)
