"""The marching face kernel against the face-plane path.

cueva_step's default face pass (k_faces_march) reconstructs, solves the
Riemann problems and forms the flux differences of a strip of 63 columns in
one launch, keeping the face fluxes in LDS; on the last stage it also applies
sumfin + varfin + electricfield, and the primitive recovery follows in its
own pass.  CUEVA_MARCH=0 selects the face-plane path (k_faces writes every
face flux to HBM, k_fluxandsource / k_final_fused difference them), which the
per-subroutine entry points also use.

* FP64: both paths evaluate the same expressions in the same order, so the
  fields are bit-identical, for every solver, reconstruction order and IMEX
  scheme, on grids of one and of two (ragged) strips, with segments of 3
  rows, 23 rows (ragged last segment), the whole grid, and the length sized
  for the device.
* Mixed form: the FP32 faces are the same code in the same unit, the FP32
  flux differences the same operations, and the FP64 sources and update are
  evaluated without contraction in both units: bit-identical as well.
* Launch shape: the XCD-aware and the plain 2D block order, and launches
  cut into 1-4 resident-block rounds, march the same tiles with the same
  arithmetic: bit-identical fields and counters.
"""
import os
import sys

import numpy as np
import pytest

import cueva
from parity import interior, ndiff, rel_l1

pytestmark = pytest.mark.gpu

# name: (flux solver, rec_mp, IMEX order, sigma, init keywords)
CASES = {
    "hll_mp5": (cueva.HLL, 0, 2, 1.0e3, {}),
    "hllc_mp5": (cueva.HLLC, 0, 2, 1.0e6, {}),
    "hllc_mp7": (cueva.HLLC, 1, 2, 1.0e6, {}),
    "hlle_mp7_bz": (cueva.HLLE, 1, 2, 1.0e6, {"bz0": 0.1}),
    "hll_mp9_ssp3": (cueva.HLL, 2, 4, 1.0e3, {}),
    "hllc_mp9": (cueva.HLLC, 2, 2, 1.0e6, {}),
    "hll_mcl_bz": (cueva.HLL_MCL, 0, 2, 1.0e6, {"bz0": 0.1}),
    "hll_mcl_ssp3": (cueva.HLL_MCL, 0, 4, 1.0e3, {}),
    "lax_bz": (cueva.LAX, 0, 2, 1.0e6, {"bz0": 0.1}),
    "hllc_mcl_bz": (cueva.HLLC_MCL, 0, 2, 1.0e6, {"bz0": 0.1}),
    "hlle_mcl_bz": (cueva.HLLE_MCL, 0, 2, 1.0e6, {"bz0": 0.1}),
    "hllc_mp5_primrec": (cueva.HLLC, 0, 2, 1.0e6, {"primitive_rec": 1}),
    "hllc_mcl_superbee": (cueva.HLLC_MCL, 0, 2, 1.0e6, {"slop": 3}),
    "hllc_mp7_lum": (cueva.HLLC, 1, 3, 1.0e6, {}),
    "lax_minmod_primrec_superbee": (cueva.LAX, 0, 2, 1.0e3, {"slop": 1, "primitive_rec": 1, "limit": 3}),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(p, prims, steps, march, seg=None, xcd=None, rounds=None):
    knobs = {"CUEVA_MARCH_SEG": seg, "CUEVA_MARCH_XCD": xcd, "CUEVA_MARCH_ROUNDS": rounds}
    old = {k: os.environ.get(k) for k in ("CUEVA_MARCH", *knobs)}
    os.environ["CUEVA_MARCH"] = "1" if march else "0"
    for k, v in knobs.items():  # read when the context is created
        if v is not None:
            os.environ[k] = str(v)
        else:
            os.environ.pop(k, None)
    try:
        with cueva.Solver(p) as s:
            s.upload(prims)
            counters = []
            for _ in range(steps):
                s.step(1)
                counters.append(s.step_counters())
            return interior(*s.download(), p), counters
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("shape", [(70, 37), (130, 68)])
def test_march_equals_face_planes_fp64(name, shape):
    fs, rec, order, sigma, kw = CASES[name]
    ni, nj = shape
    p, prims = cueva.kh_problem(ni, nj, sigma=sigma, flux_solver=fs, rec_mp=rec, order=order, rotated=True, **kw)
    steps = 3 if order == 4 else 5  # SSP3(4,3,3) stops within 8 steps on this problem
    say("case", name, "grid %dx%d" % (p.imax, p.jmax), "steps", steps)
    ref, cref = run(p, prims, steps, march=False)
    for seg in (None, 3, 23, 4095):  # sized for the device, shortest, ragged, one segment
        got, cgot = run(p, prims, steps, march=True, seg=seg)
        nd = ndiff(got, ref)
        say("  segment rows", seg, "differing points", sum(nd.values()))
        assert cgot == cref
        assert np.array_equal(got, ref), nd


@pytest.mark.parametrize("name", ["hllc_mp5", "hlle_mp7_bz", "hll_mp9_ssp3", "lax_bz"])
@pytest.mark.parametrize("precision", [cueva.FP64, cueva.MIXED])
def test_block_order_and_rounds_do_not_change_bits(name, precision):
    """The XCD-aware block order (the default), the plain 2D order, and the
    launch cut into 1, 2, 3 or 4 resident-block rounds march the same tiles
    with the same arithmetic: bit-identical fields."""
    fs, rec, order, sigma, kw = CASES[name]
    p, prims = cueva.kh_problem(197, 150, sigma=sigma, flux_solver=fs, rec_mp=rec, order=order, rotated=True,
                                precision=precision, crmath=0, **kw)
    steps = 3 if order == 4 else 4
    say("case", name, "FP64" if precision == cueva.FP64 else "mixed", "grid %dx%d" % (p.imax, p.jmax))
    ref, cref = run(p, prims, steps, march=True)
    for xcd, rounds, seg in ((0, None, None), (0, 1, None), (1, 3, None), (1, 1, 23), (0, 2, 23)):
        got, cgot = run(p, prims, steps, march=True, seg=seg, xcd=xcd, rounds=rounds)
        nd = ndiff(got, ref)
        say("  xcd", xcd, "rounds", rounds, "segment rows", seg, "differing points", sum(nd.values()))
        assert cgot == cref
        assert np.array_equal(got, ref), nd


@pytest.mark.parametrize("name", ["hllc_mp5", "hlle_mp7_bz", "hll_mp9_ssp3", "hll_mcl_bz", "lax_bz", "hllc_mcl_bz",
                                  "hlle_mcl_bz"])
def test_march_matches_face_planes_mixed(name):
    fs, rec, order, sigma, kw = CASES[name]
    p, prims = cueva.kh_problem(130, 68, sigma=sigma, flux_solver=fs, rec_mp=rec, order=order, rotated=True,
                                precision=cueva.MIXED, crmath=0, **kw)
    say("case", name, "mixed", "grid %dx%d" % (p.imax, p.jmax))
    steps = 3 if order == 4 else 5
    ref, _ = run(p, prims, steps, march=False)
    got, _ = run(p, prims, steps, march=True)
    nd = ndiff(got, ref)
    say("  differing points", sum(nd.values()), "max rel L1 %.2e" % max(rel_l1(got, ref).values()))
    assert np.array_equal(got, ref), nd


def test_march_learned_matches_face_planes():
    from cueva import learned_c2p as lc
    if not os.path.exists(lc.PARAMS_FILE):
        pytest.skip("no trained parameters shipped")
    packed = lc.load_params()
    p, prims = cueva.kh_problem(130, 68, sigma=1.0e6, flux_solver=cueva.HLLC, rotated=True, c2p_mode=cueva.C2P_LEARNED)
    out = {}
    for march in (False, True):
        os.environ["CUEVA_MARCH"] = "1" if march else "0"
        try:
            with cueva.Solver(p) as s:
                s.set_c2p_params(packed)
                s.upload(prims)
                s.step(5)
                c, fb = s.step_counters()
                out[march] = (interior(*s.download(), p), c, fb)
        finally:
            os.environ.pop("CUEVA_MARCH", None)
    say("fallbacks", out[False][2], out[True][2])
    assert out[True][1] == out[False][1] and out[True][2] == out[False][2]
    assert np.array_equal(out[True][0], out[False][0])


@pytest.mark.parametrize("name", ["hll_mp5", "hllc_mp7", "hlle_mp7_bz", "hll_mp9_ssp3", "hll_mcl_bz", "lax_bz",
                                  "hllc_mcl_bz", "hlle_mcl_bz", "hllc_mp5_primrec"])
@pytest.mark.parametrize("precision", [cueva.FP64, cueva.MIXED])
def test_multi_step_call_matches_face_planes(name, precision):
    """Within one cueva_step call the recovery pass of step h also forms the
    stage-0 state of step h+1 (and its ghost band by copy); across calls the
    stage kernel does.  Both must give the face-plane path's fields."""
    fs, rec, order, sigma, kw = CASES[name]
    p, prims = cueva.kh_problem(130, 68, sigma=sigma, flux_solver=fs, rec_mp=rec, order=order, rotated=True,
                                precision=precision, crmath=1 - precision, **kw)
    steps = 3 if order == 4 else 6
    say("case", name, "precision", precision, "steps", steps, "in one call")
    out = {}
    for march in (False, True):
        os.environ["CUEVA_MARCH"] = "1" if march else "0"
        try:
            with cueva.Solver(p) as s:
                s.upload(prims)
                s.step(steps)
                c = s.step_counters()
                out[march] = (interior(*s.download(), p), c)
        finally:
            os.environ.pop("CUEVA_MARCH", None)
    nd = ndiff(out[True][0], out[False][0])
    say("  differing points", sum(nd.values()))
    assert out[True][1] == out[False][1]
    assert np.array_equal(out[True][0], out[False][0]), nd


@pytest.mark.parametrize("name", ["hllc_mp5", "hlle_mp7_bz", "hll_mp9_ssp3", "hllc_mcl_bz"])
@pytest.mark.parametrize("precision", [cueva.FP64, cueva.MIXED])
def test_recovery_in_last_marching_launch(name, precision):
    """The Cardano recovery (and the next step's stage-0 state) inside the last
    marching launch (CUEVA_C2P_IN_MARCH=1, measured slower than its own pass
    and so not the default): the same fields and counters as the face-plane
    path with its separate recovery."""
    fs, rec, order, sigma, kw = CASES[name]
    p, prims = cueva.kh_problem(130, 68, sigma=sigma, flux_solver=fs, rec_mp=rec, order=order, rotated=True,
                                precision=precision, crmath=1 - precision, **kw)
    steps = 3 if order == 4 else 6
    say("case", name, "precision", precision, "recovery in the marching launch", steps, "steps")
    out = {}
    for march in (False, True):
        os.environ["CUEVA_MARCH"] = "1" if march else "0"
        os.environ["CUEVA_C2P_IN_MARCH"] = "1"
        try:
            with cueva.Solver(p) as s:
                s.upload(prims)
                s.step(steps - 1)
                s.step(1)  # a call of one step: no stage-0 state formed ahead
                c = s.step_counters()
                out[march] = (interior(*s.download(), p), c)
        finally:
            os.environ.pop("CUEVA_MARCH", None)
            os.environ.pop("CUEVA_C2P_IN_MARCH", None)
    assert out[True][1] == out[False][1]
    assert np.array_equal(out[True][0], out[False][0]), ndiff(out[True][0], out[False][0])
