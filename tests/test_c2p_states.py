"""Primitive recovery on crafted states: the fix-up, drift and stop branches.

The KH fixtures never reach the recovery's failure branches (every one
records fix-up counters 0, 0, 0).  tests/golden/c2p_states.* come from the
reference's varprimitivecardano run alone (the -c2p mode of
oracle/refbuild/ref_harness.f95) on a 64 x 32 grid of states chosen to pass
through every branch of 12_varprimitive.f95:

* rho < 0 after the quartic -> max(0.1 rho^n, rho), counter r (:536-548);
* p < 0 after the pressure loop -> max(1e-4 p^n, p), counter p (:592-607);
* |v| >= 1 after the recovery with B^2/rho >= 10 -> the drift branch,
  counter v (:627-666), whose W from the superluminal v makes the result NaN;
* a superluminal previous velocity, clamped with the module's eps1, which
  is 0 when no hll_flow_mpx ran in the process (:236-241);
* and, one grid each, two cells that stop the reference with different
  messages, of which it reports the first in its loop order (j outer, i
  inner): no admissible resolvent root (:419-422), det < 0 (:374-415),
  both quartic roots <= 1 (:482-486), superluminal with B^2/rho < 10
  (:684-693).

The fix-up records are the reference's own unit 301-303 files, parsed.
This file pins the oracle (CPU); tests/test_gpu_c2p_states.py holds the GPU
to the same fixture.
"""
import json
import os

import numpy as np
import pytest

from oracle.binding import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
KIND = ("rho", "p", "v")


def load():
    meta = json.load(open(os.path.join(GOLD, "c2p_states.json")))
    arrays = dict(np.load(os.path.join(GOLD, "c2p_states.npz")))
    return meta, arrays


class _Grid:
    """Parameters of the grid the states sit on (only sizes and gamma enter)."""

    def __init__(self, imax, jmax):
        import cueva
        self.p = cueva.default_params(imax=imax, jmax=jmax, flux_solver=cueva.HLLC, bc_x=cueva.BC_COPY,
                                      bc_y=cueva.BC_PERIODIC, sigma=1.0e6)


def set_cell(prims, cons, i, j, st):
    jj, ii = j + 6, i + 6
    cons[:, jj, ii] = st[0:5]
    prims[8:11, jj, ii] = st[5:8]
    prims[2:5, jj, ii] = st[8:11]
    prims[5:8, jj, ii] = st[11:14]
    prims[12, jj, ii] = st[14]
    prims[13, jj, ii] = st[15]


def planes(meta, arrays):
    """Host planes (14 prims, 5 cons) of the fixture's state grid."""
    imax, jmax = meta["grid"]
    prims = np.zeros((14, jmax + 13, imax + 13))
    cons = np.zeros((5, jmax + 13, imax + 13))
    prims[12] = 1.0
    prims[13] = 1.0
    for n, st in enumerate(arrays["states"]):
        j, i = divmod(n, imax + 1)
        set_cell(prims, cons, i, j, st)
    return prims, cons


def result_of(prims, meta):
    imax, jmax = meta["grid"]
    return prims[[5, 6, 7, 12, 13], 6:6 + jmax + 1, 6:6 + imax + 1]


def same_bits(a, b):
    return np.array_equal(np.asarray(a, dtype="<f8").view("<u8"), np.asarray(b, dtype="<f8").view("<u8"))


def records_match(ij, vals, arrays, name):
    """Records equal in cell and order; values equal as the reference's
    shortest round-trip text reads back (NaN as NaN)."""
    rij, rval = arrays["fixup_%s_ij" % name], arrays["fixup_%s_val" % name]
    if ij.shape != rij.shape or not np.array_equal(ij, rij):
        return False
    both_nan = np.isnan(vals) & np.isnan(rval)
    return bool(np.all(both_nan | (vals == rval)))


def test_fixture_covers_every_branch():
    meta, arrays = load()
    assert all(c > 100 for c in meta["counters"]), meta["counters"]
    assert meta["superluminal_old_v_messages"] == 4
    assert sorted(s["code"] for s in meta["stops"]) == [21, 22, 23, 25]
    # the drift branch leaves NaN velocities (W of the superluminal v)
    res = arrays["result"]
    assert np.isnan(res[0]).sum() >= meta["counters"][2] // 2


def test_oracle_recovery_matches_reference_bitwise():
    meta, arrays = load()
    imax, jmax = meta["grid"]
    prims, cons = planes(meta, arrays)
    o = Oracle(_Grid(imax, jmax).p, libm="c")
    o.upload(prims)
    o.upload_cons(cons)
    assert o.recover() == 0
    got, _ = o.download()
    res = result_of(got, meta)
    assert same_bits(res, arrays["result"]), int((res.view("u8") != arrays["result"].view("u8")).sum())
    assert o.counters() == meta["counters"]
    for kind, name in enumerate(KIND):
        ij, vals = o.fixups(kind)
        assert records_match(ij, vals, arrays, name), name


@pytest.mark.parametrize("k", range(4))
def test_oracle_stops_where_the_reference_stops(k):
    meta, arrays = load()
    imax, jmax = meta["grid"]
    stop = meta["stops"][k]
    prims, cons = planes(meta, arrays)
    for c in stop["cells"]:
        set_cell(prims, cons, c["i"], c["j"], np.array(c["state"]))
    o = Oracle(_Grid(imax, jmax).p, libm="c")
    o.upload(prims)
    o.upload_cons(cons)
    assert o.recover() == stop["code"], stop["message"]
    first = stop["cells"][0]
    assert o.last_stop()[2:] == (first["i"], first["j"])


def test_planted_mistake_fails():
    """A wrong fix-up floor in the oracle's inputs (p^n scaled) must show."""
    meta, arrays = load()
    imax, jmax = meta["grid"]
    prims, cons = planes(meta, arrays)
    prims[13] *= 2.0  # previous pressure: enters only the p fix-up floor
    o = Oracle(_Grid(imax, jmax).p, libm="c")
    o.upload(prims)
    o.upload_cons(cons)
    o.recover()
    got, _ = o.download()
    assert not same_bits(result_of(got, meta), arrays["result"])


def test_oracle_clamp_after_an_hll_face_solve():
    """After hll_flow_mpx has run, eps1 = 1e-6 clamps a superluminal previous
    velocity to v^2 = 1 - 1e-6 instead of 1 (the reference run with one
    hll_flow_mpx call before the recovery)."""
    import ctypes
    from oracle.binding import _dp
    from oracle.binding import load as load_oracle
    lib = load_oracle()
    meta = json.load(open(os.path.join(GOLD, "c2p_states.json")))
    arrays = dict(np.load(os.path.join(GOLD, "c2p_states.npz")))
    imax, jmax = meta["grid"]
    res = np.zeros_like(arrays["result_eps1"])
    for n, st in enumerate(arrays["states"]):
        j, i = divmod(n, imax + 1)
        out = np.array([0.0, 0.0, 0.0, st[14], st[15]])
        fl = (ctypes.c_int * 3)()
        rec = np.zeros(6)
        assert lib.oracle_c2p_point(meta["gamma"], 1.0e-6, _dp(st[:14].copy()), _dp(out), fl, _dp(rec)) == 0
        res[:, j, i] = out
    assert meta["cells_changed_by_eps1"] >= 1
    assert same_bits(res, arrays["result_eps1"])
