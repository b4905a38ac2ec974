"""KH transverse-velocity growth in the three forms.

A 512x1024 KH layer (v0 = 0.5, sigma = 1e6, MP5 + HLLC, SSP2(2,2,2)) in the
native frame, stepped by plain FP64, the mixed form and the mixed form with
the learned recovery start.  The amplitude A = (max vy - min vy)/2 grows
exponentially until the reference's own HLLC contact solve stops the run
(get_hllc_vel: no convergence, 16_hllc_flow_mpx.f95 / funciones.f95:238-244)
near step 300 in every form; the growth rate of ln A over the samples where
A has passed 1.5 A(0) must agree with plain FP64 within 1 %, and so must
the rate the reference's own measure gives on the same samples (half the
log slope of the vy^2 integral, 22_writer.f95:893, fit_growth.gnu:32,53).
"""
import sys

import numpy as np
import pytest

import cueva
from cueva import diagnostics, learned_c2p as lc

pytestmark = pytest.mark.gpu
FORMS = {
    "plain_fp64": dict(precision=cueva.FP64, crmath=1, c2p_mode=cueva.C2P_CARDANO),
    "mixed": dict(precision=cueva.MIXED, crmath=0, c2p_mode=cueva.C2P_CARDANO),
    "mixed_learned": dict(precision=cueva.MIXED, crmath=0, c2p_mode=cueva.C2P_LEARNED),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def series(form, nsteps=290, every=10):
    p, prims = cueva.kh_problem(512, 1024, sigma=1.0e6, flux_solver=cueva.HLLC, **FORMS[form])
    dt = cueva.derived(p)["dt"]
    with cueva.Solver(p) as s:
        if form == "mixed_learned":
            s.set_c2p_params(lc.load_params())
        s.upload(prims)
        return diagnostics.kh_series(s, nsteps, every, dt, integral=True)


def test_growth_rate_of_every_form_matches_fp64():
    t, a, vi, stop = series("plain_fp64")
    say("plain: %d samples, stop %s, A %.3e -> %.3e" % (len(a), stop, a[0], a[-1]))
    win = diagnostics.growth_window(a)
    assert win is not None, a
    g0 = diagnostics.growth_rate(t, a, win)
    h0 = diagnostics.growth_rate_of_integral(t, vi, win)
    say("plain growth rate %.4f (integral %.4f) over t in [%.3f, %.3f]" % (g0, h0, t[win[0]], t[win[1] - 1]))
    assert g0 > 0 and h0 > 0
    for form in ("mixed", "mixed_learned"):
        tf, af, vf, stopf = series(form)
        assert len(af) >= win[1], (form, stopf)
        g = diagnostics.growth_rate(tf, af, win)
        h = diagnostics.growth_rate_of_integral(tf, vf, win)
        say("%s growth rate %.4f (%.3f %% from plain), integral %.4f (%.3f %%)"
            % (form, g, 100 * (g / g0 - 1), h, 100 * (h / h0 - 1)))
        assert abs(g / g0 - 1) <= 0.01
        assert abs(h / h0 - 1) <= 0.01
