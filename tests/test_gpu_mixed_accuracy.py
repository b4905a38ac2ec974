"""Accuracy of the mixed form (FP32 reconstruction, Riemann solve and flux
difference; FP64 recovery, E solve, sums and reductions) against the FP64
oracle on Kelvin-Helmholtz data, per field.

Two schemes on a 128 x 256 grid (rotated frame), sigma = 1e6, 10 steps from
the seeded data: MP7 + HLLE with a guide field Bz = 0.1 (the scheme of the
mixed-precision configuration) and MP5 + HLLC (the headline scheme).  (MP7 +
HLLE stops in step 20 at this size, det < 0 at (h, l, i, j) = (20, 2, 126,
1), in the FP64 and the mixed form alike -- the scheme's stop, as at 2048 x
4096.)  Each field's bound is its relative L1 distance as measured on MI355X
for this build, times 4, rounded up: room for the order of operations a
compiler keeps in the FP32 unit, and nothing more.  The divergence-cleaning
potential phi is a round-off residue in both forms (|phi| ~ 1e-10): it is
held to an absolute bound on its mean difference instead.  Fields that stay
exactly zero in both forms (HLLC: psi, Bz, vz, Ex, Ey, q, Sz) must stay so.
"""
import numpy as np
import pytest

import cueva
from oracle.binding import Oracle
from parity import FIELDS, interior

pytestmark = pytest.mark.gpu

# per scheme: relative L1 bound per field; absolute (mean |x - x_ref|) bound
# for the residue fields
SCHEMES = {
    # measured at step 10: psi 6.7e-3, Bx 2.4e-3, By 3.2e-7, Bz 1.2e-6, vx 3.8e-4, vy 1.2e-6,
    # vz 1.3e-2, Ex 4.1e-5, Ey 1.7e-2, Ez 5.8e-4, q 1.3e-2, rho 7.1e-7, p 8.2e-7, D 8.7e-7,
    # tau 1.3e-6, Sx 3.8e-4, Sy 2.2e-6, Sz 7.1e-6; mean |phi diff| 2.5e-10
    "hlle_mp7_bz": dict(fs=cueva.HLLE, rec=1, bz0=0.1,
                        rel=dict(psi=3e-2, Bx=1e-2, By=1.3e-6, Bz=5e-6, Vx=1.6e-3, Vy=5e-6, Vz=6e-2, Ex=1.7e-4,
                                 Ey=7e-2, Ez=2.4e-3, q=6e-2, rho=3e-6, p=3.3e-6, D=3.5e-6, tau=5.2e-6, Sx=1.6e-3,
                                 Sy=9e-6, Sz=3e-5),
                        abs=dict(phi=1e-9)),
    # measured at step 10: Bx 2.8e-3, By 4.1e-7, vx 4.6e-4, vy 8.2e-7, Ez 7.4e-4, rho 1.0e-6,
    # p 1.2e-6, D 9.7e-7, tau 1.0e-6, Sx 4.7e-4, Sy 1.4e-6; psi Bz vz Ex Ey q Sz exactly 0;
    # mean |phi diff| 1.9e-10 (the bounds below are those of an earlier FP32 build, 4x its
    # measurements: Bx 2.8e-3 ... rho 1.3e-6, Sy 1.5e-6)
    "hllc_mp5": dict(fs=cueva.HLLC, rec=0, bz0=0.0,
                     rel=dict(psi=0.0, Bx=1.2e-2, By=1.6e-6, Bz=0.0, Vx=1.9e-3, Vy=3.7e-6, Vz=0.0, Ex=0.0, Ey=0.0,
                              Ez=3e-3, q=0.0, rho=5.2e-6, p=4.8e-6, D=5.2e-6, tau=4.4e-6, Sx=2e-3, Sy=6e-6, Sz=0.0),
                     abs=dict(phi=1e-9)),
}
RESIDUE = ("phi",)


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", sorted(SCHEMES))
def test_mixed_form_against_fp64_oracle(name):
    sc = SCHEMES[name]
    p, prims = cueva.kh_problem(128, 256, sigma=1.0e6, flux_solver=sc["fs"], rec_mp=sc["rec"], rotated=True,
                                bz0=sc["bz0"], precision=cueva.MIXED)
    steps = 10
    say("mixed vs FP64 oracle", name, "128x256,", steps, "steps")
    with cueva.Solver(p) as s:
        s.upload(prims)
        s.step(steps)
        cnt, _ = s.step_counters()
        g = interior(*s.download(), p)
    o = Oracle(p)  # FP64 throughout, correctly rounded Cardano
    o.upload(prims)
    o.step(steps)
    r = interior(*o.download(), p)
    assert cnt == o.counters()
    rel, ab = {}, {}
    for k, f in enumerate(FIELDS):
        d = np.abs(g[k] - r[k])
        den = np.abs(r[k]).sum()
        rel[f] = float(d.sum() / den) if den > 0 else (0.0 if d.sum() == 0 else np.inf)
        ab[f] = float(d.mean())
    say(name, "relative L1:", {f: "%.2e" % v for f, v in rel.items() if f not in RESIDUE})
    say(name, "mean |diff| of the residue fields:", {f: "%.2e" % ab[f] for f in RESIDUE})
    assert np.isfinite(g).all()
    bad = {f: (rel[f], sc["rel"][f]) for f in FIELDS if f not in RESIDUE and rel[f] > sc["rel"][f]}
    bad.update({f: (ab[f], sc["abs"][f]) for f in RESIDUE if ab[f] > sc["abs"][f]})
    assert not bad, bad
