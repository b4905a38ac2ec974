"""Properties of the step at the benchmark configurations' full sizes (2048
x 4096 up to 8192 x 16384), where the oracle would take hours.

* The marching path (reconstruction, Riemann solve and flux difference in
  one launch) and the face-plane path (face fluxes through HBM, the
  per-subroutine entry points' kernels) give the same bits after two steps
  of the KH layer of every configuration (MP5 + HLLC at 2048 x 4096 and
  4096 x 8192, MP7 + HLLE with a guide field, MP9 + HLLC with SSP3(4,3,3)
  and v0 = 0.9 at 8192 x 16384), in FP64 and in the mixed form.
* The same for the other flux solvers at 2048 x 4096 (HLL and HLLE with MP5,
  Lax-Friedrichs and HLL/HLLE/HLLC with the MC limiter), whose marching and
  face-plane kernels are each built for their own waves per SIMD.
* In a doubly periodic box the flux differences telescope: over the cells
  that are not duplicates of the periodic edge, the step's update changes
  the sums of D, tau, S (from the state startconvar forms out of the
  primitives, 06_conserved_var.f95) and of B and q only by rounding
  (17_fluxandsource.f95: none of them has a source in the flux-difference
  form, B only curl and gradient terms).  The primitive recovery between
  steps is not part of this: its pressure loop stops at its own fixed
  point, so the next startconvar need not return the updated D, tau, S.
"""
import os

import numpy as np
import pytest

import cueva
from cueva import problems
from parity import FIELDS, interior

pytestmark = pytest.mark.gpu

NI, NJ = 2048, 4096


def say(*a):
    print(*a, flush=True)


def digests(p, prims, steps, march):
    """128-bit hash (xxh3) of every interior field after `steps` steps: one
    copy of the fields on the host at a time."""
    import xxhash
    old = os.environ.get("CUEVA_MARCH")
    os.environ["CUEVA_MARCH"] = "1" if march else "0"  # read when the context is created
    try:
        with cueva.Solver(p) as s:
            s.upload(prims)
            s.step(steps)
            cnt, _ = s.step_counters()
            g = interior(*s.download(), p)
    finally:
        if old is None:
            del os.environ["CUEVA_MARCH"]
        else:
            os.environ["CUEVA_MARCH"] = old
    assert np.isfinite(g).all()
    out = {f: xxhash.xxh3_128(np.ascontiguousarray(g[k])).hexdigest() for k, f in enumerate(FIELDS)}
    return out, cnt


@pytest.mark.parametrize("cfg,precision", [("cfg2", cueva.FP64), ("cfg3", cueva.FP64), ("cfg4", cueva.FP64),
                                           ("cfg5", cueva.FP64), ("cfg2", cueva.MIXED), ("cfg3", cueva.MIXED)])
def test_march_equals_face_planes_at_full_size(cfg, precision):
    p, prims = problems.config(cfg, precision=precision)
    say("%s (%dx%d), %s, 2 steps: marching path" % (cfg, p.imax, p.jmax, "FP64" if precision == cueva.FP64 else "mixed"))
    a, ca = digests(p, prims, 2, march=True)
    say("%s: face-plane path" % cfg)
    b, cb = digests(p, prims, 2, march=False)
    del prims
    assert ca == cb
    bad = sorted(f for f in FIELDS if a[f] != b[f])
    assert not bad, bad


# the other flux solvers at 2048 x 4096 (each kernel built for its own register bound)
SOLVERS = {"hll_mp5": (cueva.HLL, 0.0), "lax_mc": (cueva.LAX, 0.0), "hll_mc": (cueva.HLL_MCL, 0.0),
           "hlle_mc": (cueva.HLLE_MCL, 0.1), "hllc_mc": (cueva.HLLC_MCL, 0.0), "hlle_mp5": (cueva.HLLE, 0.1)}


@pytest.mark.parametrize("name,precision", [("hll_mp5", cueva.FP64), ("lax_mc", cueva.FP64), ("hlle_mc", cueva.FP64),
                                            ("hllc_mc", cueva.FP64), ("hll_mc", cueva.MIXED), ("lax_mc", cueva.MIXED),
                                            ("hlle_mp5", cueva.MIXED)])
def test_march_equals_face_planes_other_solvers(name, precision):
    fs, bz = SOLVERS[name]
    p, prims = cueva.kh_problem(NI, NJ, sigma=1.0e6, flux_solver=fs, bz0=bz, precision=precision,
                                **({} if precision == cueva.FP64 else {"crmath": 0}))
    say("%s (%dx%d), %s, 2 steps: marching path" % (name, p.imax, p.jmax, "FP64" if precision == cueva.FP64 else "mixed"))
    a, ca = digests(p, prims, 2, march=True)
    say("%s: face-plane path" % name)
    b, cb = digests(p, prims, 2, march=False)
    del prims
    assert ca == cb
    bad = sorted(f for f in FIELDS if a[f] != b[f])
    assert not bad, bad


def periodic_shear(p):
    """Smooth shear flow, periodic in x and y, with the duplicated edges
    (i = imax, j = jmax) copied from i = 0, j = 0 bit for bit."""
    ni, nj = p.imax, p.jmax
    x = -0.5 + np.arange(ni + 1) * (p.longx / ni)
    y = -1.0 + np.arange(nj + 1) * (p.longy / nj)
    X, Y = np.meshgrid(x, y)  # (nj+1, ni+1), x fastest
    vx = 0.3 * np.sin(np.pi * Y)
    vy = 0.01 * np.sin(2 * np.pi * X) * np.cos(np.pi * Y)
    vz = np.zeros_like(X)
    bx, by, bz = np.full_like(X, 0.01), np.zeros_like(X), np.full_like(X, 0.05)
    ex, ey, ez = -(vy * bz - vz * by), -(vz * bx - vx * bz), -(vx * by - vy * bx)  # E = -v x B
    zero, one = np.zeros_like(X), np.ones_like(X)
    planes = [zero, zero, bx, by, bz, vx, vy, vz, ex, ey, ez, zero, one, one]  # restart order
    prims = np.zeros((14, nj + 13, ni + 13))
    for k, a in enumerate(planes):
        a = a.copy()
        a[:, ni] = a[:, 0]
        a[nj, :] = a[0, :]
        prims[k, 6:7 + nj, 6:7 + ni] = a
    return prims


def test_periodic_sums_are_conserved_at_full_size():
    p = cueva.default_params(imax=NI, jmax=NJ, longx=1.0, longy=2.0, bc_x=cueva.BC_PERIODIC,
                             bc_y=cueva.BC_PERIODIC, sigma=1.0e6, flux_solver=cueva.HLLC, cfl=0.4)
    prims = periodic_shear(p)
    names = ("D", "tau", "Sx", "Sy", "Sz", "Bx", "By", "Bz", "q")
    idx = [FIELDS.index(n) for n in names]

    def sums(s):
        # without the periodic duplicates; summed in extended precision, so
        # that the check's own rounding stays far below the update's
        g = interior(*s.download(), p)[:, :-1, :-1]
        return ([np.sum(g[k], dtype=np.longdouble) for k in idx], [float(np.abs(g[k]).sum()) for k in idx])

    with cueva.Solver(p) as s:
        s.upload(prims)
        for step in range(2):
            cueva.abi.check(s.lib.cueva_startconvar(s.ctx), "startconvar")  # D, tau, S of this step's start
            before, _ = sums(s)
            s.step(1)
            cnt, _ = s.step_counters()
            assert cnt == [0, 0, 0]
            after, scale = sums(s)
            drift = {n: float(abs(a - b)) / max(c, 1e-300) for n, a, b, c in zip(names, after, before, scale)}
            say("step", step + 1, "relative change of the sums:", {n: "%.1e" % v for n, v in drift.items()})
            assert all(v <= 1e-15 for v in drift.values()), drift
