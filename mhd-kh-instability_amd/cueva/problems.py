"""Named problem setups of the benchmark configurations.

The reference has no Kelvin-Helmholtz setup with periodic x and outflow y
(its BOUND 6 is the transpose, 04_boundary.f95:2867-3033).  ``rotated=True``
gives the same physical problem in the frame x' = y, y' = -x, which runs on
the reference's own BOUND 6 and is the frame of every reference comparison.
"""
from . import abi

# name -> (native imax, native jmax, sigma, flux solver, rec_mp, IMEX order, v0, Bx, Bz)
CONFIGS = {
    "cfg1": (128, 256, 1.0e3, abi.HLL, 0, 2, 0.5, 0.01, 0.0),
    "cfg2": (2048, 4096, 1.0e6, abi.HLLC, 0, 2, 0.5, 0.01, 0.0),
    "cfg3": (2048, 4096, 1.0e6, abi.HLLE, 1, 2, 0.5, 0.01, 0.1),
    "cfg4": (4096, 8192, 1.0e6, abi.HLLC, 0, 2, 0.5, 0.01, 0.0),
    "cfg5": (8192, 16384, 1.0e6, abi.HLLC, 2, 4, 0.9, 0.001, 0.0),
}


def kh_params(imax, jmax, sigma=1.0e3, flux_solver=abi.HLL, rec_mp=0, order=2, rotated=False, **kw):
    """Parameters of the KH problem on a native imax x jmax grid (x periodic, y outflow),
    or of the same problem in the rotated frame (grid jmax x imax, x copy, y periodic)."""
    if rotated:
        return abi.default_params(imax=jmax, jmax=imax, longx=2.0, longy=1.0, bc_x=abi.BC_COPY,
                                  bc_y=abi.BC_PERIODIC, sigma=sigma, flux_solver=flux_solver,
                                  rec_mp=rec_mp, order=order, cfl=0.4, **kw)
    return abi.default_params(imax=imax, jmax=jmax, longx=1.0, longy=2.0, bc_x=abi.BC_PERIODIC,
                              bc_y=abi.BC_COPY, sigma=sigma, flux_solver=flux_solver, rec_mp=rec_mp,
                              order=order, cfl=0.4, **kw)


def kh_problem(imax, jmax, sigma=1.0e3, flux_solver=abi.HLL, rec_mp=0, order=2, rotated=False,
               v0=0.5, bx0=0.01, bz0=0.0, seed=20261202, **kw):
    """(params, initial primitives) for the seeded KH shear layer."""
    p = kh_params(imax, jmax, sigma, flux_solver, rec_mp, order, rotated, **kw)
    prims = abi.kh_init(p.imax, p.jmax, p.longx, p.longy, rotated, v0, bx0, bz0, seed)
    return p, prims


def config(name, rotated=False, **kw):
    ni, nj, sigma, fs, rec, order, v0, bx, bz = CONFIGS[name]
    return kh_problem(ni, nj, sigma, fs, rec, order, rotated, v0=v0, bx0=bx, bz0=bz, **kw)


def cpaw_problem(n, sigma=1.0e6, flux_solver=abi.HLLC, rec_mp=0, order=2, **kw):
    """Circularly polarised Alfven wave on an n x n periodic unit box (rho = 1,
    p = 0.1, B0 = 1, amplitude 0.1): (params, primitives, Alfven speed)."""
    p = abi.default_params(imax=n, jmax=n, longx=1.0, longy=1.0, bc_x=abi.BC_PERIODIC, bc_y=abi.BC_PERIODIC,
                           sigma=sigma, flux_solver=flux_solver, rec_mp=rec_mp, order=order, cfl=0.4, **kw)
    prims, va = abi.cpaw_init(n, n, 1.0, p.gamma)
    return p, prims, va
