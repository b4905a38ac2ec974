"""Crafted initial data that drive the step through its failure branches.

The seeded Kelvin-Helmholtz data never reach the fix-up and fallback
branches of the reference (every KH fixture records counters 0, 0, 0).
These modifications of the KH primitives (in the rotated frame the
fixtures use, 64 x 32 cells) do; tests/golden/make_golden.py runs the
reference on them and the tests run the oracle and the GPU on the same data.

* ``lowp``: pressure 1e-6 in a 7 x 7 patch: the recovery's pressure loop
  ends below zero there every step (the p fix-up, 12_varprimitive.f95:592-607).
* ``negrho``: one cell with density -0.2: HLLC face densities below zero (the
  face density fix, 16_hllc_flow_mpx.f95:2663-2693).
* ``spin``: rows 10..22 move at |v| = 0.97 in a direction that turns by 1.3
  rad per cell along x; the MP5 face states of such data exceed the speed
  of light (the superluminal face fallback, 16_hll_flow_mpx.f95:1392-1461),
  and the reference then stops within the first step (Cardano det < 0 after
  HLL faces, get_hllc_vel non-convergence in HLLC).
"""
import numpy as np


def _grid(prims):
    ny, nx = prims.shape[1:]
    j, i = np.meshgrid(np.arange(ny) - 6, np.arange(nx) - 6, indexing="ij")
    return i, j


def _e_ideal(prims):
    prims[8:11] = -np.cross(prims[5:8], prims[2:5], axis=0)


def lowp(prims):
    i, j = _grid(prims)
    m = (abs(i - 32) < 4) & (abs(j - 16) < 4)
    prims[13][m] = 1.0e-6


def negrho(prims):
    prims[12][6 + 10, 6 + 20] = -0.2


def spin(prims):
    i, j = _grid(prims)
    m = (j >= 10) & (j <= 22)
    th = 1.3 * i
    prims[5][m] = (0.97 * np.cos(th))[m]
    prims[6][m] = (0.97 * np.sin(th))[m]
    prims[7][m] = 0.0
    _e_ideal(prims)


MODS = {"lowp": lowp, "negrho": negrho, "spin": spin}


def problem(meta):
    """(params, primitives) of a fixture's initial data: the seeded KH
    problem of its grid and scheme, then its crafted modification, if any."""
    import cueva
    init = dict(meta.get("init", {}))
    mod = init.pop("crafted", None)
    ni, nj = meta["native_grid"]
    p, prims = cueva.kh_problem(ni, nj, sigma=meta["sigma"], flux_solver=meta["flux_solver"],
                                rec_mp=meta["rec_mp"], order=meta["order"], rotated=True, seed=meta["seed"], **init)
    if mod:
        MODS[mod](prims)
    return p, prims
