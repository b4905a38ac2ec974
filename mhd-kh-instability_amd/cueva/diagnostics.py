"""Kelvin-Helmholtz growth diagnostics on a running solver.

Two measures of the transverse velocity (native frame: vy is normal to the
shear layer), both read on the device, so a series costs no host copies of
the fields:
* the amplitude A = (max vy - min vy)/2 over the interior, one min/max
  reduction (cueva_plane_range);
* the reference's own measure, I = sum of vy^2 dx dy over the interior
  (22_writer.f95:893, written to data_global/2d_data_vy_integral_vol.dat),
  from the global diagnostics (cueva_reduce_globals).
The growth rate is the least-squares slope of ln A against time over a
window of samples, the fit the reference's scripts make
(fit_scripts/fit_growth.gnu:28-53); I grows as A^2, so its rate is half the
slope of ln I.  Forms of the same run are compared on the same samples.
"""
import numpy as np

from . import abi

VY = (0, abi.PRIM_NAMES.index("Vy"))


def vy_amplitude(solver):
    lo, hi = solver.plane_range(*VY)
    return 0.5 * (hi - lo)


def vy_integral(solver):
    return solver.globals()["sum_vy"]


def kh_series(solver, nsteps, every, dt, integral=False):
    """Step `solver` up to `nsteps` steps, sampling every `every` steps.

    Returns (t, A, I, stop): I is None unless `integral`; stop is None, or
    (step, CuevaError) for the check that failed first, after which
    sampling ends."""
    ts, amps = [0.0], [vy_amplitude(solver)]
    ints = [vy_integral(solver)] if integral else None
    stop = None
    for n in range(every, nsteps + 1, every):
        solver.step(every)
        try:
            solver.step_counters()
        except abi.CuevaError as e:
            stop = (n, e)
            break
        ts.append(n * dt)
        amps.append(vy_amplitude(solver))
        if integral:
            ints.append(vy_integral(solver))
    return np.asarray(ts), np.asarray(amps), None if ints is None else np.asarray(ints), stop


def amplitude_series(solver, nsteps, every, dt):
    """kh_series without the integral: (t, A, stop)."""
    t, a, _, stop = kh_series(solver, nsteps, every, dt)
    return t, a, stop


def growth_window(amps, factor=1.5):
    """Sample indices from the first A >= factor * A[0] to the last sample."""
    above = np.nonzero(amps >= factor * amps[0])[0]
    if above.size < 3:
        return None
    return int(above[0]), int(len(amps))


def growth_rate(t, amps, window):
    """Slope of ln A against t over samples [window[0], window[1])."""
    a, b = window
    return float(np.polyfit(t[a:b], np.log(amps[a:b]), 1)[0])


def growth_rate_of_integral(t, ints, window):
    """Amplitude growth rate from a quadratic integral such as I: half the
    slope of ln I (fit_growth.gnu:53 halves its fit of a squared field the
    same way)."""
    return 0.5 * growth_rate(t, ints, window)
