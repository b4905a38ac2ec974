"""Kelvin-Helmholtz growth diagnostics on a running solver.

The transverse-velocity amplitude A = (max vy - min vy)/2 over the interior
(native frame: vy is normal to the shear layer) is read with one device
reduction (cueva_plane_range), so a series costs no host copies of the
fields.  The growth rate is the least-squares slope of ln A against time
over a window of samples; forms of the same run are compared on the same
samples.
"""
import numpy as np

from . import abi

VY = (0, abi.PRIM_NAMES.index("Vy"))


def vy_amplitude(solver):
    lo, hi = solver.plane_range(*VY)
    return 0.5 * (hi - lo)


def amplitude_series(solver, nsteps, every, dt):
    """Step `solver` up to `nsteps` steps, sampling A every `every` steps.

    Returns (t, A, stop): stop is None, or (step, CuevaError) for the check
    that failed first, after which sampling ends."""
    ts, amps, stop = [0.0], [vy_amplitude(solver)], None
    for n in range(every, nsteps + 1, every):
        solver.step(every)
        try:
            solver.step_counters()
        except abi.CuevaError as e:
            stop = (n, e)
            break
        ts.append(n * dt)
        amps.append(vy_amplitude(solver))
    return np.asarray(ts), np.asarray(amps), stop


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
